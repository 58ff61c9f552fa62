"""MI355X-native process-tensor engine behind OQuPy's public API.

The names below keep the reference's API (``oqupy/__init__.py:53-151``) for the
process-tensor path: TEMPO and PT-TEMPO computations and ``compute_dynamics``
run on the hand-written HIP engine ``libptengine.so`` (gfx950); baths,
correlations, systems and influence tensors are host-side FP64 code.
"""

from oqupy_amd import operators
from oqupy_amd.bath import Bath
from oqupy_amd.control import Control
from oqupy_amd.correlations import (compute_correlations,
                                    compute_correlations_nt)
from oqupy_amd.bath_correlations import (CustomCorrelations, CustomSD,
                                         PowerLawSD)
from oqupy_amd.dynamics import Dynamics, MeanFieldDynamics
from oqupy_amd.process_tensor import (SimpleProcessTensor,
                                      TrivialProcessTensor)
from oqupy_amd.pt_tempo import PtTempo, pt_tempo_compute, pt_tempo_compute_batch
from oqupy_amd.system import (MeanFieldSystem, System, TimeDependentSystem,
                              TimeDependentSystemWithField)
from oqupy_amd.system_dynamics import compute_dynamics, compute_dynamics_batch
from oqupy_amd.tempo import (MeanFieldTempo, Tempo, TempoParameters,
                             influence_matrix, tempo_compute)
from oqupy_amd.version import __version__

__all__ = [
    'Bath', 'Control', 'CustomCorrelations', 'CustomSD', 'Dynamics',
    'MeanFieldDynamics', 'MeanFieldSystem', 'MeanFieldTempo', 'PowerLawSD',
    'PtTempo', 'SimpleProcessTensor', 'System', 'Tempo', 'TempoParameters',
    'TimeDependentSystem', 'TimeDependentSystemWithField',
    'TrivialProcessTensor', 'compute_correlations',
    'compute_correlations_nt', 'compute_dynamics',
    'compute_dynamics_batch',
    'influence_matrix', 'operators', 'pt_tempo_compute',
    'pt_tempo_compute_batch', 'tempo_compute',
    '__version__',
]
