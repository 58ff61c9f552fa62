"""Package-wide constants and the backend selection surface.

Mirrors the constants of the reference's ``oqupy/config.py:16-75`` that the
process-tensor path reads. The backend dictionaries keep their names and their
role (``Tempo``/``PtTempo`` fall back to them when ``backend_config`` is None),
but here every configuration selects the MI355X engine: the keys understood are

* ``"backend"``: only ``"mi355x"`` (the default) is accepted;
* ``"precision"``: ``"c128"`` (complex128 everywhere, the default) or ``"c64"``
  (complex64 contractions with FP32 accumulation; FP64 for the
  orthogonalisation, singular values and the truncation decision);
* ``"chi_cap"``: an optional limit on every bond of the device-resident
  chain (default: none, as in the reference, which never caps a bond); when
  given, a truncation that would keep more singular values raises an error.
"""

from numpy import complex128, float64

# numpy datatypes of every host-side tensor (oqupy/config.py:19-20)
NpDtype = complex128
NpDtypeReal = float64

# separator for the __str__ of API objects (oqupy/config.py:23)
SEPERATOR = "----------------------------------------------\n"

# relative precision of scipy.integrate.quad (oqupy/config.py:26)
INTEGRATE_EPSREL = 2**(-26)

# maximal number of subdivisions of the adaptive quadrature (oqupy/config.py:29)
SUBDIV_LIMIT = 256

# default progress report: 'silent', 'simple' or 'bar' (oqupy/config.py:32)
PROGRESS_TYPE = 'silent'

# default TEMPO backend configuration (oqupy/config.py:38)
TEMPO_BACKEND_CONFIG = {"backend": "mi355x", "precision": "c128"}

# default PT-TEMPO backend configuration (oqupy/config.py:53)
PT_TEMPO_BACKEND_CONFIG = {"backend": "mi355x", "precision": "c128"}

# rounding (in decimals) used to detect degenerate coupling eigenvalue sums and
# differences (oqupy/config.py:75)
DEFAULT_TOLERANCE_DEGENERACY = 12

# bond limit of the device-resident tensor chain when ``chi_cap`` is not given:
# none (the largest C int; device buffers grow with the bonds on demand)
DEFAULT_CHI_CAP = 2**31 - 1
