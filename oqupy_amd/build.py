"""Build the MI355X engine library in-tree: ``python -m oqupy_amd.build``.

Compiles ``oqupy_amd/csrc/*.hip`` with hipcc for gfx950 into
``oqupy_amd/libptengine.so`` (no JIT cache: the library travels with the
source tree).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["pte_kernels.hip", "pte_rank.hip", "pte_eta.hip", "pte_engine.hip"]
TARGET = os.path.join(HERE, "libptengine.so")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile the engine; returns the library path."""
    cmd = [hipcc(), "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
           "-shared", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include")]
    cmd += [os.path.join(HERE, "csrc", s) for s in SOURCES]
    tmp = TARGET + ".tmp"
    cmd += ["-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libptengine.so")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
