"""Build the HIP library (gfx950) in-tree: multigrid_amd/libpmg_mi355x.so.

The library is compiled with hipcc directly (no JIT cache): the built .so
sits beside this file so that it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpmg_mi355x.so")
SOURCES = ["pmg_kernels.hip", "pmg_pbe.hip", "pmg_band.hip", "pmg_solver.hip", "pmg_serve.hip", "pbinput.cpp",
           "dxio.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-pthread",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out)
    if verbose and out.strip():
        print(out)


def build(verbose=False):
    """Compile every source to an object (in parallel), then link."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    hipcc = hipcc_path()
    with ThreadPoolExecutor(max_workers=min(len(srcs), 4)) as ex:
        list(ex.map(lambda so: _run([hipcc] + HIPCC_FLAGS + ["-c", "-o", so[1], so[0]], verbose),
                    zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-pthread", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
