"""Build libsymcode_hip.so for gfx950 with hipcc (one object per source,
compiled in parallel, then linked).  Usage: python build.py [-j N]"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "symcode", "libsymcode_hip.so")
SOURCES = ["api.hip", "lib_f64.hip", "lib_f32.hip", "cnn.hip", "pmlp.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(jobs=8, verbose=False, defines=(), out=None, bdir=None):
    """Compile and link.  ``defines`` are extra -D tuning knobs (tile sizes,
    waves-per-SIMD hints); ``out`` / ``bdir`` redirect the library and objects."""
    bdir = bdir or os.path.join(HERE, "_build")
    out = out or OUT
    os.makedirs(bdir, exist_ok=True)
    cc = hipcc()

    def one(src):
        obj = os.path.join(bdir, src + ".o")
        cmd = [cc, *FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-8000:]}")
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = out + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("-D", action="append", default=[], help="extra define, e.g. SC_WPE_X=3")
    ap.add_argument("-o", default=None, help="output library path")
    ap.add_argument("--bdir", default=None)
    a = ap.parse_args()
    print(build(a.j, a.v, a.D, a.o, a.bdir))
