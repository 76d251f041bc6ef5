"""Build libstaccato_sweep.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m staccato_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "stc_api.hip")
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libstaccato_sweep.so")
HOST_SRC = [os.path.join(HERE, "csrc", "host", f) for f in ("KrylovROMSweep.cpp", "IJCSACoupling.cpp")]
HOST_OUT = os.path.join(OUT_DIR, "libstaccato_host.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off: the compiler would otherwise pair
    # float arithmetic into v_pk_* forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    # C++ mirror of the reference's sweep entry, over the C ABI
    hcmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", HOST_OUT + ".tmp", *HOST_SRC,
            "-L" + OUT_DIR, "-lstaccato_sweep", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(hcmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    os.replace(HOST_OUT + ".tmp", HOST_OUT)
    if verbose:
        print(" ".join(cmd))
        print(" ".join(hcmd))
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
    sys.exit(0)
