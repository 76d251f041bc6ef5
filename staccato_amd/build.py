"""Build libstaccato_sweep.so for gfx950 with hipcc (in-tree, no JIT cache).

Every sweep-kernel instantiation (factor type x rows per lane) is compiled in
its own hipcc process, in parallel, then linked with the host API.

    python -m staccato_amd.build
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
OUT = os.path.join(OUT_DIR, "libstaccato_sweep.so")
HOST_SRC = [os.path.join(CSRC, "host", f) for f in ("KrylovROMSweep.cpp", "IJCSACoupling.cpp")]
HOST_OUT = os.path.join(OUT_DIR, "libstaccato_host.so")

# (refine, rows per lane, panel kind) triples of sweep_kernel, one object each
# (panel kind 0: register panel, 1: wide LDS panel, 2: wide panel with Gauss 3M
# products, complex64 only, 3: wide panel with the two-level update); the list matches STC_INSTANCES in csrc/stc_dispatch.hip
INSTANCES = ([(r, m, 0) for r in (0, 1) for m in (1, 2, 4, 8, 16)] +
             [(0, m, w) for w in (1, 3) for m in (2, 4, 8)] +
             [(1, m, w) for w in (1, 2, 3) for m in (2, 4, 8, 16)])

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: the compiler would otherwise pair
    # float arithmetic into v_pk_* forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return cmd


def build_library(verbose=False, jobs=None):
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    jobs = jobs or min(16, max(1, os.cpu_count() or 1))
    cmds, objs = [], []
    for r, m, wd in INSTANCES:
        o = os.path.join(OBJ_DIR, f"sweep_kernel_{r}_{m}_{wd}.o")
        cmds.append([cc, *HIPCC_FLAGS, "-c", f"-DSTC_INST_REFINE={r}", f"-DSTC_INST_MAXR={m}",
                     f"-DSTC_INST_WIDE={wd}", "-o", o, os.path.join(CSRC, "stc_kern_inst.hip")])
        objs.append(o)
    for name in ("stc_api", "stc_dispatch"):
        o = os.path.join(OBJ_DIR, name + ".o")
        cmds.append([cc, *HIPCC_FLAGS, "-c", "-o", o, os.path.join(CSRC, name + ".hip")])
        objs.append(o)
    # the longest compilations first
    def cost(cmd):
        m = [a for a in cmd if a.startswith("-DSTC_INST_MAXR=")]
        return int(m[0].split("=")[1]) if m else 0
    order = sorted(range(len(cmds)), key=lambda i: -cost(cmds[i]))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, [cmds[i] for i in order]))
    tmp = OUT + ".tmp"
    link = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    _run(link)
    os.replace(tmp, OUT)
    # C++ mirror of the reference's sweep entry, over the C ABI
    hcmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", HOST_OUT + ".tmp", *HOST_SRC,
            "-L" + OUT_DIR, "-lstaccato_sweep", "-Wl,-rpath,$ORIGIN"]
    _run(hcmd)
    os.replace(HOST_OUT + ".tmp", HOST_OUT)
    if verbose:
        for c in cmds + [link, hcmd]:
            print(" ".join(c))
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
    sys.exit(0)
