"""ctypes binding of libstaccato_sweep.so (C ABI: include/staccato_sweep.h).

The library is built in-tree by ``staccato_amd.build.build_library()`` into
``staccato_amd/lib/libstaccato_sweep.so``.  Loading fails loudly when it is
missing: there is no CPU fallback for the sweep.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libstaccato_sweep.so")
# the C++ mirror of the reference's sweep entry (KrylovROMSweep, IJCSACoupling);
# STC_HOST_LIB selects another build of it (the sanitizer build of the CPU tests)
HOST_LIB_PATH = os.environ.get("STC_HOST_LIB") or os.path.join(HERE, "lib", "libstaccato_host.so")

STC_OK = 0
STC_EINVAL = -1
STC_ENOMEM = -2
STC_EDEVICE = -3
STC_ENOROM = -4
STC_EBUSY = -5
STC_ETOOBIG = -6
STC_ESINGULAR = 1
STC_ENOCONV = 2

STC_PLAIN_FP64 = 0
STC_REFINED_FP32 = 1

# every symbol include/staccato_sweep.h declares
EXPORTED = (
    "stc_create", "stc_destroy", "stc_strerror", "stc_device_ok", "stc_set_policy", "stc_clear_policies",
    "stc_rom_set", "stc_rom_set_io", "stc_rom_clear",
    "stc_sweep", "stc_sweep_dev", "stc_sweep_dev_noalloc", "stc_getrf",
    "stc_sweep_batch", "stc_plan_create", "stc_plan_execute", "stc_plan_destroy",
    "stc_adjoint", "stc_back_transform", "stc_gesv_batched", "stc_gesv_batched_dev",
)


class Stats(ctypes.Structure):
    _fields_ = [
        ("seconds", ctypes.c_double),
        ("points", ctypes.c_longlong),
        ("max_refine_steps", ctypes.c_int),
        ("n_singular", ctypes.c_int),
        ("n_not_converged", ctypes.c_int),
        ("max_backward_error", ctypes.c_double),
    ]


class Policy(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int),
        ("lds_budget", ctypes.c_int),
        ("jb_max", ctypes.c_int),
    ]


class SysDesc(ctypes.Structure):
    _fields_ = [
        ("rom_id", ctypes.c_int),
        ("nfreq", ctypes.c_int),
        ("freq_hz", ctypes.c_void_p),
        ("nrhs", ctypes.c_int),
        ("F", ctypes.c_void_p),
        ("ldf", ctypes.c_int),
        ("X", ctypes.c_void_p),
        ("ldx", ctypes.c_int),
        ("info", ctypes.c_void_p),
        ("berr", ctypes.c_void_p),
    ]


_lib = None


def load(path=None):
    """Load the library once; raises OSError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    # torch ships its own libamdhip64.so.7 (same SONAME as /opt/rocm's).  The
    # first copy loaded serves the whole process, and torch does not start on
    # a newer runtime than the one it was built with, while this library runs
    # on either: let torch's copy load first when torch is installed.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(p):
        raise OSError(
            f"{p} is missing: build it with staccato_amd.build.build_library() "
            "(hipcc --offload-arch=gfx950); the sweep has no CPU fallback")
    lib = ctypes.CDLL(p)
    vp, i, d, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    pi = ctypes.POINTER(ctypes.c_int)
    sig = {
        "stc_create": (i, [ctypes.POINTER(vp), i]),
        "stc_destroy": (i, [vp]),
        "stc_strerror": (ctypes.c_char_p, [i]),
        "stc_device_ok": (i, [vp]),
        "stc_set_policy": (i, [vp, i, i, vp]),
        "stc_clear_policies": (i, [vp]),
        "stc_rom_set": (i, [vp, i, i, vp, vp, vp, i, d]),
        "stc_rom_set_io": (i, [vp, i, i, vp, i, i, vp, i]),
        "stc_rom_clear": (i, [vp, i]),
        "stc_sweep": (i, [vp, i, i, vp, i, vp, i, vp, i, vp, vp, i, d, i, vp]),
        "stc_sweep_dev": (i, [vp, i, i, vp, i, vp, i, vp, i, vp, vp, i, d, i, vp]),
        "stc_sweep_dev_noalloc": (i, [vp, i, i, vp, i, vp, i, vp, i, vp, vp, i, d, i, vp]),
        "stc_getrf": (i, [vp, i, i, vp, i, vp, vp, vp]),
        "stc_sweep_batch": (i, [vp, i, vp, i, d, i, vp]),
        "stc_plan_create": (i, [vp, i, vp, i, d, i, ctypes.POINTER(vp)]),
        "stc_plan_execute": (i, [vp, vp]),
        "stc_plan_destroy": (i, [vp]),
        "stc_adjoint": (i, [vp, i, i, vp, i, vp, i, vp, i, vp, i, vp, i, d, i,
                            vp, vp, vp, vp]),
        "stc_back_transform": (i, [vp, i, i, vp, vp, i, vp, vp, vp, i, d, i, vp]),
        "stc_gesv_batched": (i, [vp, i, i, vp, i, ll, i, vp, i, ll, vp, i, ll, vp, vp, i, d,
                                 i, vp]),
        "stc_gesv_batched_dev": (i, [vp, i, i, vp, ll, i, vp, i, ll, vp, i, vp, vp, i, d, i,
                                     vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    del pi
    _lib = lib
    return lib


def strerror(code):
    return load().stc_strerror(code).decode()
