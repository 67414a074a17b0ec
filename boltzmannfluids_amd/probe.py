"""ctypes binding of libbfprobe.so: device bandwidth probes for the benchmark
(csrc/bf_probe.h).  Measurement infrastructure, not part of the solver's ABI:
the probes replay the memory traffic of a tick without its arithmetic, giving
the ceilings that bench.py reports beside the solver's roofline figures."""
import ctypes as C
import os

from .capi import BFError, _share_hip_runtime_with_torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbfprobe.so")
_lib = None


def load_library(path=None):
    global _lib
    if _lib is not None and path is None:
        return _lib
    _share_hip_runtime_with_torch()
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with boltzmannfluids_amd/build.py (hipcc, gfx950)")
    lib = C.CDLL(p)
    i32, i64, dp = C.c_int32, C.c_int64, C.POINTER(C.c_double)
    for name, args in {"bfp_copy_bandwidth": [i32, i64, i32, dp],
                       "bfp_inplace_bandwidth": [i32, i64, i32, i32, dp],
                       "bfp_aa_bandwidth": [i32, i32, i32, i32, i32, i32, dp],
                       "bfp_pull_bandwidth": [i32, i32, i32, i32, i32, dp]}.items():
        fn = getattr(lib, name)
        fn.restype = i32
        fn.argtypes = args
    _lib = lib
    return lib


def _call(name, *args):
    out = C.c_double()
    st = getattr(load_library(), name)(*args, C.byref(out))
    if st:
        raise BFError(st, name)
    return out.value


def copy_bandwidth(nbytes=2 << 30, iters=20, device=0):
    """Sustained device copy bandwidth in GB/s (read + write)."""
    return _call("bfp_copy_bandwidth", int(device), int(nbytes), int(iters))


def inplace_bandwidth(cells, slots=19, iters=20, device=0):
    """GB/s (read + write) of the tick's access pattern without arithmetic:
    `slots` FP32 slots of `cells` cells in the lattice's slab layout, read and
    rewritten in place."""
    return _call("bfp_inplace_bandwidth", int(device), int(cells), int(slots), int(iters))


def aa_bandwidth(X, Y, Z, even, iters=20, device=0):
    """GB/s (read + write) of one D3Q19 AA tick's memory pattern on a periodic
    X x Y x Z grid in the slab layout, without arithmetic: the odd tick
    (even=False, every slot at the cell) or the even tick (even=True, slot i
    read in row (y - c_iy, z - c_iz), slot opp(i) written in row (y + c_iy,
    z + c_iz))."""
    return _call("bfp_aa_bandwidth", int(device), int(X), int(Y), int(Z), int(bool(even)), int(iters))


def pull_bandwidth(X, Y, Z, iters=20, device=0):
    """GB/s (read + write) of the two-lattice pull pattern for D3Q19 on a
    periodic X x Y x Z grid in the slab layout, without arithmetic: slot i read
    from one lattice in row (y - c_iy, z - c_iz), written to a second lattice
    at the cell itself."""
    return _call("bfp_pull_bandwidth", int(device), int(X), int(Y), int(Z), int(iters))
