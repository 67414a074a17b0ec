"""bf_marching_cubes_noalloc (device field in, device mesh out, work buffers
from the library's fixed device array) against bf_marching_cubes (host
arrays, buffers of its own), through the C ABI: identical vertices and
indices on two calls in a row, the second on a smaller grid, then the first
grid again (the shared array keeps the last call's contents, which no pass
may read); a grid above BF_MC_NOALLOC_MAX_POINTS is rejected before any
device work."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 4194304


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def field(res, seed):
    """Two overlapping blobs plus seeded noise: closed surfaces, many cases."""
    X, Y, Z = res
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    rng = np.random.default_rng(seed)
    f = (np.exp(-((x - 0.4 * X) ** 2 + (y - 0.5 * Y) ** 2 + (z - 0.5 * Z) ** 2) / (0.05 * X * X))
         + np.exp(-((x - 0.65 * X) ** 2 + (y - 0.45 * Y) ** 2 + (z - 0.55 * Z) ** 2) / (0.03 * X * X))
         + 0.05 * rng.random((Z, Y, X)))
    return f.astype(np.float32).ravel()


def noalloc_lib(bfa):
    lib = bfa.load_library()
    fn = lib.bf_marching_cubes_noalloc
    fn.restype = C.c_int32
    fn.argtypes = [C.c_int32, C.c_void_p, C.POINTER(C.c_int32), C.c_float, C.POINTER(C.c_float), C.c_void_p,
                   C.c_int64, C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p]
    return lib, fn


def run_noalloc(bfa, vol, res, thr, scale):
    import torch
    lib, fn = noalloc_lib(bfa)
    d_vol = torch.from_numpy(vol).cuda()
    r = (C.c_int32 * 3)(*res)
    sc = (C.c_float * 3)(*scale)
    nv, ni = C.c_int64(), C.c_int64()
    torch.cuda.synchronize()
    st = fn(0, d_vol.data_ptr(), r, thr, sc, None, 0, None, 0, C.byref(nv), C.byref(ni), None)
    assert st == 0, lib.bf_marching_cubes_last_error()
    d_v = torch.empty(max(3 * nv.value, 1), dtype=torch.float32, device="cuda")
    d_i = torch.empty(max(ni.value, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = fn(0, d_vol.data_ptr(), r, thr, sc, d_v.data_ptr(), nv.value, d_i.data_ptr(), ni.value, C.byref(nv),
            C.byref(ni), None)
    assert st == 0, lib.bf_marching_cubes_last_error()
    torch.cuda.synchronize()
    v = d_v.cpu().numpy()[: 3 * nv.value].reshape(-1, 3)
    i = d_i.cpu().numpy()[: ni.value].view(np.uint32).reshape(-1, 3)
    return v, i


@pytest.mark.parametrize("thr", [0.5])
def test_noalloc_equals_allocating_entry_point(bfa, thr):
    scale = (2.0, 1.5, 1.0)
    cases = [((96, 80, 72), 1), ((40, 33, 21), 2), ((96, 80, 72), 1)]
    for res, seed in cases:
        vol = field(res, seed)
        say(f"case: noalloc marching cubes {res}, threshold {thr}")
        v_ref, i_ref = bfa.marching_cubes(vol.reshape(res[2], res[1], res[0]), thr, mesh_scale=scale)
        v, i = run_noalloc(bfa, vol, res, thr, scale)
        say(f"  {len(v)} vertices, {len(i)} triangles (reference {len(v_ref)}, {len(i_ref)})")
        assert len(v) > 1000 and len(i) > 1000
        assert np.array_equal(v, np.asarray(v_ref, np.float32).reshape(-1, 3))
        assert np.array_equal(i, np.asarray(i_ref, np.uint32).reshape(-1, 3))


def test_noalloc_rejects_grids_above_the_limit(bfa):
    lib, fn = noalloc_lib(bfa)
    res = (256, 256, 65)
    assert res[0] * res[1] * res[2] > MAX_POINTS
    r = (C.c_int32 * 3)(*res)
    sc = (C.c_float * 3)(1.0, 1.0, 1.0)
    nv, ni = C.c_int64(), C.c_int64()
    # a non-null (never dereferenced) volume pointer: the size check comes first
    st = fn(0, C.c_void_p(256), r, 0.5, sc, None, 0, None, 0, C.byref(nv), C.byref(ni), None)
    assert st == 1
    assert b"allocation-free" in lib.bf_marching_cubes_last_error()
