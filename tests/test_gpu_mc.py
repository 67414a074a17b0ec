"""Iso-surface extraction on the device (bf_marching_cubes / bf_isosurface)
against the numpy oracle: vertices and triangle indices bit for bit."""
import numpy as np
import pytest

import mc_oracle as mc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table():
    return mc.case_table()


def smooth_random(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.random(shape)
    for ax in range(3):
        a = (a + np.roll(a, 1, ax) + np.roll(a, -1, ax)) / 3.0
    return a.astype(np.float32)


def check_equal(bfa, table, vol, thr, scale):
    v, t = bfa.marching_cubes(vol, thr, scale)
    ov, ot = mc.marching_cubes(vol, thr, scale, table=table)
    assert v.shape == ov.shape and t.shape == ot.shape
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(t, ot)
    return v, t


def test_sphere_matches_oracle(bfa, table):
    n = 48
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    vol = (15.0 - np.sqrt((x - 23.3) ** 2 + (y - 24.1) ** 2 + (z - 23.7) ** 2)).astype(np.float32)
    v, t = check_equal(bfa, table, vol, 0.0, (1.0, 1.0, 1.0))
    assert len(t) > 8000


@pytest.mark.parametrize("shape,thr,seed", [((33, 37, 40), 0.5, 1), ((17, 64, 65), 0.47, 2), ((5, 6, 7), 0.5, 3)])
def test_random_fields_match_oracle(bfa, table, shape, thr, seed):
    check_equal(bfa, table, smooth_random(shape, seed), thr, (2.0, 0.5, 1.25))


def test_degenerate_grids(bfa, table):
    # a single z-plane has vertices on its in-plane edges but no cubes
    vol = smooth_random((1, 30, 31), 4)
    v, t = check_equal(bfa, table, vol, 0.5, (1, 1, 1))
    assert len(v) > 0 and len(t) == 0
    # nothing crosses the level
    v, t = bfa.marching_cubes(np.ones((8, 8, 8), np.float32), 0.5)
    assert len(v) == 0 and len(t) == 0


def test_capacity_checks(bfa):
    import ctypes as C
    lib = bfa.load_library()
    vol = smooth_random((16, 16, 16), 5)
    res = (C.c_int32 * 3)(16, 16, 16)
    sc = np.ones(3, np.float32)
    nv, ni = C.c_int64(), C.c_int64()
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    assert lib.bf_marching_cubes(0, fp(vol), res, 0.5, fp(sc), None, 0, None, 0, C.byref(nv), C.byref(ni)) == 0
    assert nv.value > 0 and ni.value > 0
    small = np.zeros((nv.value - 1, 3), np.float32)
    idx = np.zeros(ni.value, np.uint32)
    st = lib.bf_marching_cubes(0, fp(vol), res, 0.5, fp(sc), fp(small), nv.value - 1,
                               idx.ctypes.data_as(C.POINTER(C.c_uint32)), ni.value, C.byref(nv), C.byref(ni))
    assert st == 1 and b"output too small" in lib.bf_marching_cubes_last_error()
    mv, mi = C.c_int64(), C.c_int64()
    assert lib.bf_marching_cubes_capacity(res, C.byref(mv), C.byref(mi)) == 0
    assert mv.value >= nv.value and mi.value >= ni.value


def test_isosurface_of_a_shan_chen_droplet(bfa, table):
    """The solver's resident density (last save tick) gives the same mesh as
    the oracle applied to the density read back to the host."""
    n = 40
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    r = np.sqrt((x - 19.5) ** 2 + (y - 19.5) ** 2 + (z - 19.5) ** 2)
    rho = (0.12 + 1.77 * 0.5 * (1 - np.tanh((r - 10) / 1.5))).astype(np.float32).ravel()
    s = bfa.Solver((n, n, n), 1.0, velocity_set="D3Q19", multiphase=True)
    s.set_param("interaction_strength", -5.0)
    s.init_fields(rho, np.zeros((n ** 3, 3), np.float32))
    s.step(200, save=True)
    dens = s.read("velocity_density")[:, 3].reshape(n, n, n).copy()
    v, t = s.isosurface(1.0, (n, n, n))
    s.close()
    ov, ot = mc.marching_cubes(dens, 1.0, (n, n, n), table=table)
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32)) and np.array_equal(t, ot)
    area, volume = mc.area_volume(v, t)
    radius = (3 * volume / (4 * np.pi)) ** (1 / 3)
    assert 8.5 < radius < 11.5 and abs(area / (4 * np.pi * radius ** 2) - 1) < 0.05
