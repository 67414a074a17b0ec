"""Marching-cubes oracle and case table, pinned by properties every correct
table has (no reference output exists: MarchingCubes::compute,
MarchingCubes.cpp:3-48, is never dispatched by the reference).

- the library's table equals the oracle's independently generated one;
- every one of the 256 corner patterns, embedded in an outside neighbourhood,
  gives a closed surface whose triangles are consistently oriented (each
  directed edge once, its reverse once);
- random fields: the same, with the surface crossing many cubes of every case;
- a sphere: area and enclosed volume within 0.5 % of 4 pi r^2 and 4/3 pi r^3,
  positive volume (normals point toward lower values), Euler characteristic 2;
- a torus: Euler characteristic 0.
"""
import numpy as np
import pytest

import mc_oracle as mc


@pytest.fixture(scope="module")
def table():
    return mc.case_table()


def closed_and_oriented(tri):
    e = mc.mesh_edges(tri).astype(np.int64)
    key = e[:, 0] * (1 << 32) + e[:, 1]
    rev = e[:, 1] * (1 << 32) + e[:, 0]
    return len(np.unique(key)) == len(key) and bool(np.isin(rev, key).all())


def euler(vertices, tri):
    return len(vertices) - len(mc.mesh_edges(tri)) // 2 + len(tri)


def test_library_table_matches_oracle(bfa, table):
    count, edges = bfa.marching_cubes_table()
    assert np.array_equal(count, table[0])
    assert np.array_equal(edges, table[1])


def test_table_shape(table):
    count, edges = table
    assert count[0] == 0 and count[255] == 0
    assert count.max() == 5
    for c in range(256):
        used = edges[c, :3 * count[c]]
        assert np.all((used >= 0) & (used < 12)) and np.all(edges[c, 3 * count[c]:] == -1)
        # every triangle edge id is an edge whose end corners differ in state
        for e in used:
            a, j = divmod(int(e), 4)
            k0 = [k for k in range(8) if not (k >> a) & 1][j]
            assert ((c >> k0) & 1) != ((c >> (k0 | 1 << a)) & 1)


def test_every_case_gives_a_closed_oriented_surface(table):
    for c in range(1, 255):
        vol = np.zeros((4, 4, 4), np.float32)
        for k in range(8):
            vol[1 + ((k >> 2) & 1), 1 + ((k >> 1) & 1), 1 + (k & 1)] = 1.0 if (c >> k) & 1 else 0.0
        v, tri = mc.marching_cubes(vol, 0.5, table=table)
        assert len(tri) > 0 and closed_and_oriented(tri), c
        _, volume = mc.area_volume(v, tri)
        assert volume > 0, c


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_fields_are_closed(table, seed):
    rng = np.random.default_rng(seed)
    vol = rng.random((22, 19, 17)).astype(np.float32)
    vol[0] = vol[-1] = 0
    vol[:, 0] = vol[:, -1] = 0
    vol[:, :, 0] = vol[:, :, -1] = 0
    v, tri = mc.marching_cubes(vol, 0.55, table=table)
    assert len(tri) > 1000 and closed_and_oriented(tri)
    assert mc.area_volume(v, tri)[1] > 0


def sphere(n=48, r=15.0, c=(23.3, 24.1, 23.7)):
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    return (r - np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)).astype(np.float32)


def test_sphere_area_volume_euler(table):
    n, r = 48, 15.0
    v, tri = mc.marching_cubes(sphere(n, r), 0.0, (n, n, n), table=table)
    area, volume = mc.area_volume(v, tri)
    assert abs(area / (4 * np.pi * r * r) - 1) < 5e-3
    assert abs(volume / (4 / 3 * np.pi * r ** 3) - 1) < 5e-3
    assert closed_and_oriented(tri) and euler(v, tri) == 2


def test_torus_euler(table):
    n = 48
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64) - 23.5
    q = np.sqrt(x * x + y * y) - 12.0
    vol = (5.0 - np.sqrt(q * q + z * z)).astype(np.float32)
    v, tri = mc.marching_cubes(vol, 0.0, table=table)
    assert closed_and_oriented(tri) and euler(v, tri) == 0


def test_vertices_lie_on_the_iso_level(table):
    """Trilinear interpolation is exact along an edge: the field sampled at each
    vertex by linear interpolation along its edge equals the threshold."""
    rng = np.random.default_rng(7)
    vol = rng.random((9, 8, 7)).astype(np.float32)
    v, _ = mc.marching_cubes(vol, 0.4, (7, 8, 9), table=table)
    p = v + np.array([3.5, 4.0, 4.5], np.float32)         # back to grid units
    base = np.floor(p + 1e-4).astype(int)
    frac = p - base
    axis = np.argmax(frac > 1e-4, axis=1)
    lo = vol[base[:, 2], base[:, 1], base[:, 0]]
    hi_idx = base.copy()
    hi_idx[np.arange(len(p)), axis] += 1
    hi = vol[hi_idx[:, 2], hi_idx[:, 1], hi_idx[:, 0]]
    t = frac[np.arange(len(p)), axis]
    assert np.allclose(lo + t * (hi - lo), 0.4, atol=2e-5)
