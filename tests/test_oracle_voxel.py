"""Mesh voxeliser oracle, pinned by geometry (no reference output exists: the
reference's MeshRasterizer3D is only called from commented-out code).

- an axis-aligned box gives exactly the voxels the closed form predicts;
- columns through the box's edges and vertices (exact ties) are decided by the
  perturbation rule: every column crosses the closed mesh an even number of
  times and the box occupies exactly the predicted columns;
- a sphere mesh (marching cubes of a distance field) fills 4/3 pi r^3 + pi r^2
  voxels (the z-overlap rule adds half a voxel at each end of a column) to 1 %;
- a hollow shell (outer sphere plus inner sphere) leaves its cavity empty.
"""
import numpy as np

import mc_oracle as mc
import voxel_oracle as vx


def expected_box(res, lo, hi):
    X, Y, Z = res
    cx = (np.arange(X) * 2 + 1) / X - 1
    cy = (np.arange(Y) * 2 + 1) / Y - 1
    k = np.arange(Z)
    zhi, zlo = 1 - 2 * k / Z, 1 - (2 * k + 2) / Z
    inz = (zhi > lo) & (zlo < hi)
    inx = (cx > lo) & (cx < hi)
    iny = (cy > lo) & (cy < hi)
    return (inz[:, None, None] & iny[None, :, None] & inx[None, None, :]).astype(np.uint8)


def test_box_matches_closed_form():
    # faces away from voxel boundaries (a face on a boundary is a float tie in z)
    v, t = vx.cube_mesh(-0.43, 0.57)
    res = (32, 24, 20)
    vol = vx.voxelize(v, t, res)
    assert np.array_equal(vol, expected_box(res, -0.43, 0.57))


def test_ties_on_edges_and_vertices_keep_parity():
    # columns at x, y in {-0.75, -0.25, 0.25, 0.75}; the box faces sit on them
    v, t = vx.cube_mesh(-0.25, 0.25)
    res = (4, 4, 8)
    cols = vx.crossings(v, t, res)
    assert all(len(c) % 2 == 0 for c in cols)
    vol = vx.voxelize(v, t, res)
    # perturbed columns: x = -0.25 + e is inside, x = 0.25 + e is outside
    solid_cols = np.argwhere(vol.any(axis=0))
    assert solid_cols.tolist() == [[1, 1]]
    assert vol[:, 1, 1].sum() == 2           # z in (-0.25, 0.25) covers two of 8 slices


def sphere_mesh(n=40, r=14.0):
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    vol = (r - np.sqrt((x - 19.6) ** 2 + (y - 20.2) ** 2 + (z - 19.9) ** 2)).astype(np.float32)
    v, t = mc.marching_cubes(vol, 0.0, (2.0, 2.0, 2.0))     # coordinates in [-1, 1]
    return v, t


def test_sphere_volume():
    """A voxel counts when its z range touches the inside interval of its
    column, so each column gains half a voxel at either end on average: the
    count is the ball's volume plus its projected area, pi r^2 (in voxels)."""
    v, t = sphere_mesh()
    res = (48, 48, 48)
    vol = vx.voxelize(v, t, res)
    r = 14.0 / 40 * 2 * 24
    expected = 4 / 3 * np.pi * r ** 3 + np.pi * r ** 2
    assert abs(vol.sum() / expected - 1) < 0.01


def test_hollow_shell():
    v1, t1 = sphere_mesh(40, 14.0)
    v2, t2 = sphere_mesh(40, 7.0)
    v = np.concatenate([v1, v2])
    t = np.concatenate([t1, t2 + len(v1)])
    vol = vx.voxelize(v, t, (40, 40, 40))
    assert vol[20, 20, 20] == 0 and vol.sum() > 0
    ring = vol[20, 20]
    assert ring[20 + 10] == 1 and ring[20 - 10] == 1
