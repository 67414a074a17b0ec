"""Mesh voxeliser on the device (bf_voxelize_mesh) against the numpy oracle,
voxel for voxel, including transformed meshes and exact ties."""
import numpy as np
import pytest

import mc_oracle as mc
import voxel_oracle as vx

pytestmark = pytest.mark.gpu


def sphere_mesh(n, r, c):
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    vol = (r - np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)).astype(np.float32)
    return mc.marching_cubes(vol, 0.0, (2.0, 2.0, 2.0))


def rotation(ax, ay, az, shift):
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    m = np.eye(4)
    m[:3, :3] = rz @ ry @ rx
    m[:3, 3] = shift
    return m.astype(np.float32)


@pytest.mark.parametrize("case", ["box", "ties", "sphere_rotated", "shell"])
def test_matches_oracle(bfa, case):
    model = None
    if case == "box":
        v, t = vx.cube_mesh(-0.43, 0.57)
        res = (32, 24, 20)
    elif case == "ties":
        v, t = vx.cube_mesh(-0.25, 0.25)
        res = (4, 4, 8)
    elif case == "sphere_rotated":
        v, t = sphere_mesh(36, 12.0, (17.3, 18.1, 17.7))
        res = (40, 44, 36)
        model = rotation(0.3, -0.7, 1.1, (0.1, -0.05, 0.2))
    else:
        v1, t1 = sphere_mesh(40, 14.0, (19.6, 20.2, 19.9))
        v2, t2 = sphere_mesh(40, 7.0, (19.6, 20.2, 19.9))
        v, t = np.concatenate([v1, v2]), np.concatenate([t1, t2 + len(v1)])
        res = (40, 40, 40)
    print(f"case: voxelise {case} {res} {len(t)} triangles", flush=True)
    got = bfa.voxelize_mesh(v, t, res, value=3, model=model)
    ref = vx.voxelize(v, t, res, value=3, model=model)
    assert np.array_equal(got, ref) and got.sum() > 0


def test_paints_over_existing_ids(bfa):
    v, t = vx.cube_mesh(-0.43, 0.57)
    res = (16, 16, 16)
    vol = np.zeros((16, 16, 16), np.uint8)
    vol[0] = 1                      # e.g. a wall plane
    out = bfa.voxelize_mesh(v, t, res, value=2, volume=vol)
    assert out is vol and np.all(vol[0] == 1)
    assert np.array_equal(vol == 2, vx.voxelize(v, t, res, value=2) == 2)


def test_rejects_bad_indices(bfa):
    v, t = vx.cube_mesh()
    t = t.copy()
    t[3, 1] = 99
    with pytest.raises(bfa.BFError, match="out of range"):
        bfa.voxelize_mesh(v, t, (8, 8, 8))
