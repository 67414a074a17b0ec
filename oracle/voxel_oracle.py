"""Mesh voxeliser oracle (test infrastructure only; numpy, float32).

CPU restatement of bf_voxelize_mesh (bf_voxel.hip), which replaces
MeshRasterizer3D::rasterize (BoltzmannFluids/Source/MeshRasterizer3D/
MeshRasterizer3D.cpp:7-100, GLSL/MeshRasterizer3D/raster_unifier.comp):
  - vertices are transformed by the 4x4 model matrix (x' = ((m00 x + m01 y) + m02 z) + m03);
  - column (i, j) is the line x = (2i+1)/X - 1, y = (2j+1)/Y - 1;
  - each triangle's edge functions use the edge's endpoints in index order;
    an exact zero takes the sign of the perturbed line (x + e, y + e^2):
    -(yb - ya) when yb != ya, else (xb - xa);
  - a triangle is crossed when its three edge signs agree; the crossing is
    z = ((e0 z0 + e1 z1) + e2 z2) / ((e0 + e1) + e2);
  - sorted crossings pair into inside intervals; voxel k (z in
    [1 - (2k+2)/Z, 1 - 2k/Z]) is solid when an interval (za, zb) has
    za < hi and zb > lo.
All arithmetic in float32 in the same order as the device code.
"""
import numpy as np


def transform(vertices, model):
    v = np.asarray(vertices, np.float32)
    m = np.asarray(model, np.float32)
    out = np.empty_like(v)
    for r in range(3):
        out[:, r] = ((m[r, 0] * v[:, 0] + m[r, 1] * v[:, 1]) + m[r, 2] * v[:, 2]) + m[r, 3]
    return out


def crossings(vertices, triangles, resolution, model=None):
    """List of sorted crossing z per column (j, i) and the column grid."""
    X, Y, _ = resolution
    tv = transform(vertices, np.eye(4) if model is None else model)
    tri = np.asarray(triangles, np.int64).reshape(-1, 3)
    px = (np.arange(X, dtype=np.float32) * 2 + 1).astype(np.float32) / np.float32(X) - np.float32(1)
    py = (np.arange(Y, dtype=np.float32) * 2 + 1).astype(np.float32) / np.float32(Y) - np.float32(1)
    PY, PX = np.meshgrid(py, px, indexing="ij")
    PX = PX.ravel()[:, None]
    PY = PY.ravel()[:, None]
    e = []
    pos = []
    for c in range(3):
        c1 = (c + 1) % 3
        ia, ib = tri[:, c], tri[:, c1]
        rev = ia > ib
        a = np.where(rev[:, None], tv[ib], tv[ia])
        b = np.where(rev[:, None], tv[ia], tv[ib])
        ev = (b[:, 0] - a[:, 0])[None, :] * (PY - a[:, 1][None, :]) - (b[:, 1] - a[:, 1])[None, :] * (PX - a[:, 0][None, :])
        tie = np.where(b[:, 1] != a[:, 1], b[:, 1] < a[:, 1], b[:, 0] > a[:, 0])[None, :]
        p = np.where(ev == 0, tie, ev > 0)
        e.append(np.where(rev[None, :], -ev, ev))
        pos.append(np.where(rev[None, :], ~p, p))
    # weights: edge (c -> c+1) belongs to vertex (c + 2) % 3
    w = [e[1], e[2], e[0]]
    ps = [pos[1], pos[2], pos[0]]
    hit = (ps[0] & ps[1] & ps[2]) | (~ps[0] & ~ps[1] & ~ps[2])
    s = (w[0] + w[1]) + w[2]
    hit &= s != 0
    z = tv[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        zz = ((w[0] * z[tri[:, 0]][None, :] + w[1] * z[tri[:, 1]][None, :]) + w[2] * z[tri[:, 2]][None, :]) / s
    cols = []
    for r in range(hit.shape[0]):
        cols.append(np.sort(zz[r][hit[r]]))
    return cols


def voxelize(vertices, triangles, resolution, value=1, model=None, volume=None):
    X, Y, Z = resolution
    vol = np.zeros((Z, Y, X), np.uint8) if volume is None else volume.copy()
    k = np.arange(Z)
    hi = np.float32(1) - (2 * k).astype(np.float32) / np.float32(Z)
    lo = np.float32(1) - (2 * k + 2).astype(np.float32) / np.float32(Z)
    for col, zc in enumerate(crossings(vertices, triangles, resolution, model)):
        j, i = divmod(col, X)
        inside = np.zeros(Z, bool)
        for a in range(0, len(zc) - 1, 2):
            inside |= (zc[a] < hi) & (zc[a + 1] > lo)
        vol[inside, j, i] = value
    return vol


def cube_mesh(lo=-0.5, hi=0.5):
    """Axis-aligned box as 12 outward triangles."""
    v = np.array([[x, y, z] for z in (lo, hi) for y in (lo, hi) for x in (lo, hi)], np.float32)
    # corner index = x + 2y + 4z
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tri = []
    for a, b, c, d in quads:
        tri += [(a, b, c), (a, c, d)]
    return v, np.array(tri, np.uint32)
