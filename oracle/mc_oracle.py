"""Marching-cubes oracle (test infrastructure only; numpy).

CPU restatement of the iso-surface extraction that the library runs on the
device (bf_mc.hip), replacing MarchingCubes::compute
(BoltzmannFluids/Source/MarchingCubes/MarchingCubes.cpp:3-48,
GLSL/MarchingCubes/marching_cubes.comp:41-98), which the reference never
dispatches.  Conventions (the same as the library's documented output):
  corner k of a cube at (k & 1, k >> 1 & 1, k >> 2 & 1); edge 4a + j along axis
  a from the j-th corner without bit a; a point is inside when value >= threshold;
  vertices ordered by (grid point, axis), t = (thr - v0) / (v1 - v0) in float32,
  placed at (p / resolution - 1/2) * mesh_scale; triangles ordered by (cube,
  triangle of the case); normals point toward lower values.

The case table is generated here independently of the library: faces are
oriented by their outward normal (corners sorted by angle around it), each
inside run of a face's corners contributes the segment from its exit crossing
to its entry crossing, the segments chain into loops (starting from the
smallest edge id), and each loop is split into a fan from the first vertex
whose chords stay off the cube's faces.  The tests pin the table
by what must hold for any correct one: a closed, consistently oriented surface
around every configuration, the area and enclosed volume of a sphere, and the
Euler characteristic of a sphere and a torus.
"""
import math

import numpy as np


def _corner(k):
    return np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1], float)


def _edge_id(k0, k1):
    a = int(np.log2(k0 ^ k1))
    base = min(k0, k1)
    lower = [k for k in range(8) if not (k >> a) & 1]
    return 4 * a + lower.index(base)


def _edge_midpoint(e):
    a, j = divmod(e, 4)
    base = [k for k in range(8) if not (k >> a) & 1][j]
    p = _corner(base)
    p[a] = 0.5
    return p


def _faces(e):
    """The two cube faces (axis, side) that contain edge e."""
    a, j = divmod(e, 4)
    base = [k for k in range(8) if not (k >> a) & 1][j]
    return {(b, (base >> b) & 1) for b in range(3) if b != a}


def _face_cycle(a, side):
    """Corners of face (axis a, side) counter-clockwise seen from outside."""
    n = np.zeros(3)
    n[a] = 1.0 if side else -1.0
    corners = [k for k in range(8) if ((k >> a) & 1) == side]
    centre = np.mean([_corner(k) for k in corners], axis=0)
    # right-handed basis (u, v, n) of the face plane
    u = np.zeros(3)
    u[(a + 1) % 3] = 1.0
    v = np.cross(n, u)
    ang = [math.atan2(np.dot(_corner(k) - centre, v), np.dot(_corner(k) - centre, u)) for k in corners]
    return [k for _, k in sorted(zip(ang, corners))]


def case_table():
    """(triangle_count[256] uint8, edges[256, 15] int8, -1 padded)."""
    faces = [_face_cycle(a, s) for a in range(3) for s in (0, 1)]
    tris = []
    for case in range(256):
        inside = [(case >> k) & 1 for k in range(8)]
        succ = {}
        for cyc in faces:
            kinds = []
            for i in range(4):
                k0, k1 = cyc[i], cyc[(i + 1) % 4]
                kinds.append(0 if inside[k0] == inside[k1] else (1 if inside[k0] else -1))
            for i in range(4):
                if kinds[i] != 1:
                    continue
                j = next((i - s) % 4 for s in range(1, 5) if kinds[(i - s) % 4] == -1)
                succ[_edge_id(cyc[i], cyc[(i + 1) % 4])] = _edge_id(cyc[j], cyc[(j + 1) % 4])
        out, seen = [], set()
        for e0 in sorted(succ):
            if e0 in seen:
                continue
            loop, e = [], e0
            while e not in seen:
                seen.add(e)
                loop.append(e)
                e = succ[e]
            # fan apex: the first loop vertex none of whose chords joins two
            # edges of one cube face (such a chord would lie in the face and
            # could coincide with the neighbouring cube's)
            L = len(loop)
            for r in range(L):
                rot = loop[r:] + loop[:r]
                if not any(_faces(rot[0]) & _faces(rot[i]) for i in range(2, L - 1)):
                    loop = rot
                    break
            out += [(loop[0], loop[i], loop[i + 1]) for i in range(1, L - 1)]
        tris.append(out)
    # orientation: the triangle around corner 0 alone must face away from it
    p = [_edge_midpoint(e) for e in tris[1][0]]
    if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[0]) < 0:
        tris = [[(a, c, b) for a, b, c in t] for t in tris]
    count = np.array([len(t) for t in tris], np.uint8)
    edges = -np.ones((256, 15), np.int8)
    for c, t in enumerate(tris):
        flat = [e for tri in t for e in tri]
        edges[c, :len(flat)] = flat
    return count, edges


def marching_cubes(volume, threshold, mesh_scale=(1.0, 1.0, 1.0), table=None):
    """volume: (Z, Y, X) float32.  Returns vertices (V, 3) float32, triangles (T, 3) uint32."""
    vol = np.ascontiguousarray(volume, np.float32)
    Z, Y, X = vol.shape
    thr = np.float32(threshold)
    scale = np.asarray(mesh_scale, np.float32)
    res = np.array([X, Y, Z], np.float32)
    count, edges = table if table is not None else case_table()
    inside = vol >= thr
    flags = np.zeros(vol.shape, np.uint8)
    flags[:, :, :-1] |= (inside[:, :, :-1] != inside[:, :, 1:]).astype(np.uint8)
    flags[:, :-1, :] |= (inside[:, :-1, :] != inside[:, 1:, :]).astype(np.uint8) << 1
    flags[:-1, :, :] |= (inside[:-1, :, :] != inside[1:, :, :]).astype(np.uint8) << 2
    f = flags.ravel()
    nvert = ((f & 1) + ((f >> 1) & 1) + ((f >> 2) & 1)).astype(np.int64)
    first = np.concatenate([[0], np.cumsum(nvert)[:-1]])
    v = vol.ravel()
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    coords = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1).astype(np.float32)
    step = [1, X, X * Y]
    keys, pos = [], []
    for a in range(3):
        pts = np.nonzero((f >> a) & 1)[0]
        v0, v1 = v[pts], v[pts + step[a]]
        t = (thr - v0) / (v1 - v0)
        q = coords[pts].copy()
        q[:, a] = q[:, a] + t
        pos.append((q / res - np.float32(0.5)) * scale)
        keys.append(pts * 3 + a)
    keys = np.concatenate(keys)
    order = np.argsort(keys, kind="stable")
    vertices = np.concatenate(pos)[order].astype(np.float32)
    # cubes
    cs = np.zeros((Z - 1, Y - 1, X - 1), np.int64)
    for k in range(8):
        dx, dy, dz = k & 1, (k >> 1) & 1, (k >> 2) & 1
        cs |= inside[dz:Z - 1 + dz, dy:Y - 1 + dy, dx:X - 1 + dx].astype(np.int64) << k
    cs = cs.ravel()
    active = np.nonzero(count[cs] > 0)[0]
    cz, rem = np.divmod(active, (X - 1) * (Y - 1))
    cy, cx = np.divmod(rem, X - 1)
    origin = cx + X * (cy + Y * cz)
    ent = edges[cs[active]].astype(np.int64)           # (n, 15)
    valid = ent >= 0
    a = np.where(valid, ent >> 2, 0)
    j = np.where(valid, ent & 3, 0)
    b0 = np.where(a == 0, 1, 0)
    b1 = np.where(a == 2, 1, 2)
    off = np.zeros(ent.shape + (3,), np.int64)
    np.put_along_axis(off, b0[..., None], (j & 1)[..., None], axis=2)
    np.put_along_axis(off, b1[..., None], (j >> 1)[..., None], axis=2)
    q = origin[:, None] + off[..., 0] + X * (off[..., 1] + Y * off[..., 2])
    fq = f[q].astype(np.int64)
    below = fq & ((1 << a) - 1)
    idx = first[q] + (below & 1) + ((below >> 1) & 1)
    tri = idx[valid].astype(np.uint32).reshape(-1, 3)
    return vertices, tri


def mesh_edges(tri):
    """Directed edges of all triangles as (n, 2)."""
    return np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])


def area_volume(vertices, tri):
    p0, p1, p2 = (vertices[tri[:, k]].astype(np.float64) for k in range(3))
    cr = np.cross(p1 - p0, p2 - p0)
    area = 0.5 * np.linalg.norm(cr, axis=1).sum()
    volume = np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6.0
    return area, volume
