"""Synthetic inputs of the benchmark workloads (seeded, host arrays in cell
order z*X*Y + y*X + x), shared by bench.py and the tests.

  taylor_green      3-D Taylor-Green vortex, rho = 1 + 3 p' (D3Q19 256^3 workload)
  lid_cavity        D2Q9 lid-driven cavity ids (walls id 1, moving lid id 2)
  porous_spheres    non-overlapping solid spheres, radius U[rmin, rmax], periodic
                    in x, walls on the y and z boundary planes (D3Q19 384^3 workload)
  spinodal          rho = rho0 + U(-a, a) (Shan-Chen D3Q27 192^3 workload) with hot
                    (id 1, T=1) / cold (id 2, T=0) walls at y = 0 / Y-1
  porosity          channel porosity: scale x 3x3(x3) box-smoothed U[0,1]
"""
import numpy as np

SEED = 20261203


def taylor_green(n, u0=0.05, rho0=1.0):
    """u = u0 (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0), k = 2 pi / n;
    p' = (rho0 u0^2 / 16)(cos 2kx + cos 2ky)(cos 2kz + 2), rho = rho0 + 3 p'."""
    k = 2 * np.pi / n
    idx = np.arange(n, dtype=np.float64)
    sx, cx, c2 = np.sin(k * idx), np.cos(k * idx), np.cos(2 * k * idx)
    ux = u0 * sx[None, None, :] * cx[None, :, None] * cx[:, None, None]
    uy = -u0 * cx[None, None, :] * sx[None, :, None] * cx[:, None, None]
    p = (rho0 * u0 ** 2 / 16.0) * (c2[None, None, :] + c2[None, :, None]) * (c2[:, None, None] + 2.0)
    rho = (rho0 + 3.0 * p).astype(np.float32).ravel()
    u = np.zeros((n ** 3, 3), np.float32)
    u[:, 0] = ux.ravel()
    u[:, 1] = uy.ravel()
    return rho, u


def lid_cavity(X, Y, lid_id=2, wall_id=1):
    ids = np.zeros((Y, X), np.uint8)
    ids[:, 0] = wall_id
    ids[:, -1] = wall_id
    ids[0, :] = wall_id
    ids[-1, :] = lid_id
    return ids.ravel()


def porous_spheres(X, Y, Z, fraction=0.05, rmin=4.0, rmax=12.0, seed=SEED, walls=True, max_tries=200000):
    """Rejection sampling of non-overlapping spheres (integer centres, radius
    U[rmin, rmax], periodic in x) until the solid fraction reaches `fraction`;
    each sphere only touches its bounding box.  Returns uint8 ids (1 = solid)."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((Z, Y, X), bool)
    centres = np.zeros((0, 3))
    radii = np.zeros(0)
    count = 0
    target = fraction * X * Y * Z
    tries = 0
    while count < target and tries < max_tries:
        tries += 1
        r = rng.uniform(rmin, rmax)
        c = np.array([rng.integers(0, X), rng.integers(0, Y), rng.integers(0, Z)], float)
        if len(radii):
            d = centres - c
            d[:, 0] = np.minimum(np.abs(d[:, 0]), X - np.abs(d[:, 0]))
            if np.any(np.sqrt((d ** 2).sum(1)) < radii + r + 1.0):
                continue
        centres = np.vstack([centres, c])
        radii = np.append(radii, r)
        R = int(np.ceil(r))
        xs = np.arange(int(c[0]) - R, int(c[0]) + R + 1)
        ys = np.arange(max(0, int(c[1]) - R), min(Y, int(c[1]) + R + 1))
        zs = np.arange(max(0, int(c[2]) - R), min(Z, int(c[2]) + R + 1))
        dz, dy, dx = np.meshgrid(zs - c[2], ys - c[1], xs - c[0], indexing="ij")
        inside = dx ** 2 + dy ** 2 + dz ** 2 <= r * r
        box = solid[np.ix_(zs, ys, xs % X)]
        count += int(np.count_nonzero(inside & ~box))
        solid[np.ix_(zs, ys, xs % X)] = box | inside
    ids = solid.astype(np.uint8)
    if walls:
        ids[:, 0, :] = 1
        ids[:, -1, :] = 1
        ids[0, :, :] = 1
        ids[-1, :, :] = 1
    return ids.ravel()


def spinodal(X, Y, Z, rho0=2.0, amplitude=0.01, seed=SEED, thermal_walls=True):
    """rho = rho0 + U(-amplitude, amplitude); optional hot (y = 0, id 1) and
    cold (y = Y-1, id 2) walls.  Returns rho, u (zeros), ids or None."""
    rng = np.random.default_rng(seed)
    n = X * Y * Z
    rho = (rho0 + amplitude * (2.0 * rng.random(n) - 1.0)).astype(np.float32)
    u = np.zeros((n, 3), np.float32)
    ids = None
    if thermal_walls:
        ids = np.zeros((Z, Y, X), np.uint8)
        ids[:, 0, :] = 1
        ids[:, -1, :] = 2
        ids = ids.ravel()
    return rho, u, ids


def porosity(X, Y, Z=1, seed=SEED, scale=0.02):
    """scale x (3x3(x3) box-smoothed seeded U[0,1]), clipped to [0, 1], zero on
    the wall rows y = 0 and y = Y-1."""
    rng = np.random.default_rng(seed)
    a = rng.random((Z, Y, X))
    k = np.ones(3) / 3.0
    for ax in ((1, 2) if Z == 1 else (0, 1, 2)):
        a = np.apply_along_axis(lambda v: np.convolve(np.pad(v, 1, mode="edge"), k, mode="valid"), ax, a)
    g = np.clip(scale * a, 0.0, 1.0)
    g[:, 0, :] = 0.0
    g[:, -1, :] = 0.0
    return g.ravel().astype(np.float32)
