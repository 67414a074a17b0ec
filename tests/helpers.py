"""Scenario builders shared by the tests (synthetic inputs, seeded)."""
import numpy as np

SEED = 20261203


def taylor_green(n, u0, velocity_set="D3Q19", rho0=1.0):
    """3-D Taylor-Green vortex on an n^3 periodic box (analytic pressure field)."""
    k = 2 * np.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ux = u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    uy = -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    uz = np.zeros_like(ux)
    p = (rho0 * u0 ** 2 / 16.0) * (np.cos(2 * k * x) + np.cos(2 * k * y)) * (np.cos(2 * k * z) + 2.0)
    rho = rho0 + 3.0 * p
    u = np.stack([ux, uy, uz], axis=-1).reshape(-1, 3)
    return rho.ravel().astype(np.float32), u.astype(np.float32)


def random_state(n_cells, seed=SEED, du=0.02, drho=0.02):
    rng = np.random.default_rng(seed)
    rho = (1.0 + drho * (rng.random(n_cells) - 0.5)).astype(np.float32)
    u = (du * (rng.random((n_cells, 3)) - 0.5)).astype(np.float32)
    return rho, u


def cavity_ids(X, Y, lid_id=2, wall_id=1):
    """Lid-driven cavity: walls at x=0, x=X-1, y=0; moving lid at y=Y-1 (cell order y*X+x)."""
    ids = np.zeros((Y, X), np.uint8)
    ids[:, 0] = wall_id
    ids[:, -1] = wall_id
    ids[0, :] = wall_id
    ids[-1, :] = lid_id
    return ids.ravel()


def channel_ids(X, Y, Z=1, wall_id=1):
    """Walls at y=0 and y=Y-1 (periodic along x and z)."""
    ids = np.zeros((Z, Y, X), np.uint8)
    ids[:, 0, :] = wall_id
    ids[:, -1, :] = wall_id
    return ids.ravel()


def spheres_ids(X, Y, Z, fraction, rmin, rmax, seed=SEED, walls_yz=True):
    """Non-overlapping solid spheres (integer centres, radius uniform in [rmin, rmax])
    added by rejection sampling until the solid fraction reaches `fraction`;
    optional walls on the y and z boundary planes."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((Z, Y, X), bool)
    spheres = []
    zz, yy, xx = np.ogrid[:Z, :Y, :X]
    tries = 0
    while solid.mean() < fraction and tries < 100000:
        tries += 1
        r = rng.uniform(rmin, rmax)
        c = np.array([rng.integers(0, X), rng.integers(0, Y), rng.integers(0, Z)])
        if any(np.linalg.norm(c - c2) < r + r2 + 1 for c2, r2 in spheres):
            continue
        spheres.append((c, r))
        dx = np.minimum(np.abs(xx - c[0]), X - np.abs(xx - c[0]))  # periodic in x
        solid |= (dx ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2) <= r * r
    ids = solid.astype(np.uint8)
    if walls_yz:
        ids[:, 0, :] = 1
        ids[:, -1, :] = 1
        ids[0, :, :] = 1
        ids[-1, :, :] = 1
    return ids.ravel()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def porosity_field(X, Y, Z=1, seed=SEED, scale=1.0):
    """Porosity gamma in [0, 1]: 3x3(x3) box-smoothed seeded uniform noise,
    clipped, zero on the wall rows y = 0 and y = Y-1."""
    rng = np.random.default_rng(seed)
    a = rng.random((Z, Y, X))
    k = np.ones(3) / 3.0
    for ax in (1, 2) if Z == 1 else (0, 1, 2):
        a = np.apply_along_axis(lambda v: np.convolve(np.pad(v, 1, mode="edge"), k, mode="valid"), ax, a)
    g = np.clip(scale * a, 0.0, 1.0)
    g[:, 0, :] = 0.0
    g[:, -1, :] = 0.0
    return g.ravel().astype(np.float32)
