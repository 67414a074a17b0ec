"""GPU parity of the stability branches: density floor (collide.comp:704-705),
velocity clamp (:722-725) and equilibrium reset (:727-745), against the FP32
oracle (whose branches are pinned by closed forms in test_oracle_limits.py).

The seeded states mix three kinds of cells: blocks of density 0.003 (their
streamed density stays below the floor), and cells whose velocity has a random
direction and a magnitude uniform in [0, 0.45] (after streaming, cells fall
below the clamp, between velocity_limit and velocity_limit_extreme, and above
it).  Before the compared tick the test classifies every fluid cell from the
populations that tick reads (BF_FIELD_POPULATIONS, no force: u = sum c f / rho
with the floored rho) and requires each branch to be taken by several cells.

Bars: FP32 storage bit-identical to the FP32 oracle (populations and saved
fields); FP16/BF16 storage identical to the oracle's emulated storage except
for one-quantum rounding flips (at least 98 % identical, every difference
within two quanta of the value's own magnitude).
"""
import sys

import numpy as np
import pytest

from helpers import SEED

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def limit_state(shape, vs, seed):
    X, Y, Z = shape
    n = X * Y * Z
    rng = np.random.default_rng(seed)
    rho = (1.0 + 0.05 * (rng.random(n) - 0.5)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    if vs == "D2Q9":
        d[:, 2] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u = (d * (0.45 * rng.random(n))[:, None]).astype(np.float32)
    r3 = rho.reshape(Z, Y, X)
    for k in range(max(2, n // (150 if Z == 1 else 400))):
        x0, y0 = rng.integers(0, X - 3), rng.integers(0, Y - 3)
        z0 = 0 if Z == 1 else rng.integers(0, Z - 3)
        r3[z0:z0 + (1 if Z == 1 else 3), y0:y0 + 3, x0:x0 + 3] = 0.003
    return rho, u


def classify(g, vs, fluid, ulim=0.25, uext=0.30):
    """Counts of fluid cells by branch for the next tick (no force)."""
    import boltzmannfluids_amd as bfa
    f = g.read("populations").astype(np.float64)
    tab = bfa.velocity_table(vs)
    r = f.sum(axis=0)
    rf = np.maximum(r, 0.01)
    un = np.linalg.norm((tab[:, :3].T.astype(np.float64) @ f) / rf, axis=0)
    floor = (r < 0.01) & fluid
    clamp = (un > ulim) & (un <= uext) & fluid
    reset = (un > uext) & fluid
    return int(floor.sum()), int(clamp.sum()), int(reset.sum())


def compare(g, o, storage, vs, orc):
    a, b = g.read("populations_raw").astype(np.float64), o.read("populations_raw")
    d = np.abs(a - b)
    same = float(np.mean(a == b))
    say(f"  raw max|diff|={d.max():.3e} identical={same:.5f}")
    if storage == "fp32":
        assert np.array_equal(a, b)
        assert np.array_equal(g.read("velocity_density").astype(np.float64), o.read("velocity_density"))
    else:
        q = 2.0 ** -11 if storage == "fp16" else 2.0 ** -8
        w = orc.velocity_table(vs)[:, 3][:, None]
        assert same > 0.98
        assert np.all(d <= 2 * q * np.maximum(np.abs(b - w), 2.0 ** -14) + 1e-9)


CASES = [("D2Q9", "fp32", False), ("D2Q9", "fp32", True), ("D2Q9", "fp16", True),
         ("D3Q19", "fp32", False), ("D3Q19", "fp32", True), ("D3Q19", "fp16", False), ("D3Q19", "fp16", True),
         ("D3Q19", "bf16", True), ("D3Q27", "fp32", False), ("D3Q27", "fp32", True), ("D3Q27", "fp16", True)]


@pytest.mark.parametrize("vs,storage,walls", CASES)
def test_floor_clamp_reset_against_oracle(bfa, orc, vs, storage, walls):
    shape = (48, 40, 1) if vs == "D2Q9" else (20, 16, 12)
    X, Y, Z = shape
    n = X * Y * Z
    rho, u = limit_state(shape, vs, SEED + 41)
    init = dict(rho=rho, u=u)
    periodic = (1, 1, 1)
    objects = []
    fluid = np.ones(n, bool)
    if walls:
        rng = np.random.default_rng(SEED + 43)
        ids = np.where(rng.random(n) < 0.1, rng.integers(1, 3, n), 0).astype(np.uint8)
        init["boundary_id"] = ids
        objects = [(2, dict(velocity=(0.05, -0.02, 0.0)))]
        periodic = (1, 0, 1) if vs == "D2Q9" else (1, 0, 0)
        fluid = ids == 0
    say(f"case: {vs} {storage} walls={walls} {shape}")
    g = bfa.Solver(shape, 0.6, periodic=periodic, velocity_set=vs, storage=storage)
    o = orc.Oracle(shape, 0.6, periodic=periodic, velocity_set=vs, storage=storage, precision="f32")
    for oid, kw in objects:
        g.set_object(oid, **kw)
        o.set_object(oid, **kw)
    g.init_fields(**init)
    o.init_fields(**init)
    seen = np.zeros(3, int)
    for t, ticks in enumerate((1, 1, 3)):
        counts = classify(g, vs, fluid)
        say(f"  before tick {g.ticks}: floor {counts[0]}, clamp band {counts[1]}, reset {counts[2]}")
        if t == 0:
            assert min(counts) >= 3, counts
        seen += counts
        g.step(ticks, save=True)
        o.step(ticks, save=True)
        compare(g, o, storage, vs, orc)
    g.close()
    o.close()


@pytest.mark.parametrize("vs", ["D2Q9", "D3Q19"])
def test_lowered_limits_against_oracle(bfa, orc, vs):
    """The reference's thresholds as parameters: velocity_limit 0.03 and
    velocity_limit_extreme 0.04 on a low-Mach state (velocity components in
    +-0.08), so that clamps and resets happen throughout a 20-tick run with a
    constant force; bit-identical to the oracle."""
    shape = (40, 24, 1) if vs == "D2Q9" else (20, 12, 10)
    n = int(np.prod(shape))
    rng = np.random.default_rng(SEED + 47)
    rho = (1.0 + 0.04 * (rng.random(n) - 0.5)).astype(np.float32)
    u = (0.16 * (rng.random((n, 3)) - 0.5)).astype(np.float32)
    if vs == "D2Q9":
        u[:, 2] = 0
    force = np.tile(np.array([2e-4, -1e-4, 0], np.float32), (n, 1))
    g = bfa.Solver(shape, 0.7, velocity_set=vs)
    o = orc.Oracle(shape, 0.7, velocity_set=vs, precision="f32")
    for s in (g, o):
        s.set_param("velocity_limit", 0.03)
        s.set_param("velocity_limit_extreme", 0.04)
        s.init_fields(rho, u, force=force)
    g.step(1)
    o.step(1)
    _, clamp, reset = classify(g, vs, np.ones(n, bool), 0.03, 0.04)
    say(f"  {vs}: before tick 1 (force not included): clamp band {clamp}, reset {reset}")
    assert clamp > 10 and reset > 10
    g.step(19, save=True)
    o.step(19, save=True)
    compare(g, o, "fp32", vs, orc)
    un = np.linalg.norm(g.read("velocity_density")[:, :3].astype(np.float64), axis=1)
    assert un.max() <= 0.03 * (1 + 1e-6)


def droplets(n):
    """demo3d::multiphase_droplet_collision (LBMDemo3D.cpp:94-131) scaled from
    256^3 to n^3: gas density 0.056, two droplets of density 2.659 and radius
    24 n / 256 at (1.2 n/4, n/2, n/2 + 10 n/256) and (2.8 n/4, n/2, n/2 - 10 n/256),
    moving at +-1.5 along x."""
    s = n / 256.0
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = np.full((n, n, n), 0.056, np.float32)
    u = np.zeros((n, n, n, 3), np.float32)
    for cx, cz, vx in ((n * 1.2 / 4, n / 2 + 10 * s, 1.5), (n * 2.8 / 4, n / 2 - 10 * s, -1.5)):
        inside = (x - cx) ** 2 + (y - n / 2) ** 2 + (z - cz) ** 2 < (24 * s) ** 2
        rho[inside] = 2.659
        u[inside, 0] = vx
    return rho.ravel(), u.reshape(-1, 3)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_droplet_collision_demo_32(bfa, orc, storage):
    """The reference demo's settings at 32^3: D3Q19, Shan-Chen (default G), tau
    0.51, velocity_limit 0.32 / velocity_limit_extreme 0.35, droplets launched
    at |u| = 1.5, fully periodic; FP16 as in the demo and FP32.  The first
    tick resets the droplets to equilibrium at the clamped speed.

    psi = 1 - exp(-rho) differs from the oracle's in the last bit now and then
    (device and host exp).  In FP32 that stays a rounding difference (bar
    1e-5 on the populations over 20 ticks, measured 5e-7), though in this
    violent flow it reaches most cells by tick 20.  In FP16 a one-quantum flip
    in a cell near a threshold (the reset, or the nonlinear force) can grow
    to a visible difference within ~20 ticks, so the quantum-level bar is
    applied to the first 5 ticks (measured: tick 1 identical, tick 5 99.999 %
    identical within 1.5e-5) and afterwards only the clamp and the FP64 mass
    are compared."""
    n = 32
    rho, u = droplets(n)
    shape = (n, n, n)
    g = bfa.Solver(shape, 0.51, velocity_set="D3Q19", storage=storage, multiphase=True)
    o = orc.Oracle(shape, 0.51, velocity_set="D3Q19", storage=storage, multiphase=True, precision="f32")
    for s in (g, o):
        s.set_param("velocity_limit", 0.32)
        s.set_param("velocity_limit_extreme", 0.35)
        s.init_fields(rho, u)
    fluid = np.ones(n ** 3, bool)
    _, clamp, reset = classify(g, "D3Q19", fluid, 0.32, 0.35)
    say(f"  before tick 0 (force not included): clamp band {clamp}, reset {reset}")
    assert reset > 100
    w = orc.velocity_table("D3Q19")[:, 3][:, None]
    for ticks in (1, 4, 15):
        g.step(ticks, save=True)
        o.step(ticks, save=True)
        a, b = g.read("populations_raw").astype(np.float64), o.read("populations_raw")
        d = np.abs(a - b)
        same = float(np.mean(a == b))
        un = np.linalg.norm(g.read("velocity_density")[:, :3].astype(np.float64), axis=1)
        say(f"  tick {g.ticks}: raw max|diff|={d.max():.3e} identical={same:.5f} "
            f"cells at the clamp {int(np.sum(np.abs(un - 0.32) < 1e-6))}")
        assert un.max() <= 0.32 * (1 + 1e-6)
        if storage == "fp32":
            assert d.max() <= 1e-5
            if g.ticks == 1:
                assert same > 0.99
        elif g.ticks <= 5:
            assert same > 0.99
            assert np.all(d <= 2 * 2.0 ** -11 * np.maximum(np.abs(b - w), 2.0 ** -14) + 1e-9)
        else:
            mg, mo = g.conservation()[0], o.conservation()[0]
            say(f"  mass GPU {mg:.9g} oracle {mo:.9g}")
            assert abs(mg - mo) <= 1e-4 * mo
    g.close()
    o.close()
