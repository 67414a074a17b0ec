"""Chunked tick pairs (BF_PARAM_TICK_PAIR_PLANES): an even tick and the odd tick
after it run z-chunk by z-chunk, chunk j's even tick in one launch with chunk
j-2's odd tick.  Every cell sees the same operations as with whole-grid ticks,
so the results must be bit-identical to the whole-grid ticks and to the oracle
(sweeps of two pairs, BF_PARAM_TICK_SWEEP_TICKS, likewise)
(FP32), or identical to the oracle's emulated 16-bit storage up to one-quantum
rounding flips, whatever the chunk size: chunks of 1, 2 and 3 planes, a ragged
last chunk, both parities at the start and end of a bf_step call, saves and
FP64 sums on the last tick, walls, open z faces and the thermal lattice."""
import sys

import numpy as np
import pytest

from helpers import SEED, channel_ids, random_state, spheres_ids

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(bfa, orc, res, vs, storage, planes, *, periodic=(1, 1, 1), walls=False, thermal=False, seed=SEED + 51,
              sweep=0):
    X, Y, Z = res
    n = X * Y * Z
    rho, u = random_state(n, seed=seed, du=0.06, drho=0.04)
    init = dict(rho=rho, u=u)
    objects, params = [], []
    if walls:
        rng = np.random.default_rng(seed)
        ids = np.where(rng.random(n) < 0.1, rng.integers(1, 3, n), 0).astype(np.uint8)
        init["boundary_id"] = ids
        objects.append((2, dict(velocity=(0.02, 0.0, -0.01))))
    if thermal:
        ids = channel_ids(X, Y, Z)
        ids.reshape(Z, Y, X)[:, -1, :] = 2
        init["boundary_id"] = ids
        init["temperature"] = (0.5 + 0.02 * np.random.default_rng(seed).random(n)).astype(np.float32)
        init["force"] = np.tile(np.array([0, -1e-4, 0], np.float32), (n, 1))
        objects = [(1, dict(temperature=1.0)), (2, dict(temperature=0.0))]
        params = [("reference_temperature", 0.5), ("thermal_relaxation_time", 0.9)]
    g = bfa.Solver(res, 0.6, periodic=periodic, velocity_set=vs, storage=storage)
    o = orc.Oracle(res, 0.6, periodic=periodic, velocity_set=vs, storage=storage, precision="f32")
    g.set_param("tick_pair_planes", planes)
    g.set_param("tick_sweep_ticks", sweep)
    for oid, kw in objects:
        g.set_object(oid, **kw)
        o.set_object(oid, **kw)
    for k, v in params:
        g.set_param(k, v)
        o.set_param(k, v)
    g.init_fields(**init)
    o.init_fields(**init)
    return g, o


def check(g, o, storage, vs, orc, thermal=False):
    a, b = g.read("populations_raw").astype(np.float64), o.read("populations_raw")
    same = float(np.mean(a == b))
    say(f"  raw max|diff|={np.abs(a - b).max():.3e} identical={same:.5f}")
    if storage == "fp32":
        assert np.array_equal(a, b)
        assert np.array_equal(g.read("velocity_density").astype(np.float64), o.read("velocity_density"))
    else:
        q = 2.0 ** -11 if storage == "fp16" else 2.0 ** -8
        w = orc.velocity_table(vs)[:, 3][:, None]
        assert same > 0.98
        assert np.all(np.abs(a - b) <= 2 * q * np.maximum(np.abs(b - w), 2.0 ** -14) + 1e-9)
    if thermal:
        assert np.array_equal(g.read("thermal_raw").astype(np.float64), o.read("thermal_raw"))


CASES = [("D3Q19", "fp32", 1, (12, 10, 8)), ("D3Q19", "fp32", 3, (12, 10, 13)), ("D3Q19", "fp16", 2, (16, 6, 9)),
         ("D3Q19", "bf16", 2, (16, 6, 9)), ("D3Q15", "fp32", 2, (10, 8, 8)), ("D3Q27", "fp32", 2, (10, 8, 11)),
         ("D3Q27", "fp16", 3, (10, 8, 12)), ("D3Q19", "fp32", 2, (64, 5, 9)), ("D3Q19", "fp16", 2, (128, 3, 8))]


@pytest.mark.parametrize("vs,storage,planes,res", CASES)
def test_pairs_periodic_match_oracle(bfa, orc, vs, storage, planes, res):
    """Fully periodic grids (the plain kernels, with and without the row
    buffer): 2 + 3 + 4 ticks in three calls, so pairs start on both parities
    and calls end on both; saved fields on the last tick of each call."""
    say(f"case: pairs {vs} {storage} {res} planes={planes}")
    g, o = make_pair(bfa, orc, res, vs, storage, planes)
    for n in (2, 3, 4):
        g.step(n, save=True)
        o.step(n, save=True)
        check(g, o, storage, vs, orc)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("periodic", [(1, 0, 1), (1, 0, 0), (0, 1, 0)])
def test_pairs_walls_and_open_faces_match_oracle(bfa, orc, storage, periodic):
    """Random solid cells with a moving wall, open y or z faces (0.01 reads
    and dropped writes at z = 0 and z = Z-1 cross chunk boundaries)."""
    res = (14, 9, 11)
    say(f"case: pairs walls {storage} periodic={periodic}")
    g, o = make_pair(bfa, orc, res, "D3Q19", storage, 2, periodic=periodic, walls=True)
    for n in (4, 5):
        g.step(n, save=True)
        o.step(n, save=True)
        check(g, o, storage, "D3Q19", orc)


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("periodic", [(1, 1, 0), (0, 1, 1), (1, 0, 1)])
def test_pairs_open_faces_without_walls_match_oracle(bfa, orc, storage, periodic):
    """No solid cells, one open axis (0.01 reads and dropped writes at its
    faces; on z they cross chunk boundaries): the plain kernels with
    neighbour checks, which a lattice of this kind above 384 MB pairs by
    default."""
    res = (13, 9, 11)
    say(f"case: pairs open faces {storage} periodic={periodic}")
    g, o = make_pair(bfa, orc, res, "D3Q19", storage, 3, periodic=periodic)
    for n in (4, 3):
        g.step(n, save=True)
        o.step(n, save=True)
        check(g, o, storage, "D3Q19", orc)


@pytest.mark.parametrize("vs,storage", [("D3Q19", "fp32"), ("D3Q27", "fp32"), ("D3Q19", "fp16")])
def test_pairs_thermal_match_oracle(bfa, orc, vs, storage):
    """The thermal lattice (D3Q7, FP32, in place like the main lattice) in
    pairs: hot and cold walls with buoyancy; the main lattice in FP32 or FP16
    (a 16-bit lattice with walls above 384 MB pairs by default)."""
    say(f"case: pairs thermal {vs} {storage}")
    g, o = make_pair(bfa, orc, (12, 10, 10), vs, storage, 2, periodic=(1, 0, 1), thermal=True)
    assert g.info()["thermal"] == 1
    for n in (6, 3):
        g.step(n, save=True)
        o.step(n, save=True)
        check(g, o, storage, vs, orc, thermal=True)


SWEEP_CASES = [("D3Q19", "fp32", 1, (12, 10, 8), {}), ("D3Q19", "fp32", 3, (12, 10, 13), {}),
               ("D3Q19", "fp16", 2, (16, 6, 9), {}), ("D3Q19", "bf16", 2, (16, 6, 9), {}),
               ("D3Q15", "fp32", 2, (10, 8, 8), {}), ("D3Q27", "fp16", 3, (10, 8, 12), {}),
               ("D3Q19", "fp32", 2, (64, 5, 9), {}),
               ("D3Q19", "fp32", 2, (14, 9, 11), dict(periodic=(1, 0, 0), walls=True)),
               ("D3Q19", "fp16", 2, (14, 9, 11), dict(periodic=(0, 1, 0), walls=True)),
               ("D3Q19", "bf16", 3, (13, 9, 11), dict(periodic=(1, 1, 0))),
               ("D3Q19", "fp32", 2, (12, 10, 10), dict(periodic=(1, 0, 1), thermal=True)),
               ("D3Q27", "fp16", 2, (12, 10, 10), dict(periodic=(1, 0, 1), thermal=True))]


@pytest.mark.parametrize("vs,storage,planes,res,kw", SWEEP_CASES)
def test_two_pair_sweeps_match_oracle(bfa, orc, vs, storage, planes, res, kw):
    """Sweeps of two tick pairs (BF_PARAM_TICK_SWEEP_TICKS 4): each launch
    also runs the next pair's even and odd ticks two planes behind the tick
    before each.  Calls of 5, 8, 3 and 6 ticks (sweeps start on both
    parities; calls that save end with a pair or a single tick, one without
    saves ends with a sweep), against the oracle."""
    say(f"case: two-pair sweeps {vs} {storage} {res} planes={planes} {kw}")
    g, o = make_pair(bfa, orc, res, vs, storage, planes, sweep=4, **kw)
    for n, save in ((5, True), (8, False), (3, True), (6, True)):
        g.step(n, save=save)
        o.step(n, save=save)
    if storage == "fp32":
        g.step(1, save=True)
        o.step(1, save=True)
    check(g, o, storage, vs, orc, thermal=kw.get("thermal", False))


SC_CASES = [("D3Q27", "fp32", 3, (16, 12, 14), "thermal", 2), ("D3Q27", "fp32", 2, (16, 12, 14), "thermal", 4),
            ("D3Q19", "fp32", 2, (14, 10, 12), None, 2), ("D3Q19", "fp32", 1, (14, 10, 12), None, 4),
            ("D3Q19", "fp16", 3, (14, 10, 12), None, 4), ("D3Q19", "bf16", 2, (14, 10, 12), "walls", 2),
            ("D3Q15", "fp32", 2, (12, 10, 10), "walls", 4)]


@pytest.mark.parametrize("vs,storage,planes,res,kind,sweep", SC_CASES)
def test_shan_chen_pairs(bfa, orc, vs, storage, planes, res, kind, sweep):
    """The lagged Shan-Chen kernels in chunked pairs and sweeps of two pairs
    (tick t reads the psi that tick t - 1 wrote, one plane around, like the
    populations): a dense block in a vapour, G = -5, with hot and cold walls
    and the D3Q7 thermal lattice, walls of effective density 1.5, or fully
    periodic.  Calls of 5, 8 and 3 ticks: bit-identical to whole-grid ticks
    (populations, saved fields, psi); FP32 also against the oracle, to the
    last bit of exp (1e-6 on the populations, 2e-6 on psi)."""
    X, Y, Z = res
    N = X * Y * Z
    rng = np.random.default_rng(SEED + 57)
    rho = (0.3 + 0.05 * rng.random(N)).astype(np.float32)
    rho.reshape(Z, Y, X)[2:Z - 3, 3:Y - 3, 4:X - 5] = 2.2
    init = dict(rho=rho, u=np.zeros((N, 3), np.float32))
    objects, params, periodic = [], [("interaction_strength", -5.0)], (1, 1, 1)
    if kind == "thermal":
        ids = np.zeros((Z, Y, X), np.uint8)
        ids[:, 0, :] = 1
        ids[:, -1, :] = 2
        init["boundary_id"] = ids.ravel()
        init["temperature"] = (0.5 + 0.02 * rng.random(N)).astype(np.float32)
        init["force"] = np.tile(np.array([0, -1e-4, 0], np.float32), (N, 1))
        objects = [(1, dict(temperature=1.0)), (2, dict(temperature=0.0))]
        params += [("reference_temperature", 0.5)]
        periodic = (1, 0, 1)
    elif kind == "walls":
        init["boundary_id"] = channel_ids(X, Y, Z)
        objects, periodic = [(1, dict(effective_density=1.5))], (1, 0, 1)
    say(f"case: Shan-Chen pairs {vs} {storage} {res} planes={planes} {kind} sweep={sweep}")

    def build(solver):
        for oid, kw in objects:
            solver.set_object(oid, **kw)
        for k, v in params:
            solver.set_param(k, v)
        solver.init_fields(**init)
        return solver

    out = []
    for p in (-1, planes):
        g = bfa.Solver(res, 0.9, periodic=periodic, velocity_set=vs, storage=storage, multiphase=True)
        g.set_param("tick_pair_planes", p)
        g.set_param("tick_sweep_ticks", sweep)
        build(g)
        for n, save in ((5, True), (8, False), (3, True)):
            g.step(n, save=save)
        out.append([g.read("populations_raw"), g.read("velocity_density"), g.read("psi")])
        if kind == "thermal":
            out[-1].append(g.read("thermal_raw"))
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    if storage == "fp32":
        o = build(orc.Oracle(res, 0.9, periodic=periodic, velocity_set=vs, storage=storage, precision="f32",
                             multiphase=True))
        for n, save in ((5, True), (8, False), (3, True)):
            o.step(n, save=save)
        d = float(np.abs(out[1][0].astype(np.float64) - o.read("populations_raw")).max())
        dpsi = float(np.abs(out[1][2] - o.read("psi")).max())
        say(f"  against the oracle: raw max|diff|={d:.3e} psi max|diff|={dpsi:.3e}")
        assert d <= 1e-6 and dpsi <= 2e-6


def test_pairs_conservation_sums(bfa, orc):
    """FP64 sums recorded on the last tick of a pair: each odd chunk writes
    the partials of its own planes."""
    g, o = make_pair(bfa, orc, (12, 10, 9), "D3Q19", "fp32", 2, periodic=(1, 0, 1), walls=True)
    g.step(6, conservation=True)
    o.step(6, conservation=True)
    np.testing.assert_allclose(g.last_step_sums()[:4], o.last_step_sums()[:4], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_pairs_equal_whole_grid_ticks_porous_96(bfa, storage):
    """A 96^3 porous duct (walls, constant force) run with chunks of 4 planes
    and with whole-grid ticks: bit-identical after 10 ticks."""
    n = 96
    cells = n ** 3
    ids = spheres_ids(n, n, n, 0.05, 4, 8, seed=SEED + 53)
    out = []
    for planes in (-1, 4):
        s = bfa.Solver((n, n, n), 0.56, periodic=(1, 0, 0), velocity_set="D3Q19", storage=storage)
        s.set_param("tick_pair_planes", planes)
        s.init_fields(np.ones(cells, np.float32), np.zeros((cells, 3), np.float32),
                      force=np.tile(np.array([1e-5, 0, 0], np.float32), (cells, 1)), boundary_id=ids)
        s.step(10, save=True)
        out.append((s.read("populations_raw"), s.read("velocity_density")))
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_automatic_mode_on_the_periodic_tgv_256(bfa, storage):
    """The automatic mode pairs the ticks of the 256^3 box (1.3 GB in FP32,
    0.64 GB in FP16; no walls; FP16 in sweeps of two pairs), bit-identical to
    whole-grid ticks (6 ticks, saved fields on the last: FP16 runs one sweep
    of two pairs, then a pair)."""
    from boltzmannfluids_amd.scenarios import taylor_green
    n = 256
    say(f"case: automatic tick pairs TGV {n}^3 {storage}")
    rho, u = taylor_green(n)
    out = []
    for planes in (0, -1):
        s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=storage)
        s.set_param("tick_pair_planes", planes)
        s.init_fields(rho, u)
        s.step(6, save=True)
        out.append((s.read("populations_raw"), s.read("velocity_density")))
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_pairs_parameter(bfa):
    s = bfa.Solver((8, 8, 8), 0.6)
    assert s.get_param("tick_pair_planes") == 0
    s.set_param("tick_pair_planes", 3)
    assert s.get_param("tick_pair_planes") == 3
    for bad in (-2, 1.5):
        with pytest.raises(bfa.BFError):
            s.set_param("tick_pair_planes", bad)
    assert s.get_param("tick_sweep_ticks") == 0
    s.set_param("tick_sweep_ticks", 4)
    assert s.get_param("tick_sweep_ticks") == 4
    for bad in (1, 3, -1, 8):
        with pytest.raises(bfa.BFError):
            s.set_param("tick_sweep_ticks", bad)
