"""The oracle's stability branches, pinned by closed forms (CPU only).

collide.comp:704-705 floors the density at 0.01; :722-725 clamps the velocity
(after the half-force shift) to |u| = velocity_limit along u; :727-745 replaces
the BGK update by the equilibrium f_eq(rho, clamped u) when the unclamped |u|
exceeds velocity_limit_extreme.  The macroscopic save (:784-792) stores the
clamped velocity and the floored density.

Set-up: a single cell on a fully periodic 1 x 1 x 1 grid.  On the even tick it
reads slot i at x - c_i = x and writes the post-collision value of direction i
to slot opp(i) at x + c_i = x, so one tick maps the raw slots f to
f*_i = raw[opp(i)].  The input is the first-order state f_i = w_i (rho + 3 c_i.m),
whose moments are exactly rho and m (sum w = 1, sum w c c = I/3).

What the tests check, with no force (tau = 0.6 unless stated):
  - mass: sum f* = rho, or rho_raw (1 - 1/tau) + 0.01 / tau below the floor;
  - momentum: sum f* c = m (1 - 1/tau) + rho u_c / tau in the clamp band, and
    exactly rho u_c past the reset threshold (no memory of m);
  - second moment past the reset: sum f* c c = rho (I/3 + u_c u_c) (the
    second-order equilibrium's closed form on these lattices);
  - the saved velocity is u_c, |u_c| = velocity_limit, parallel to m;
  - below the floor the velocity is m / 0.01, not m / rho_raw.
Each check is also run on an oracle given the wrong thresholds (the two
limits swapped, or no reset: the same as testing the reset after the clamp),
where it must fail.
"""
import numpy as np
import pytest

SETS = ["D2Q9", "D3Q15", "D3Q19", "D3Q27"]
TOL = {"f64": 1e-12, "f32": 2e-6}


def single_cell_tick(orc, vs, precision, f_raw, tau=0.6, ulim=0.25, uext=0.30):
    """One even tick of a single periodic cell from raw slots f_raw; returns
    (post-collision populations f*_i, saved (u, rho))."""
    o = orc.Oracle((1, 1, 1), tau, velocity_set=vs, precision=precision)
    o.set_param("velocity_limit", ulim)
    o.set_param("velocity_limit_extreme", uext)
    o.init_fields(np.ones(1, np.float32), np.zeros((1, 3), np.float32))
    o.write_populations(f_raw)
    o.step(1, save=True)
    raw = o.read("populations_raw")[:, 0]
    q = len(raw)
    post = np.array([raw[orc_opp(i)] for i in range(q)])
    vd = o.read("velocity_density")[0]
    o.close()
    return post, vd


def orc_opp(i):
    return 0 if i == 0 else (i + 1 if i & 1 else i - 1)


def first_order_state(tab, rho, m):
    """f_i = w_i (rho + 3 c_i . m): moments rho and m exactly."""
    c, w = tab[:, :3], tab[:, 3]
    return w * (rho + 3.0 * (c @ m))


def moments(tab, f):
    c = tab[:, :3]
    return f.sum(), c.T @ f, np.einsum("i,ia,ib->ab", f, c, c)


def direction(vs, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(3)
    if vs == "D2Q9":
        d[2] = 0.0
    return d / np.linalg.norm(d)


def limit_failures(orc, vs, precision, speed, rho=1.1, tau=0.6, ulim=0.25, uext=0.30, expect=(0.25, 0.30)):
    """Run one cell at unclamped speed `speed` (no force) and return the list of
    closed-form checks that fail, judged against the reference's thresholds
    `expect` = (velocity_limit, velocity_limit_extreme).  The oracle runs with
    (ulim, uext), which a planted-mistake test sets wrong."""
    tab = orc.velocity_table(vs)
    d = direction(vs, int(speed * 1000) + len(tab))
    m = rho * speed * d
    f = first_order_state(tab, rho, m)
    post, vd = single_cell_tick(orc, vs, precision, f, tau=tau, ulim=ulim, uext=uext)
    M0, M1, M2 = moments(tab, post)
    tau = float(np.float32(tau))      # the solver's parameters are FP32
    el, ee = expect
    uc = (el * d) if speed > el else speed * d
    tol = TOL[precision]
    fails = []
    if abs(M0 - rho) > tol * rho:
        fails.append("mass")
    if speed > ee:
        if np.abs(M1 - rho * uc).max() > tol:
            fails.append("reset momentum")
        pi = rho * (np.eye(3) / 3.0 + np.outer(uc, uc))
        if vs == "D2Q9":
            pi[2, 2] = 0.0
        if np.abs(M2 - pi).max() > tol:
            fails.append("reset second moment")
    else:
        if np.abs(M1 - (m * (1 - 1 / tau) + rho * uc / tau)).max() > tol:
            fails.append("BGK momentum with the clamped velocity")
    # the saved fields are FP32 (RGBA32F texels)
    if np.abs(vd[:3] - uc).max() > max(tol, 1e-7) or abs(vd[3] - rho) > max(tol, 2e-7):
        fails.append("saved velocity")
    return fails


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("vs", SETS)
@pytest.mark.parametrize("speed", [0.1, 0.27, 0.29, 0.31, 0.5, 1.5])
def test_clamp_and_reset_closed_forms(orc, vs, precision, speed):
    assert limit_failures(orc, vs, precision, speed) == []


@pytest.mark.parametrize("vs", SETS)
def test_clamped_saved_velocity_has_the_limit_norm(orc, vs):
    """In the clamp band the saved |u| is velocity_limit to rounding, along m."""
    tab = orc.velocity_table(vs)
    d = direction(vs, 7)
    f = first_order_state(tab, 1.0, 0.28 * d)
    _, vd = single_cell_tick(orc, vs, "f32", f)
    u = vd[:3].astype(np.float64)
    assert abs(np.linalg.norm(u) - 0.25) < 1e-7
    assert np.abs(u / np.linalg.norm(u) - d).max() < 1e-6


@pytest.mark.parametrize("vs", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("speed", [0.27, 0.4])
def test_swapped_limits_are_caught(orc, vs, speed):
    """Planted mistake (in the oracle's parameters): velocity_limit and
    velocity_limit_extreme swapped.  A cell at 0.27 is then reset instead of
    clamped, a cell at 0.4 reset to the wrong speed; the closed-form checks
    against the reference's thresholds must fail."""
    assert limit_failures(orc, vs, "f64", speed, ulim=0.30, uext=0.25) != []


@pytest.mark.parametrize("vs", ["D2Q9", "D3Q19", "D3Q27"])
def test_reset_after_clamp_is_caught(orc, vs):
    """Planted mistake: the reset test made on the clamped velocity (never
    above 0.25 < 0.30, the same as no reset).  A cell at 0.4 then relaxes by
    BGK and keeps part of its momentum; the reset checks must fail."""
    fails = limit_failures(orc, vs, "f64", 0.4, uext=1e9)
    assert "reset momentum" in fails and "reset second moment" in fails


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("vs", SETS)
def test_density_floor(orc, vs, precision):
    """Sum f = 0.005 < 0.01: the collision uses rho = 0.01 (collide.comp:705).
    With tau = 0.6 the mass becomes 0.005 (1 - 1/tau) + 0.01 / tau, the
    momentum m (1 - 1/tau) + m / tau = m (u = m / 0.01 relaxes to itself), and
    the saved density and velocity are 0.01 and m / 0.01."""
    tab = orc.velocity_table(vs)
    tau, r = 0.6, 0.005
    m = 0.0004 * direction(vs, 3)
    post, vd = single_cell_tick(orc, vs, precision, first_order_state(tab, r, m), tau=tau)
    M0, M1, _ = moments(tab, post)
    tol = TOL[precision]
    tau = float(np.float32(tau))
    assert abs(M0 - (r * (1 - 1 / tau) + 0.01 / tau)) < tol
    assert np.abs(M1 - m).max() < tol
    assert abs(vd[3] - 0.01) < max(tol, 1e-9)             # saved as FP32
    assert np.abs(vd[:3] - m / 0.01).max() < max(10 * tol, 1e-8)


@pytest.mark.parametrize("vs", ["D2Q9", "D3Q19"])
def test_density_floor_with_tau_one_gives_the_floor_equilibrium(orc, vs):
    """tau = 1 and m = 0 below the floor: f* = f_eq(0.01, 0) = 0.01 w_i."""
    tab = orc.velocity_table(vs)
    post, vd = single_cell_tick(orc, vs, "f64", first_order_state(tab, 0.004, np.zeros(3)), tau=1.0)
    np.testing.assert_allclose(post, 0.01 * tab[:, 3], rtol=0, atol=1e-15)
    assert vd[3] == pytest.approx(0.01, abs=1e-9)


def test_no_floor_is_caught(orc):
    """Planted mistake: a floor that only changes the saved density (the
    collision keeps rho_raw).  Emulated on the oracle's output by comparing
    with the unfloored expectations, which the real oracle must miss."""
    tab = orc.velocity_table("D3Q19")
    tau, r = 0.6, 0.005
    m = 0.0004 * direction("D3Q19", 3)
    post, vd = single_cell_tick(orc, "D3Q19", "f64", first_order_state(tab, r, m), tau=tau)
    M0, _, _ = moments(tab, post)
    assert abs(M0 - r) > 1e-3                 # the unfloored collision would keep the mass
    assert np.abs(vd[:3] - m / r).max() > 1e-2  # and save m / rho_raw
