"""Pins of the differentiable-channel oracle (oracle/bf_oracle_channel.c):
boundary closures reproduce the imposed moments, the forward model carries the
inlet flux to the outlet, and the hand-derived adjoint equals central finite
differences and satisfies the per-tick dot-product identity.  CPU only."""
import numpy as np
import pytest

from helpers import SEED, porosity_field


def near_rest_state(q, n, table, seed):
    rng = np.random.default_rng(seed)
    w = table[:, 3][:, None]
    return (w * (1.0 + 0.05 * (rng.random((q, n)) - 0.5))).ravel()


@pytest.mark.parametrize("vs,res", [("D2Q9", (10, 8, 1)), ("D3Q19", (6, 7, 4))])
def test_closures_impose_moments(orc, vs, res):
    u_in = 0.05
    ch = orc.OracleChannel(res, 0.6, u_in, velocity_set=vs)
    t = orc.velocity_table(vs)
    c = t[:, :3]
    X, Y, Z = res
    st = near_rest_state(ch.q, ch.cells, t, SEED)
    prof = ch.profile()
    for y in range(1, Y - 1):
        for z in range(Z):
            f = ch.precollision(st, 0, y, z)
            rho = f.sum()
            m = f @ c
            uin = np.float32(u_in) * prof[y]
            assert abs(m[0] - rho * uin) < 1e-13
            assert abs(m[1]) < 1e-13 and abs(m[2]) < 1e-13
            f = ch.precollision(st, X - 1, y, z)
            assert abs(f.sum() - 1.0) < 1e-13
            m = f @ c
            assert abs(m[1]) < 1e-13 and abs(m[2]) < 1e-13


def test_profile_has_unit_mean_between_halfway_walls(orc):
    ch = orc.OracleChannel((4, 130, 1), 0.6, 0.05)
    p = ch.profile()
    assert p[0] == 0 and p[-1] == 0
    # 6 s (1 - s), s = (y - 1/2)/H, has unit mean over the continuous channel;
    # sampled at the H cell centres its mean is 1 + 1/(2 H^2)
    H = 128
    assert abs(p[1:-1].mean() - (1 + 0.5 / H ** 2)) < 1e-12


def test_open_channel_carries_inlet_flux(orc):
    """gamma = 0, steady state: the mass flux rho u_x is the same in the
    columns of the upstream half (mass conservation); the
    outlet-cell velocity J (a boundary-cell moment) is within 2 % of it."""
    X, Y = 40, 14
    ch = orc.OracleChannel((X, Y, 1), 0.8, 0.04)
    ch.set_porosity(np.zeros(X * Y))
    J = ch.forward(4000)
    mom = ch.moments().reshape(Y, X, 4)
    flux = (mom[1:-1, :, 0] * mom[1:-1, :, 3]).mean(0)
    # (near the pressure outlet an odd-even flux oscillation of ~0.5 % remains)
    assert np.all(np.abs(flux[: X // 2] - flux[0]) < 1e-4 * flux[0])
    assert abs(J - flux[0]) < 2e-2 * flux[0]
    assert 0.03 < J < 0.05


@pytest.mark.parametrize("vs,res,T", [("D2Q9", (12, 8, 1), 40), ("D3Q19", (7, 6, 4), 20)])
def test_adjoint_matches_central_differences(orc, vs, res, T):
    X, Y, Z = res
    gam = porosity_field(X, Y, Z, seed=SEED + 11, scale=0.6)
    ch = orc.OracleChannel(res, 0.6, 0.05, velocity_set=vs)
    ch.set_porosity(gam)
    J, dg, du = ch.gradient(T)
    assert abs(J - ch.forward(T)) < 1e-15
    rng = np.random.default_rng(SEED)
    fluid = np.where(np.tile(np.r_[0, np.ones(Y - 2), 0].repeat(X), Z) > 0)[0]
    eps = 1e-6
    for c in rng.choice(fluid, 8, replace=False):
        gp, gm = gam.astype(np.float64).copy(), gam.astype(np.float64).copy()
        gp[c] += eps
        gm[c] -= eps
        ch._lib.och_set_porosity(ch._h, gp.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
        Jp = ch.forward(T)
        ch._lib.och_set_porosity(ch._h, gm.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
        Jm = ch.forward(T)
        fd = (Jp - Jm) / (2 * eps)
        assert abs(fd - dg[c]) <= 1e-6 * max(abs(fd), 1e-6), (c, fd, dg[c])
    ch.set_porosity(gam)
    u0 = float(np.float32(0.05))
    ch.set_inlet(u0 + eps)
    Jp = ch.forward(T)
    ch.set_inlet(u0 - eps)
    Jm = ch.forward(T)
    fd = (Jp - Jm) / (2 * eps)
    assert abs(fd - du) <= 1e-6 * abs(fd), (fd, du)


@pytest.mark.parametrize("vs,res", [("D2Q9", (9, 7, 1)), ("D3Q19", (5, 6, 3))])
def test_tick_dot_product_identity(orc, vs, res):
    """<tick'(s) v, nu> == <v, tick'(s)^T nu> with the linearisation taken by
    central differences of the FP64 tick."""
    X, Y, Z = res
    ch = orc.OracleChannel(res, 0.7, 0.05, velocity_set=vs)
    ch.set_porosity(porosity_field(X, Y, Z, seed=SEED + 3, scale=0.5))
    t = orc.velocity_table(vs)
    s = near_rest_state(ch.q, ch.cells, t, SEED + 1)
    rng = np.random.default_rng(SEED + 2)
    v = rng.standard_normal(s.size)
    nu = rng.standard_normal(s.size)
    eps = 1e-7
    op, _ = ch.tick_pair(s + eps * v, nu)
    om, _ = ch.tick_pair(s - eps * v, nu)
    _, jt_nu = ch.tick_pair(s, nu)
    lhs = ((op - om) / (2 * eps)) @ nu
    rhs = v @ jt_nu
    assert abs(lhs - rhs) < 1e-6 * abs(rhs)
