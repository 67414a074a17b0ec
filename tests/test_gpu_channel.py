"""Differentiable channel on the GPU: forward parity with the oracle, adjoint
gradient against the FP64 CPU adjoint and against central finite differences
of the FP64 forward, checkpoint/recompute invariance, and a
torch.autograd.Function wrapper (test harness only).

Bars: forward J and moments vs the FP32 oracle to 1e-6 relative; GPU gradient
vs the FP64 adjoint and vs FP64 central differences to 1e-4 relative (L2 over
cells, per-cell for the checked cells, and for dJ/du_in)."""
import sys

import numpy as np
import pytest

from helpers import SEED, porosity_field, rel_l2

pytestmark = pytest.mark.gpu

SCALE = 0.02   # porosity = 0.02 x smoothed U[0,1]: the channel stays open


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(bfa, orc, res, vs, tau=0.6, u_in=0.05, scale=SCALE, seed=SEED):
    X, Y, Z = res
    gam = porosity_field(X, Y, Z, seed=seed, scale=scale)
    g = bfa.capi.Channel(res, tau, u_in, velocity_set=vs)
    g.set_porosity(gam)
    o = orc.OracleChannel(res, tau, u_in, velocity_set=vs, precision="f64")
    o.set_porosity(gam)
    return g, o, gam


@pytest.mark.parametrize("vs,res,T", [("D2Q9", (96, 24, 1), 400), ("D3Q19", (40, 14, 8), 200)])
def test_channel_forward_parity(bfa, orc, vs, res, T):
    say(f"case: channel forward {vs} {res} T={T}")
    g, o64, gam = make_pair(bfa, orc, res, vs)
    Jg = g.forward(T)
    o32 = orc.OracleChannel(res, 0.6, 0.05, velocity_set=vs, precision="f32")
    o32.set_porosity(gam)
    J32 = o32.forward(T)
    J64 = o64.forward(T)
    mg, m32, m64 = g.moments(), o32.moments(), o64.moments()
    say(f"  J gpu {Jg:.9e} f32 {J32:.9e} f64 {J64:.9e}; moments vs f32 max {np.abs(mg - m32).max():.2e}")
    assert abs(Jg - J32) <= 1e-6 * abs(J32) + 1e-12
    assert abs(Jg - J64) <= 1e-5 * abs(J64)
    assert np.abs(mg - m32).max() <= 1e-6
    assert rel_l2(mg[:, 0], m64[:, 0]) < 1e-5


@pytest.mark.parametrize("vs,res,T,K", [("D2Q9", (96, 24, 1), 400, 50), ("D3Q19", (32, 12, 6), 150, 40)])
def test_channel_gradient_vs_fp64_adjoint(bfa, orc, vs, res, T, K):
    say(f"case: channel gradient {vs} {res} T={T} K={K}")
    g, o64, gam = make_pair(bfa, orc, res, vs)
    Jg, dg_g, du_g = g.gradient(T, checkpoint_every=K, trajectory_budget_bytes=0)
    J64, dg_o, du_o = o64.gradient(T)
    X, Y, Z = res
    fluid = np.tile(np.r_[0, np.ones(Y - 2), 0].repeat(X), Z) > 0
    e = rel_l2(dg_g[fluid], dg_o[fluid])
    say(f"  J {Jg:.6e}/{J64:.6e}  dgamma relL2 {e:.2e}  du {du_g:.6e}/{du_o:.6e}")
    assert abs(Jg - J64) <= 1e-5 * abs(J64)
    assert e < 1e-4
    assert abs(du_g - du_o) <= 1e-4 * abs(du_o)


def test_checkpointing_is_exact(bfa, orc):
    """All segments but the last two recomputed (the forward pass fills the
    resident slot and the extra slot), two or three resident, all resident:
    identical results.  The recomputations run on a second stream beside the
    adjoint sweep, alternating between the extra trajectory slot and the slot
    of the last resident segment once that segment is swept.  With two
    segments nothing is recomputed."""
    res, T, K = (80, 20, 1), 230, 40
    g, _, _ = make_pair(bfa, orc, res, "D2Q9")
    seg = 9 * ((80 * 20 + 63) // 64 * 64) * 4 * (K - 1)
    out = [g.gradient(T, checkpoint_every=K, trajectory_budget_bytes=b) for b in (0, 2 * seg, 3 * seg, 100 * seg)]
    for J, dg, du in out[1:]:
        assert J == out[0][0]
        np.testing.assert_array_equal(dg, out[0][1])
        assert du == out[0][2]
    two = [g.gradient(70, checkpoint_every=K, trajectory_budget_bytes=b) for b in (0, 100 * seg)]
    assert two[0][0] == two[1][0] and two[0][2] == two[1][2]
    np.testing.assert_array_equal(two[0][1], two[1][1])


def test_full_size_gradient_vs_central_differences(bfa, orc):
    """Workload size: D2Q9 512x128, tau 0.6, u_in 0.05, 2000 ticks, checkpoints
    every 100.  GPU adjoint vs central differences of the FP64 forward."""
    X, Y, T = 512, 128, 2000
    say(f"case: channel D2Q9 {X}x{Y} T={T} K=100")
    g, o64, gam = make_pair(bfa, orc, (X, Y, 1), "D2Q9")
    J, dg, du = g.gradient(T, checkpoint_every=100, trajectory_budget_bytes=0)
    t = g.timing()
    Jf = g.forward(T)
    tf = g.timing()[0]
    say(f"  J {J:.6e}  forward {tf:.2f} ms  gradient {t[2]:.2f} ms (recompute {t[1]:.2f})  ratio {t[2] / tf:.2f}")
    assert abs(J - Jf) <= 1e-7 * abs(J)
    rng = np.random.default_rng(SEED)
    fluid = np.where(np.r_[0, np.ones(Y - 2), 0].repeat(X) > 0)[0]
    big = fluid[np.abs(dg[fluid]) > np.quantile(np.abs(dg[fluid]), 0.5)]
    cells = rng.choice(big, 3, replace=False)
    g64 = gam.astype(np.float64)
    eps = 1e-4 * SCALE
    import ctypes as C
    for c in cells:
        vals = []
        for sgn in (1, -1):
            gp = g64.copy()
            gp[c] += sgn * eps
            o64._lib.och_set_porosity(o64._h, gp.ctypes.data_as(C.POINTER(C.c_double)))
            vals.append(o64.forward(T))
        fd = (vals[0] - vals[1]) / (2 * eps)
        say(f"  cell {c}: adjoint {dg[c]:.6e}  fd {fd:.6e}")
        assert abs(dg[c] - fd) <= 1e-4 * abs(fd)
    o64.set_porosity(gam)
    u0 = float(np.float32(0.05))
    vals = []
    for sgn in (1, -1):
        o64.set_inlet(u0 + sgn * 1e-6)
        vals.append(o64.forward(T))
    fd = (vals[0] - vals[1]) / 2e-6
    say(f"  u_in: adjoint {du:.6e}  fd {fd:.6e}")
    assert abs(du - fd) <= 1e-4 * abs(fd)


def test_torch_autograd_wrapper(bfa, orc):
    """A torch.autograd.Function around bf_channel_gradient: backward returns
    grad_output times the adjoint gradients."""
    import torch

    class ChannelObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, gamma, u_in, channel, steps, k):
            channel.set_porosity(gamma.detach().cpu().numpy())
            channel.set_inlet_velocity(float(u_in))
            J, dg, du = channel.gradient(steps, checkpoint_every=k)
            ctx.save_for_backward(torch.from_numpy(dg), torch.tensor(du, dtype=torch.float64))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            dg, du = ctx.saved_tensors
            return grad_out.to(torch.float32) * dg, (grad_out * du).to(torch.float32), None, None, None

    res, T = (64, 18, 1), 300
    g, o64, gam = make_pair(bfa, orc, res, "D2Q9")
    gamma = torch.tensor(gam, requires_grad=True)
    u_in = torch.tensor(0.05, requires_grad=True)
    J = ChannelObjective.apply(gamma, u_in, g, T, 60)
    (2.0 * J).backward()
    J64, dg_o, du_o = o64.gradient(T)
    fluid = np.tile(np.r_[0, np.ones(16), 0].repeat(64), 1) > 0
    assert rel_l2(gamma.grad.numpy()[fluid] / 2.0, dg_o[fluid]) < 1e-4
    assert abs(u_in.grad.item() / 2.0 - du_o) <= 1e-4 * abs(du_o)
