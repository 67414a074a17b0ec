"""Benchmark: lattice Boltzmann ticks on one MI355X.

Headline workload: 256^3 fully periodic D3Q19 box, Taylor-Green vortex
(u0 = 0.05, rho = 1 + 3 p'), tau = 0.55, FP32 storage.  One step = one lattice
tick over the whole grid; the populations stay resident in HBM.

Prints one JSON line:
  value / roofline        MLUPs/s of the FP32 run; HBM roofline of the tick kernel
                          (algorithmic 152 B/cell / device-timed launch duration)
  ceiling                 copy and in-place 19-slot streaming probes (achievable GB/s)
  fp16_shifted, bf16_...  the same run with 16-bit shifted storage
  error_vs_fp32           16-bit fields vs the FP32 run (same workload at 128^3)
  error_vs_fp64           every storage type vs the FP64 CPU oracle (TGV and porous duct
                          at 64^3, 1000 ticks; Shan-Chen + thermal FP32 at 32^3, 500
                          ticks; the D2Q9 channel gradient against the FP64 adjoint;
                          computed in the CPU leg)
  other_configs           porous 384^3 D3Q19 (FP32/FP16/BF16, + error at 128^3),
                          Shan-Chen D3Q27 + thermal D3Q7 192^3, D2Q9 cavity 256^2
  adjoint_over_forward    differentiable channel: gradient time / forward time
  cpu_baseline            the FP32 CPU restatement (oracle) on the host cores
  wall_clock_frac         the headline's HBM fraction from host wall clock (roofline.frac
                          uses the kernels' device time)

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from boltzmannfluids_amd import probe  # noqa: E402  (bench-only bandwidth probes, libbfprobe.so)
from boltzmannfluids_amd.scenarios import (lid_cavity, porosity, porous_spheres,  # noqa: E402
                                           spinodal, taylor_green)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E published peak
BYTES = {"fp32": 4, "fp16": 2, "bf16": 2}


def bytes_per_cell(q, storage, extra=0):
    """Algorithmic bytes per cell update: Q loads + Q stores (+ flags/aux)."""
    return 2 * q * BYTES[storage] + extra


def load_traffic(name):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration) if a summary for this workload is committed
    (profiles/<name>_traffic.json)."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def roofline(bpc, cells, kernel_ms, traffic=None):
    achieved = bpc * cells / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}


# Warm-up covers at least this much device time.  After the GPU idles (the
# host set-up of each workload), the kernels with the most instructions per
# byte (the 16-bit even tick) run a power-management transient: fast for ~1 ms,
# up to 1.6x slower ~4 ms in, settled after ~60 ms; the memory-bound odd tick
# stays within 15 %.  The timed region starts in the steady state.
WARMUP_MIN_S = 0.15


def timed_run(s, steps, warmup, conservation_every=0):
    """Warm up (`warmup` ticks and at least WARMUP_MIN_S of running), then time
    `steps` ticks with host wall clock (bracketed by device synchronisation) and
    HIP events on the solver's stream."""
    import torch
    t0 = time.perf_counter()
    s.step(warmup)
    s.synchronize()
    while time.perf_counter() - t0 < WARMUP_MIN_S:
        s.step(max(2, warmup) // 2 * 2)   # whole tick pairs: the timed region starts on the same parity
        s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if conservation_every:
        ev = 0.0
        for k in range(0, steps, conservation_every):
            ev += s.step_timed(min(conservation_every, steps - k), conservation=True)
    else:
        ev = s.step_timed(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, ev


def run_tgv(bfa, n, storage, steps, warmup):
    rho, u = taylor_green(n)
    s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=storage)
    s.init_fields(rho, u)
    wall, ev = timed_run(s, steps, warmup)
    s.close()
    cells = n ** 3
    kernel_ms = ev / steps
    return {"mlups": cells * steps / wall / 1e6, "ms_per_step": wall * 1e3 / steps, "kernel_ms": kernel_ms,
            "roofline": roofline(bytes_per_cell(19, storage), cells, kernel_ms,
                                 load_traffic(f"step_d3q19_{storage}_{n}"))}


def field_error(a, b):
    er = float(np.linalg.norm(a[:, 3] - b[:, 3]) / np.linalg.norm(b[:, 3]))
    eu = float(np.linalg.norm(a[:, :3] - b[:, :3]) / np.linalg.norm(b[:, :3]))
    return er, eu


def field_errors(a, b):
    """Norm-relative and largest absolute differences of the (u, rho) fields."""
    er, eu = field_error(a, b)
    return {"rho": er, "u": eu, "rho_max_abs": float(np.abs(a[:, 3] - b[:, 3]).max()),
            "u_max_abs": float(np.abs(a[:, :3] - b[:, :3]).max())}


def run_fields(bfa, shape, tau, storage, ticks, setup):
    """Run `ticks` ticks of a workload built by setup(solver) and return the
    (N, 4) velocity/density field."""
    g = bfa.Solver(shape[0], tau, storage=storage, **shape[1])
    setup(g)
    g.step(ticks, save=True)
    vd = g.read("velocity_density").astype(np.float64)
    g.close()
    return vd


def accuracy_vs_fp32(bfa, storages, n=128, ticks=1000):
    """||d rho||/||rho|| and ||d u||/||u|| of each 16-bit run against the FP32
    GPU run of the same TGV workload at n^3 (FP32 itself is held to the FP64
    CPU restatement by the parity tests)."""
    rho, u = taylor_green(n)
    shape = ((n, n, n), {"velocity_set": "D3Q19"})
    ref = run_fields(bfa, shape, 0.55, "fp32", ticks, lambda g: g.init_fields(rho, u))
    out = {"workload": f"TGV {n}^3 tau=0.55 u0=0.05, {ticks} ticks, vs the FP32 run"}
    for st in storages:
        if st != "fp32":
            er, eu = field_error(run_fields(bfa, shape, 0.55, st, ticks, lambda g: g.init_fields(rho, u)), ref)
            out[st] = {"rho": er, "u": eu}
    return out


def porous_config(bfa, n=384, steps=100, warmup=10, err_n=128, err_ticks=1000):
    """D3Q19 porous duct: periodic x, walls on y/z, 5 % solid spheres r in [4, 12],
    g_x = 1e-6, tau = 0.56; FP32/FP16/BF16 throughput at n^3, 16-bit field
    error vs the FP32 run at err_n^3 (same generator, same parameters)."""
    out = {"workload": f"D3Q19 {n}^3 porous duct (periodic x, walls y/z, 5% spheres r 4-12), g_x 1e-6, tau 0.56",
           "storage": {}}
    ids = porous_spheres(n, n, n)
    cells = n ** 3
    fluid = int(np.count_nonzero(ids == 0))
    out["solid_fraction"] = round(1 - fluid / cells, 4)
    rho = np.ones(cells, np.float32)
    u = np.zeros((cells, 3), np.float32)
    force = np.array([1e-6, 0, 0], np.float32)
    small_ids = porous_spheres(err_n, err_n, err_n)
    sc = err_n ** 3
    mask = small_ids == 0
    shape = ((err_n,) * 3, {"velocity_set": "D3Q19", "periodic": (1, 0, 0)})

    def setup(g):
        g.init_fields(np.ones(sc, np.float32), np.zeros((sc, 3), np.float32), force=np.tile(force, (sc, 1)),
                      boundary_id=small_ids)

    ref = run_fields(bfa, shape, 0.56, "fp32", err_ticks, setup)[mask]
    for st in ("fp32", "fp16", "bf16"):
        s = bfa.Solver((n, n, n), 0.56, periodic=(1, 0, 0), velocity_set="D3Q19", storage=st)
        s.init_fields(rho, u, boundary_id=ids)
        s.set_param("constant_force", force)
        wall, ev = timed_run(s, steps, warmup)
        s.close()
        kms = ev / steps
        out["storage"][st] = {"mlups": round(cells * steps / wall / 1e6, 1),
                              "fluid_mlups": round(fluid * steps / wall / 1e6, 1),
                              "roofline": roofline(bytes_per_cell(19, st, 1), cells, kms,
                                                   load_traffic(f"step_d3q19_porous_{st}_{n}"))}
        if st != "fp32":
            er, eu = field_error(run_fields(bfa, shape, 0.56, st, err_ticks, setup)[mask], ref)
            out["storage"][st]["error_vs_fp32"] = {"rho": er, "u": eu, "workload": f"{err_n}^3, {err_ticks} ticks"}
    return out


def spinodal_thermal_config(bfa, n=192, steps=100, warmup=10):
    """D3Q27 Shan-Chen (G = -5, psi = 1 - exp(-rho), rho = 2 + U(+-0.01), tau = 1)
    with a D3Q7 thermal lattice: hot wall T = 1 at y = 0, cold wall T = 0 at
    y = n-1, Boussinesq buoyancy with gravity (0, -1e-5, 0), T_ref = 0.5;
    FP64 mass sums every 100 ticks inside the timed region.  Run with the
    reference's lagged neighbour pseudopotential and with the synchronous one
    (a density pre-pass per tick)."""
    rho, u, ids = spinodal(n, n, n)
    cells = n ** 3
    out = {"workload": f"D3Q27 Shan-Chen {n}^3 (G -5, tau 1) + D3Q7 thermal walls, FP32"}
    for mode in ("lagged", "synchronous"):
        s = bfa.Solver((n, n, n), 1.0, periodic=(1, 0, 1), velocity_set="D3Q27", multiphase=True)
        s.set_object(1, temperature=1.0)
        s.set_object(2, temperature=0.0)
        s.set_param("interaction_strength", -5.0)
        s.set_param("reference_temperature", 0.5)
        s.set_param("shan_chen_synchronous", 1.0 if mode == "synchronous" else 0.0)
        T = np.full(cells, 0.5, np.float32)
        force = np.tile(np.array([0, -1e-5, 0], np.float32), (cells, 1))
        s.init_fields(rho, u, force=force, temperature=T, boundary_id=ids)
        m0 = s.conservation()[0]
        wall, ev = timed_run(s, steps, warmup, conservation_every=100)
        m1 = s.last_step_sums()[0]
        s.close()
        kms = ev / steps
        # the synchronous pre-pass reads the main lattice once more (27 slots),
        # the boundary id, and writes psi
        extra = 27 * 4 + 1 + 4 if mode == "synchronous" else 0
        out[mode] = {"mlups": round(cells * steps / wall / 1e6, 1),
                     "roofline": roofline(2 * 27 * 4 + 2 * 7 * 4 + 8 + 1 + extra, cells, kms,
                                          load_traffic(f"step_d3q27_shanchen_thermal_{mode}_{n}")),
                     "fp64_mass_relative_drift": abs(m1 - m0) / m0}
    return out


def cavity_config(bfa, n=256, steps=1000, warmup=50):
    """D2Q9 lid-driven cavity, tau 0.6, lid u 0.1, halfway bounce-back walls (2.4 MB:
    cache resident, reported as time per tick)."""
    s = bfa.Solver((n, n, 1), 0.6, periodic=(0, 0, 1), velocity_set="D2Q9")
    s.set_object(2, velocity=(0.1, 0.0, 0.0))
    cells = n * n
    s.init_fields(np.ones(cells, np.float32), np.zeros((cells, 3), np.float32), boundary_id=lid_cavity(n, n))
    wall, ev = timed_run(s, steps, warmup)
    s.close()
    return {"workload": f"D2Q9 lid-driven cavity {n}^2, tau 0.6, lid 0.1", "us_per_tick": round(wall * 1e6 / steps, 2),
            "kernel_us_per_tick": round(ev * 1e3 / steps, 2), "mlups": round(cells * steps / wall / 1e6, 1)}


def isosurface_config(bfa, n=256, repeats=5):
    """Marching cubes on the device of the saved density of the D3Q19 n^3 TGV
    after 100 ticks, at the level rho = 1 (the surface p' = 0); device time of
    the classification, prefix-sum and emit passes (best of `repeats`)."""
    rho, u = taylor_green(n)
    s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19")
    s.init_fields(rho, u)
    s.step(100, save=True)
    best = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        v, t = s.isosurface(1.0)
        wall = time.perf_counter() - t0
        dev = bfa.capi.marching_cubes_device_ms()
        best = dev if best is None else min(best, dev)
    s.close()
    return {"workload": f"marching cubes of the D3Q19 {n}^3 TGV density (100 ticks) at rho = 1",
            "vertices": int(len(v)), "triangles": int(len(t)), "device_ms": round(best, 3),
            "wall_ms_with_mesh_download": round(wall * 1e3, 2),
            "mcells_per_s": round(n ** 3 / best / 1e3, 1)}


def adjoint_ratio(bfa, res, vs, steps=2000, k=100, budget=0):
    """Gradient (forward with checkpoints + recomputation + adjoint) time over
    the plain forward time of the differentiable channel, device-timed.
    Returns the report and dJ/dgamma."""
    ch = bfa.capi.Channel(res, 0.6, 0.05, velocity_set=vs)
    ch.set_porosity(porosity(*res))
    ch.forward(steps)                 # warm-up
    # both sides warm and repeated, the fastest run of each: the small D2Q9
    # channel's launch-bound ticks vary by 10 % from run to run
    fwd = float("inf")
    for _ in range(3):
        J = ch.forward(steps)
        fwd = min(fwd, ch.timing()[0])
    t = None
    for _ in range(2):
        while True:
            try:
                Jg, dg, du = ch.gradient(steps, checkpoint_every=k, trajectory_budget_bytes=budget)
                break
            except bfa.capi.BFError:
                if budget == 0:
                    raise
                budget //= 2          # device memory short of the estimate
        tt = ch.timing()
        if t is None or tt[2] < t[2]:
            t = tt
    ch.close()
    n = 1
    for r in res:
        n *= r
    state_bytes = 4 * (19 if vs == "D3Q19" else 9) * ((n + 63) // 64 * 64)
    resident = min(steps // k, max(1, budget // (state_bytes * k))) if budget else 0
    # the forward pass keeps its states for the resident segments (at least
    # one) and, when others are recomputed, for the segment before them
    nseg = (steps + k - 1) // k
    kept = max(1, resident)
    kept = nseg if kept >= nseg else kept + 1
    return {"workload": f"channel {vs} {'x'.join(map(str, res))}, tau 0.6, u_in 0.05, porosity 0.02 x box-smoothed "
                        f"U[0,1], {steps} ticks, checkpoint every {k}",
            "trajectory_budget_GB": round(budget / 1e9, 1), "resident_segments": resident,
            "segments": steps // k, "segments_recomputed": nseg - kept,
            "J": J, "forward_ms": round(fwd, 3), "gradient_ms": round(t[2], 3),
            "recompute_ms": round(t[1], 3), "recompute_beside_sweep": True,
            "ratio": round(t[2] / fwd, 3), "dJ_duin": du}, dg


def _gpu_fields(bfa, shape, tau, storage, ticks, setup, periodic=(1, 1, 1)):
    g = bfa.Solver(shape, tau, periodic=periodic, velocity_set="D3Q19", storage=storage)
    setup(g)
    g.step(ticks, save=True)
    vd = g.read("velocity_density").astype(np.float64)
    g.close()
    return vd


def cpu_leg(bfa, storages, n_err=64, ticks_err=1000, n=128, ticks=400):
    """The benchmark's CPU leg, the only part of it that runs the CPU oracle
    (oracle/, test infrastructure):
      cpu_baseline  the FP32 restatement (oracle/liboracle_f32.so) timed on the
                    host cores on a bounded sample of the headline workload;
      error_vs_fp64 the FP64 oracle run on the Taylor-Green box and on the
                    porous duct at n_err^3 for ticks_err ticks, as the absolute
                    reference for the GPU runs of every storage type on the
                    same inputs: ||d rho||/||rho||, ||d u||/||u|| and the
                    largest absolute differences over the fluid cells (the
                    oracle is the checker here, never the thing measured)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rho, u = taylor_green(n)
    o = oracle.Oracle((n, n, n), 0.55, velocity_set="D3Q19", precision="f32")
    o.init_fields(rho, u)
    threads = int(oracle.load("f32").orc_threads())
    t0 = time.perf_counter()
    o.step(ticks)
    dt = time.perf_counter() - t0
    o.close()
    base = {"value": round(n ** 3 * ticks / dt / 1e6, 3), "unit": "MLUPs/s", "cores": threads, "kind": "port",
            "sample": f"D3Q19 TGV {n}^3 FP32, {ticks} ticks, oracle/liboracle_f32.so, {threads} OpenMP threads "
                      f"({dt:.1f} s)"}
    if not storages:
        return base, None
    m = n_err
    cells = m ** 3
    errors = {"workloads": f"D3Q19 {m}^3, {ticks_err} ticks, against the FP64 oracle (fluid cells)"}
    t0 = time.perf_counter()
    # Taylor-Green box (tau 0.55, u0 0.05, fully periodic)
    rho, u = taylor_green(m)
    o = oracle.Oracle((m, m, m), 0.55, velocity_set="D3Q19", precision="f64")
    o.init_fields(rho, u)
    o.step(ticks_err, save=True)
    ref = o.read("velocity_density")
    o.close()
    errors["tgv"] = {}
    for st in storages:
        vd = _gpu_fields(bfa, (m, m, m), 0.55, st, ticks_err, lambda g: g.init_fields(rho, u))
        errors["tgv"][st] = field_errors(vd, ref)
    # porous duct (periodic x, walls y/z, spheres r 4-12, g_x 1e-6, tau 0.56)
    ids = porous_spheres(m, m, m)
    fluid = ids == 0
    force = np.array([1e-6, 0, 0], np.float32)

    def setup(s):
        # (initialisation without a force array resets the force; the constant force follows it)
        s.init_fields(np.ones(cells, np.float32), np.zeros((cells, 3), np.float32), boundary_id=ids)
        s.set_param("constant_force", force)

    o = oracle.Oracle((m, m, m), 0.56, periodic=(1, 0, 0), velocity_set="D3Q19", precision="f64")
    setup(o)
    o.step(ticks_err, save=True)
    ref = o.read("velocity_density")[fluid]
    o.close()
    errors["porous"] = {"solid_fraction": round(float(1 - fluid.mean()), 4)}
    for st in storages:
        vd = _gpu_fields(bfa, (m, m, m), 0.56, st, ticks_err, setup, periodic=(1, 0, 0))[fluid]
        errors["porous"][st] = field_errors(vd, ref)
    # D3Q27 Shan-Chen (lagged psi, G -5, tau 1) + D3Q7 thermal hot/cold walls,
    # FP32 (the configuration's storage), at 32^3 for 500 ticks
    ms, ticks_sc = 32, 500
    rho_s, u_s, ids_s = spinodal(ms, ms, ms)
    cs = ms ** 3
    fluid_s = ids_s == 0

    def setup_sc(s):
        s.set_object(1, temperature=1.0)
        s.set_object(2, temperature=0.0)
        s.set_param("interaction_strength", -5.0)
        s.set_param("reference_temperature", 0.5)
        s.init_fields(rho_s, u_s, force=np.tile(np.array([0, -1e-5, 0], np.float32), (cs, 1)),
                      temperature=np.full(cs, 0.5, np.float32), boundary_id=ids_s)

    o = oracle.Oracle((ms,) * 3, 1.0, periodic=(1, 0, 1), velocity_set="D3Q27", multiphase=True, precision="f64")
    setup_sc(o)
    o.step(ticks_sc, save=True)
    ref = o.read("velocity_density")[fluid_s]
    o.close()
    g = bfa.Solver((ms,) * 3, 1.0, periodic=(1, 0, 1), velocity_set="D3Q27", multiphase=True)
    setup_sc(g)
    g.step(ticks_sc, save=True)
    vd = g.read("velocity_density").astype(np.float64)[fluid_s]
    g.close()
    errors["shan_chen_thermal"] = {"workload": f"D3Q27 Shan-Chen + D3Q7 thermal walls {ms}^3, {ticks_sc} ticks",
                                   "fp32": field_errors(vd, ref)}
    # the differentiable D2Q9 channel (512 x 128, 2000 ticks, checkpoints every
    # 100): the GPU gradient against the FP64 CPU adjoint of the same schedule
    res_c, ticks_c = (512, 128, 1), 2000
    gam = porosity(*res_c)
    oc = oracle.OracleChannel(res_c, 0.6, 0.05, velocity_set="D2Q9", precision="f64")
    oc.set_porosity(gam)
    J64, dg64, du64 = oc.gradient_checkpointed(ticks_c, 100)
    ch = bfa.capi.Channel(res_c, 0.6, 0.05, velocity_set="D2Q9")
    ch.set_porosity(gam)
    Jg, dgg, dug = ch.gradient(ticks_c, checkpoint_every=100, trajectory_budget_bytes=0)
    ch.close()
    rows = np.r_[0, np.ones(res_c[1] - 2), 0].repeat(res_c[0]) > 0
    errors["channel_gradient"] = {
        "workload": f"channel D2Q9 {res_c[0]}x{res_c[1]}, {ticks_c} ticks, checkpoint every 100, FP32 on the GPU "
                    f"against the FP64 adjoint",
        "J": abs(Jg - J64) / abs(J64),
        "dJ_dgamma": float(np.linalg.norm(dgg[rows] - dg64[rows]) / np.linalg.norm(dg64[rows])),
        "dJ_duin": abs(dug - du64) / abs(du64)}
    errors["oracle_seconds"] = round(time.perf_counter() - t0, 1)
    return base, errors


def counters_sample(bfa, n=256, ticks=12):
    """Short workload for one rocprofv3 --pmc pass: a 1 GiB device copy of known
    traffic (counter calibration), then `ticks` ticks of the FP32 and FP16 runs
    (12: six tick pairs in FP32, three sweeps of two pairs in FP16)."""
    import torch  # noqa: F401  (one HIP runtime with torch)
    print("copy GB/s", round(probe.copy_bandwidth(1 << 30, 3), 1), flush=True)
    rho, u = taylor_green(n)
    for st in ("fp32", "fp16"):
        s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=st)
        s.init_fields(rho, u)
        s.step(ticks)
        s.synchronize()
        s.close()
    print("counters sample done", flush=True)


def porous_sample(bfa, ticks, n=384):
    """Short workload for rocprofv3: `ticks` ticks of the D3Q19 porous duct
    (constant g_x) in FP32, FP16 and BF16."""
    import torch  # noqa: F401  (one HIP runtime with torch)
    ids = porous_spheres(n, n, n)
    cells = n ** 3
    for st in ("fp32", "fp16", "bf16"):
        s = bfa.Solver((n, n, n), 0.56, periodic=(1, 0, 0), velocity_set="D3Q19", storage=st)
        s.init_fields(np.ones(cells, np.float32), np.zeros((cells, 3), np.float32), boundary_id=ids)
        s.set_param("constant_force", np.array([1e-6, 0, 0], np.float32))
        s.step(ticks)
        s.synchronize()
        s.close()
    print(f"porous sample: {n}^3, {ticks} ticks per storage type", flush=True)


def tgv_sample(bfa, ticks, n=256):
    """Short workload for rocprofv3: the 256^3 TGV in FP32, FP16 and BF16,
    each 200 ticks past the power-management transient after set-up, then
    `ticks` ticks in one call (no saves: whole tick pairs, or sweeps of two
    pairs for 16-bit storage, when `ticks` is a multiple of four)."""
    import torch  # noqa: F401  (one HIP runtime with torch)
    rho, u = taylor_green(n)
    for st in ("fp32", "fp16", "bf16"):
        s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=st)
        s.init_fields(rho, u)
        s.step(200)
        s.synchronize()
        s.step(ticks)
        s.synchronize()
        s.close()
    print(f"TGV sample: {n}^3, 200 + {ticks} ticks per storage type", flush=True)


def spinodal_sample(bfa, ticks, n=192):
    """Short workload for rocprofv3: `ticks` ticks of the D3Q27 Shan-Chen +
    D3Q7 thermal workload with the lagged pseudopotential."""
    import torch  # noqa: F401  (one HIP runtime with torch)
    rho, u, ids = spinodal(n, n, n)
    cells = n ** 3
    s = bfa.Solver((n, n, n), 1.0, periodic=(1, 0, 1), velocity_set="D3Q27", multiphase=True)
    s.set_object(1, temperature=1.0)
    s.set_object(2, temperature=0.0)
    s.set_param("interaction_strength", -5.0)
    s.set_param("reference_temperature", 0.5)
    s.init_fields(rho, u, force=np.tile(np.array([0, -1e-5, 0], np.float32), (cells, 1)),
                  temperature=np.full(cells, 0.5, np.float32), boundary_id=ids)
    s.step(ticks)
    s.synchronize()
    s.close()
    print(f"spinodal sample: {n}^3, {ticks} ticks", flush=True)


def transient_sample(bfa, n=256, chunks=4, ticks=300, pause_s=1.0):
    """The FP16 256^3 TGV in `chunks` runs of `ticks` ticks with the GPU idle
    for `pause_s` between them (for a rocprofv3 kernel trace of the
    power-management transient after idle)."""
    rho, u = taylor_green(n)
    s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage="fp16")
    s.init_fields(rho, u)
    for k in range(chunks):
        s.step(ticks)
        s.synchronize()
        print(f"transient sample: chunk {k} done", flush=True)
        time.sleep(pause_s)
    s.close()


def adjoint_sample(bfa, ticks):
    """Short workload for rocprofv3: the D3Q19 128^3 channel forward and its
    gradient over `ticks` ticks with checkpoints only (K = 100)."""
    import torch  # noqa: F401  (one HIP runtime with torch)
    res = (128, 128, 128)
    ch = bfa.capi.Channel(res, 0.6, 0.05, velocity_set="D3Q19")
    ch.set_porosity(porosity(*res))
    ch.forward(ticks)
    ch.gradient(ticks, checkpoint_every=100, trajectory_budget_bytes=0)
    t = ch.timing()
    ch.close()
    print(f"adjoint sample: {ticks} ticks, gradient {t[2]:.1f} ms (recompute {t[1]:.1f} ms)", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--fp32-only", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true")
    ap.add_argument("--counters-sample", action="store_true",
                    help="run only the short workload used for rocprofv3 --pmc passes")
    ap.add_argument("--porous-sample", type=int, default=0, metavar="TICKS",
                    help="run only TICKS ticks of the 384^3 porous duct per storage type (for rocprofv3)")
    ap.add_argument("--tgv-sample", type=int, default=0, metavar="TICKS",
                    help="run only 200 + TICKS ticks of the 256^3 TGV per storage type (for rocprofv3)")
    ap.add_argument("--spinodal-sample", type=int, default=0, metavar="TICKS",
                    help="run only TICKS ticks of the 192^3 Shan-Chen + thermal workload (for rocprofv3)")
    ap.add_argument("--transient-sample", action="store_true",
                    help="run only the FP16 TGV in chunks with idle pauses (for rocprofv3)")
    ap.add_argument("--adjoint-sample", type=int, default=0, metavar="TICKS",
                    help="run only the D3Q19 128^3 channel gradient over TICKS ticks (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solver runs on a single MI355X; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import boltzmannfluids_amd as bfa
    bfa.load_library()
    if args.counters_sample:
        counters_sample(bfa, args.size)
        return
    if args.porous_sample:
        porous_sample(bfa, args.porous_sample)
        return
    if args.tgv_sample:
        tgv_sample(bfa, args.tgv_sample)
        return
    if args.spinodal_sample:
        spinodal_sample(bfa, args.spinodal_sample)
        return
    if args.adjoint_sample:
        adjoint_sample(bfa, args.adjoint_sample)
        return
    if args.transient_sample:
        transient_sample(bfa)
        return
    n = args.size
    stores = ["fp32"] if args.fp32_only else ["fp32", "fp16", "bf16"]
    res = {st: run_tgv(bfa, n, st, args.steps, args.warmup) for st in stores}
    main_run = res["fp32"]
    line = {
        "metric": "D3Q19 MLUPs/s & %HBM peak, FP16-store vs FP32 @<=1e-3 rel err; adjoint/fwd time",
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_min_s": WARMUP_MIN_S,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Taylor-Green vortex initial state)",
        "config": {"workload": f"D3Q19 BGK esoteric-pull, {n}^3 fully periodic, TGV u0=0.05, tau=0.55, FP32 storage",
                   "cells": n ** 3, "bytes_per_cell": bytes_per_cell(19, "fp32")},
        "roofline": main_run["roofline"],
        # roofline.frac divides the algorithmic bytes by the kernels' device
        # time (HIP events on the solver's stream); the same bytes over the
        # host wall clock of the timed region (value, ms_per_step):
        "wall_clock_frac": round(bytes_per_cell(19, "fp32") * main_run["mlups"] * 1e6 / (HBM_PEAK_GBS * 1e9), 4),
    }
    # achievable-bandwidth ceilings on this device: a plain copy, and the tick's
    # own access pattern (19 FP32 slots read and rewritten in place) without
    # arithmetic, at the FP32 lattice size and at the 16-bit lattice's bytes
    copy = probe.copy_bandwidth(1 << 30, 20)
    inplace = probe.inplace_bandwidth(n ** 3, 19, 20)
    inplace_half = probe.inplace_bandwidth(n ** 3 // 2, 19, 20)
    # the AA tick's two memory patterns (odd: every slot at the cell; even: the
    # y/z-shifted rows) at the workload's size; a tick pair takes 1/B_odd +
    # 1/B_even per byte, so their harmonic mean is the pair's ceiling
    aa_odd = probe.aa_bandwidth(n, n, n, False, 20)
    aa_even = probe.aa_bandwidth(n, n, n, True, 20)
    aa_pair = 2.0 / (1.0 / aa_odd + 1.0 / aa_even)
    # the two-lattice pull pattern (y/z-shifted reads, local writes to a second
    # lattice) at the same size, for comparison with the AA tick pair
    pull = probe.pull_bandwidth(n, n, n, 20)
    line["ceiling"] = {"copy_GBs": round(copy, 1), "inplace_19slot_GBs": round(inplace, 1),
                       "inplace_19slot_half_bytes_GBs": round(inplace_half, 1),
                       "fp32_kernel_frac_of_inplace": round(main_run["roofline"]["achieved"] / inplace, 4),
                       "aa_odd_GBs": round(aa_odd, 1), "aa_even_GBs": round(aa_even, 1),
                       "aa_tick_pair_GBs": round(aa_pair, 1),
                       "two_lattice_pull_GBs": round(pull, 1),
                       "fp32_kernel_frac_of_aa_pattern": round(main_run["roofline"]["achieved"] / aa_pair, 4)}
    for st in ("fp16", "bf16"):
        if st in res:
            r = res[st]
            line[f"{st}_shifted"] = {"value": round(r["mlups"], 1), "unit": "MLUPs/s",
                                     "speedup_vs_fp32": round(r["mlups"] / main_run["mlups"], 3),
                                     "roofline": r["roofline"],
                                     "frac_of_inplace_half_bytes": round(r["roofline"]["achieved"] / inplace_half, 4)}
    if not args.no_accuracy and len(res) > 1:
        line["error_vs_fp32"] = accuracy_vs_fp32(bfa, list(res))
    if not args.no_other_configs:
        line["other_configs"] = [porous_config(bfa), spinodal_thermal_config(bfa), cavity_config(bfa),
                                 isosurface_config(bfa, n)]
    if not args.no_adjoint:
        # checkpoints only (every 100 ticks, segments recomputed), then with as
        # much of the trajectory resident in HBM as the device holds
        import torch
        runs = []
        for res, vs in (((512, 128, 1), "D2Q9"), ((128, 128, 128), "D3Q19")):
            r0, dg0 = adjoint_ratio(bfa, res, vs)
            free, _ = torch.cuda.mem_get_info()
            r1, dg1 = adjoint_ratio(bfa, res, vs, budget=max(0, free - (24 << 30)))
            r1["same_gradient_as_checkpointed"] = bool(np.array_equal(dg0, dg1))
            runs += [r0, r1]
        line["adjoint_over_forward"] = runs
    if not args.no_cpu_baseline:
        line["cpu_baseline"], err64 = cpu_leg(bfa, [] if args.no_accuracy else stores)
        if err64:
            line["error_vs_fp64"] = err64
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
