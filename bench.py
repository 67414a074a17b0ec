"""Benchmark: D3Q19 BGK in-place (esoteric pull) tick on one MI355X.

Workload (default): 256^3 fully periodic D3Q19 box, Taylor-Green vortex
(u0 = 0.05, rho = 1 + 3 p'), tau = 0.55, FP32 storage.  One step = one lattice
tick over the whole grid; the populations stay resident in HBM.  Prints one
JSON line: MLUPs/s of the FP32 run as `value`, the HBM roofline of the step
kernel, the FP16/BF16 shifted-storage runs with their error against the FP64
CPU oracle, and a CPU baseline (the FP32 CPU restatement, one core).

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

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E published peak
BYTES_PER_CELL = {"fp32": 2 * 19 * 4, "fp16": 2 * 19 * 2, "bf16": 2 * 19 * 2}  # 2*Q*s, no flags


def taylor_green(n, u0=0.05, rho0=1.0):
    k = 2 * np.pi / n
    idx = np.arange(n, dtype=np.float64)
    sx, cx = np.sin(k * idx), np.cos(k * idx)
    c2 = np.cos(2 * k * idx)
    # cell order z*n*n + y*n + x
    ux = u0 * sx[None, None, :] * cx[None, :, None] * cx[:, None, None]
    uy = -u0 * cx[None, None, :] * sx[None, :, None] * cx[:, None, None]
    p = (rho0 * u0 ** 2 / 16.0) * (c2[None, None, :] + c2[None, :, None]) * (c2[:, None, None] + 2.0)
    rho = (rho0 + 3.0 * p).astype(np.float32).ravel()
    u = np.zeros((n ** 3, 3), np.float32)
    u[:, 0] = ux.ravel()
    u[:, 1] = uy.ravel()
    return rho, u


def load_traffic(storage, size):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by
    a copy-kernel calibration) if a summary for this workload is committed."""
    path = os.path.join(ROOT, "profiles", f"step_d3q19_{storage}_{size}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def run_gpu(bfa, n, storage, steps, warmup):
    import torch
    rho, u = taylor_green(n)
    s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=storage)
    s.init_fields(rho, u)
    s.step(warmup)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = s.step_timed(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = n ** 3
    out = {
        "mlups": cells * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": ev_ms / steps,
    }
    b = BYTES_PER_CELL[storage] * cells
    achieved = b / (out["kernel_ms"] * 1e-3) / 1e9
    traffic = load_traffic(storage, n)
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}
    s.close()
    return out


def accuracy_vs_fp64(bfa, storage, n=48, ticks=200):
    """||d rho||/||rho|| and ||d u||/||u|| of a GPU run against the FP64 oracle,
    same TGV workload at n^3 (the oracle's cost bounds n)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rho, u = taylor_green(n)
    g = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=storage)
    g.init_fields(rho, u)
    g.step(ticks, save=True)
    a = g.read("velocity_density").astype(np.float64)
    g.close()
    o = oracle.Oracle((n, n, n), 0.55, velocity_set="D3Q19", precision="f64")
    o.init_fields(rho, u)
    o.step(ticks, save=True)
    b = o.read("velocity_density")
    er = float(np.linalg.norm(a[:, 3] - b[:, 3]) / np.linalg.norm(b[:, 3]))
    eu = float(np.linalg.norm(a[:, :3] - b[:, :3]) / np.linalg.norm(b[:, :3]))
    return {"rho": er, "u": eu, "workload": f"TGV {n}^3 tau=0.55 u0=0.05, {ticks} ticks"}


def porosity(X, Y, Z=1, seed=20261203, scale=0.02):
    """0.02 x (3x3 box-smoothed seeded U[0,1]), zero on the wall rows."""
    rng = np.random.default_rng(seed)
    a = rng.random((Z, Y, X))
    k = np.ones(3) / 3.0
    for ax in ((1, 2) if Z == 1 else (0, 1, 2)):
        a = np.apply_along_axis(lambda v: np.convolve(np.pad(v, 1, mode="edge"), k, mode="valid"), ax, a)
    g = np.clip(scale * a, 0.0, 1.0)
    g[:, 0, :] = 0.0
    g[:, -1, :] = 0.0
    return g.ravel().astype(np.float32)


def adjoint_ratio(bfa, res, vs, steps=2000, k=100, budget=0):
    """Gradient (forward with checkpoints + recomputation + adjoint) time over
    the plain forward time of the differentiable channel, device-timed."""
    ch = bfa.capi.Channel(res, 0.6, 0.05, velocity_set=vs)
    ch.set_porosity(porosity(*res))
    ch.forward(steps)                 # warm-up
    J = ch.forward(steps)
    fwd = ch.timing()[0]
    Jg, dg, du = ch.gradient(steps, checkpoint_every=k, trajectory_budget_bytes=budget)
    t = ch.timing()
    ch.close()
    return {"workload": f"channel {vs} {'x'.join(map(str, res))}, tau 0.6, u_in 0.05, {steps} ticks, "
                        f"checkpoint every {k}, resident budget {budget} B",
            "J": J, "forward_ms": round(fwd, 3), "gradient_ms": round(t[2], 3), "recompute_ms": round(t[1], 3),
            "ratio": round(t[2] / fwd, 3), "dJ_duin": du}


def cpu_baseline(n=64, ticks=150):
    """FP32 CPU restatement (oracle/liboracle_f32.so), one thread, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rho, u = taylor_green(n)
    o = oracle.Oracle((n, n, n), 0.55, velocity_set="D3Q19", precision="f32")
    o.init_fields(rho, u)
    t0 = time.perf_counter()
    o.step(ticks)
    dt = time.perf_counter() - t0
    return {"value": round(n ** 3 * ticks / dt / 1e6, 3), "unit": "MLUPs/s", "cores": 1, "kind": "port",
            "sample": f"D3Q19 TGV {n}^3 FP32, {ticks} ticks, oracle/liboracle_f32.so single thread ({dt:.1f} s)"}


def counters_sample(bfa, n=256, ticks=10):
    """Short workload for one rocprofv3 --pmc pass: a 1 GiB device copy of known
    traffic (counter calibration), then `ticks` ticks of the FP32 and FP16 runs."""
    import torch  # noqa: F401  (one HIP runtime with torch)
    print("copy GB/s", round(bfa.capi.copy_bandwidth(1 << 30, 3), 1), flush=True)
    rho, u = taylor_green(n)
    for st in ("fp32", "fp16"):
        s = bfa.Solver((n, n, n), 0.55, velocity_set="D3Q19", storage=st)
        s.init_fields(rho, u)
        s.step(ticks)
        s.synchronize()
        s.close()
    print("counters sample done", flush=True)


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
    ap.add_argument("--counters-sample", action="store_true",
                    help="run only the short workload used for rocprofv3 --pmc passes")
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
    n = args.size
    res = {st: run_gpu(bfa, n, st, args.steps, args.warmup) for st in (["fp32"] if args.fp32_only else ["fp32", "fp16", "bf16"])}
    main = res["fp32"]
    line = {
        "metric": "D3Q19 MLUPs/s & %HBM peak, FP16-store vs FP32 @<=1e-3 rel err; adjoint/fwd time",
        "value": round(main["mlups"], 1),
        "unit": "MLUPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Taylor-Green vortex initial state)",
        "config": {"workload": f"D3Q19 BGK esoteric-pull, {n}^3 fully periodic, TGV u0=0.05, tau=0.55, FP32 storage",
                   "cells": n ** 3, "bytes_per_cell": BYTES_PER_CELL["fp32"]},
        "roofline": main["roofline"],
    }
    for st in ("fp16", "bf16"):
        if st in res:
            r = res[st]
            line[f"{st}_shifted"] = {"value": round(r["mlups"], 1), "unit": "MLUPs/s",
                                     "speedup_vs_fp32": round(r["mlups"] / main["mlups"], 3),
                                     "roofline": r["roofline"]}
    if not args.no_accuracy:
        line["error_vs_fp64"] = {st: accuracy_vs_fp64(bfa, st) for st in res}
    if not args.no_adjoint:
        line["adjoint_over_forward"] = [adjoint_ratio(bfa, (512, 128, 1), "D2Q9"),
                                        adjoint_ratio(bfa, (128, 128, 128), "D3Q19")]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
