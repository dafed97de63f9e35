"""Benchmark: MLUPS of the D3Q19 step on one MI355X.

Workload (default): air over a moving sinusoidal bed, 256 x 128 x 128, KBC
collision + Smagorinsky LES, per-link open-fraction moving-wall bounce-back,
log-law inflow / zero-gradient outflow, free-slip top (the configuration the
headline metric is quoted on).  Synthetic seeded inputs; FP32 populations.

Prints ONE JSON line:
  value / ms_per_step      whole-step throughput (MLUPS), populations in HBM
  roofline                 the fused step kernels: algorithmic bytes
                           (2 x 19 x 4 B per lattice update) over their
                           HIP-event-timed average duration, against 8 TB/s
  cpu_baseline             the CPU restatement of the same step (oracle,
                           single thread) on a bounded sample of the workload
  variants                 the same measurement for FP16 / BF16 storage and
                           the regularized collision, plus the plain-channel
                           and VOF workloads (reported, not the headline)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT / "lbm-3d-les-air-over-waves_amd"), str(ROOT / "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
BYTES_PER_LU = {"fp32": 152, "fp16": 76, "bf16": 76, "fp64": 304}


def run_device(env: dict, steps: int, warmup: int):
    import torch

    from harness import setup_case
    from lbm_config import parse_env
    from lbm_native import Case

    cfg = parse_env(env)
    case = Case(cfg.to_struct())
    setup_case(case, cfg)
    case.set_option("timing", 1)
    if warmup:
        case.step(warmup)
    torch.cuda.synchronize()
    case.sync()
    t0 = time.perf_counter()
    case.step(steps)
    case.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step_ev, ms_kernel = case.last_timing()
    cells = cfg.Nx * cfg.Ny * cfg.Nz
    out = {
        "cells": cells,
        "ms_per_step": 1e3 * wall / steps,
        "mlups": cells * steps / wall / 1e6,
        "ms_kernel": ms_kernel,
        "ms_step_events": ms_step_ev,
        "path": case.path(),
        "storage": cfg.storage,
    }
    case.close()
    return out


def cpu_baseline(env: dict, budget_s: float = 15.0):
    """Oracle (CPU restatement, one thread) on a bounded sample: as many full
    steps of the workload as fit in the budget (at least one)."""
    from harness import setup_case
    from lbm_config import parse_env
    from oracle_ref import OracleCase

    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct())
    setup_case(o, cfg)
    cells = cfg.Nx * cfg.Ny * cfg.Nz
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    o.close()
    return {
        "value": round(cells * n / wall / 1e6, 3),
        "unit": "MLUPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{n} full step(s) of the {cfg.Nx}x{cfg.Ny}x{cfg.Nz} workload on the CPU oracle (C, -O2, 1 thread)",
    }


def serving(n_cases: int, storage: str, seed: int = 7):
    """Serving stream (half air over a moving sinusoidal bed, half water
    Gaussian mounds; sizes 32^3 .. 128x64x64; U[500, 2000] steps):
    the same requests run one at a time (sequential plain runs) and through
    the ensemble server, first as a burst (its capacity), then as Poisson
    arrivals at 0.8 x that capacity (latency)."""
    import numpy as np

    from lbm_server import EnsembleServer, ensemble_stream
    from lbm_solver import Solver
    from lbm_config import parse_env

    stream = ensemble_stream(n_cases, seed=seed, storage=storage)
    lups = sum(parse_env(e).Nx * parse_env(e).Ny * parse_env(e).Nz * st for e, st in stream)
    t0 = time.perf_counter()
    for env, steps in stream:
        s = Solver(config=parse_env(env))
        s.initialize()
        s.step(steps)
        s.case.download(1)
        s.close()
    seq_wall = time.perf_counter() - t0
    budget = 200 << 30
    srv = EnsembleServer(budget, max_active=64, chunk_steps=16, n_workers=4)
    t0 = time.perf_counter()
    for env, steps in stream:
        srv.submit(env, steps)
    srv.run_until_idle()
    burst = srv.stats(t0)
    srv.close()
    capacity = n_cases / burst["wall_s"]
    rate = 0.8 * capacity
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_cases))
    srv = EnsembleServer(budget, max_active=64, chunk_steps=16, n_workers=4)
    t0 = time.perf_counter()
    nxt = 0
    while nxt < n_cases or not srv.idle():
        now = time.perf_counter() - t0
        while nxt < n_cases and arrivals[nxt] <= now:
            srv.submit(*stream[nxt])
            nxt += 1
        srv.pump()
        time.sleep(2e-4)
    poisson = srv.stats(t0)
    srv.close()
    return {
        "cases": n_cases,
        "storage_air": storage,
        "lattice_updates": lups,
        "sequential_mlups": round(lups / seq_wall / 1e6, 1),
        "ensemble_mlups": round(burst["aggregate_mlups"], 1),
        "speedup_vs_sequential": round(burst["aggregate_mlups"] / (lups / seq_wall / 1e6), 3),
        "poisson_rate_per_s": round(rate, 3),
        "poisson_mlups": round(poisson["aggregate_mlups"], 1),
        "latency_p50_s": round(poisson["latency_p50_s"], 4),
        "latency_p99_s": round(poisson["latency_p99_s"], 4),
        "peak_active": max(burst["peak_active"], poisson["peak_active"]),
        "failed": burst["failed"] + poisson["failed"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="waves_256x128x128")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-cases", type=int, default=64, help="requests per precision of the serving stream (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (3 launches): the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
            "(the multi-GPU decomposition is described in DESIGN.md, not implemented)",
            file=sys.stderr,
        )
        sys.exit(2)
    import torch

    from lbm_config import parse_env, preset_env

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        x = torch.ones(256 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "launches": 3}))
        return
    env = preset_env(args.workload)
    env["LBM_KERNEL_PATH"] = "auto"
    cfg = parse_env(env)
    main_run = run_device(env, args.steps, args.warmup)
    bpl = BYTES_PER_LU[cfg.storage]
    achieved = main_run["cells"] * bpl / (main_run["ms_kernel"] * 1e-3) / 1e9
    result = {
        "metric": "MLUPS & % HBM peak (FP32 vs FP16-storage); ensemble p50/p99 latency",
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded log-law inflow + 2% uniform noise, seed 12345)",
        "config": {
            "workload": args.workload,
            "grid": [cfg.Nx, cfg.Ny, cfg.Nz],
            "collision": cfg.collision_model,
            "les": bool(cfg.use_LES),
            "storage": cfg.storage,
            "path": "fused" if main_run["path"] == 2 else "staged",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel_ms": round(main_run["ms_kernel"], 5),
            "bytes_per_lu": bpl,
        },
    }
    if not args.no_variants:
        variants = {}
        vsteps = max(20, args.steps // 4)
        for label, extra in (
            ("fp16_storage", {"LBM_STORAGE": "fp16"}),
            ("bf16_storage", {"LBM_STORAGE": "bf16"}),
            ("regularized", {"LBM_COLLISION": "regularized"}),
        ):
            r = run_device(dict(env, **extra), vsteps, 5)
            b = BYTES_PER_LU[parse_env(dict(env, **extra)).storage]
            variants[label] = {
                "mlups": round(r["mlups"], 1),
                "hbm_frac": round(r["cells"] * b / (r["ms_kernel"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            }
        result["variants"] = variants
    if args.serving_cases > 0 and not args.no_variants:
        result["serving"] = {st: serving(args.serving_cases, st) for st in ("fp32", "fp16")}
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(env)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
