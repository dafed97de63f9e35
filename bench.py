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
                           (frac_wall: the same bytes over the whole step);
                           traffic = HBM bytes per step measured by rocprofv3
                           counters (profiles/fused_step_<workload>.md)
  cpu_baseline             the CPU restatement of the same step (oracle,
                           OpenMP over cells, all host threads; and one
                           thread) on a bounded sample of the workload
  variants                 FP16 / BF16 storage and the regularized collision
                           on the same workload
  workloads                the plain channel at 64x32x32 and 512x256x192
                           (FP32 / FP16 / BF16) and the VOF box 192x96x160
                           (FP32 / FP16)
  accuracy                 FP32 / FP16 / BF16 against FP64: the plain
                           channel and the waves case at 256x128x128 after
                           2000 steps, the VOF box after 1500 steps
  graph_launch             one small case, calls of 16 steps: direct
                           launches against HIP-graph launches
  serving                  the ensemble server against sequential runs
                           (streams, and streams + graph-launched chunks)
(all reported beside the headline, which is the default workload)
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


def progress(msg: str):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)
BYTES_PER_LU = {"fp32": 152, "fp16": 76, "bf16": 76, "fp64": 304}
# VOF: the populations once each way plus mass, fill and type (18 B)
VOF_BYTES_PER_LU = {k: v + 18 for k, v in BYTES_PER_LU.items()}
PROFILES = ROOT / "profiles"


def measured_traffic(workload: str, storage: str):
    """HBM bytes per step of the step kernels measured with rocprofv3 --pmc
    (FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated) for this workload, from
    the committed profile summary, and where that figure comes from (file
    and the first 16 hex digits of its sha256); (None, None) if there is
    none."""
    import hashlib
    import re

    if storage != "fp32":
        return None, None
    for path in sorted(PROFILES.glob(f"*_step_{workload}.md")):
        text = path.read_text()
        m = re.search(r"^traffic_bytes_per_step:\s*(\d+)", text, re.M)
        if m:
            src = {"file": str(path.relative_to(ROOT)), "sha256_16": hashlib.sha256(text.encode()).hexdigest()[:16],
                   "measured": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, not in this run"}
            return int(m.group(1)), src
    return None, None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_device(env: dict, steps: int, warmup: int):
    import torch

    from lbm_config import parse_env
    from lbm_solver import Solver

    cfg = parse_env(env)
    solver = Solver(config=cfg)
    solver.initialize()
    case = solver.case
    if warmup:
        case.step(warmup)
    torch.cuda.synchronize()
    case.sync()
    t0 = time.perf_counter()
    case.step(steps)
    case.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the per-kernel HIP events are recorded in a run of their own: an event
    # between two launches adds a few microseconds to the step
    case.set_option("timing", 1)
    case.step(max(5, min(steps, 50)))
    ms_step_ev, ms_kernel = case.last_timing()
    case.set_option("timing", 0)
    cells = cfg.Nx * cfg.Ny * cfg.Nz
    out = {
        "cells": cells,
        "ms_per_step": 1e3 * wall / steps,
        "mlups": cells * steps / wall / 1e6,
        "ms_kernel": ms_kernel,
        "ms_step_events": ms_step_ev,
        "path": case.path(),
        "storage": cfg.storage,
        "schedule": (("tile-resident" if case.get_option("resident") == 1 else
                      "wavefront" if cfg.storage == "fp32" else "three launches") if case.path() == 2 else "staged"),
    }
    case.close()
    return out


def cpu_baseline(env: dict, budget_s: float = 12.0):
    """Oracle (the C restatement, OpenMP over cells) on a bounded sample: as
    many full steps of the workload as fit in the budget (at least one),
    with all its threads (the host's cores, or OMP_NUM_THREADS) and with one
    thread.  The results do not depend on the thread count."""
    from harness import setup_case
    from lbm_config import parse_env
    from oracle_ref import OracleCase, set_threads

    cfg = parse_env(env)
    cells = cfg.Nx * cfg.Ny * cfg.Nz

    def timed(threads):
        used = set_threads(threads)
        o = OracleCase(cfg.to_struct())
        setup_case(o, cfg)
        n = 0
        t0 = time.perf_counter()
        while True:
            o.step(1)
            n += 1
            if time.perf_counter() - t0 > budget_s:
                break
        wall = time.perf_counter() - t0
        o.close()
        return used, n, cells * n / wall / 1e6

    threads, n_all, mlups_all = timed(0)
    _, n_one, mlups_one = timed(1)
    set_threads(0)
    return {
        "value": round(mlups_all, 3),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_all} full step(s) of the {cfg.Nx}x{cfg.Ny}x{cfg.Nz} workload on the CPU oracle (C, -O3 "
                  f"-march=x86-64-v3, OpenMP, {threads} threads)",
        "cpu_model": cpu_model(),
        "single_thread": {"value": round(mlups_one, 3), "unit": "MLUPS", "sample": f"{n_one} full step(s), 1 thread"},
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
    for c, (env, steps) in enumerate(stream):
        s = Solver(config=parse_env(env))
        s.initialize()
        s.step(steps)
        s.case.download(1)
        s.close()
        if (c + 1) % 32 == 0:
            progress(f"serving {storage}: {c + 1}/{n_cases} sequential cases done")
    seq_wall = time.perf_counter() - t0
    budget = 200 << 30
    errors = []

    def note_errors(srv):
        for r in srv.finished:
            if r.status == "failed" and len(errors) < 4:
                errors.append(r.error[:300])

    def run_burst(graphs=False, batch=False):
        srv = EnsembleServer(budget, max_active=64, chunk_steps=16, n_workers=4, graphs=graphs, batch=batch)
        t0 = time.perf_counter()
        for env, steps in stream:
            srv.submit(env, steps)
        srv.run_until_idle()
        st = srv.stats(t0)
        note_errors(srv)
        srv.close()
        return st

    burst_streams = run_burst()
    progress(f"serving {storage}: burst of {n_cases} on streams done in {burst_streams['wall_s']:.1f} s")
    burst_graphs = run_burst(graphs=True)
    progress(f"serving {storage}: burst with graphs done in {burst_graphs['wall_s']:.1f} s")
    burst_batch = run_burst(batch=True)
    progress(f"serving {storage}: burst with the merged launch grid done in {burst_batch['wall_s']:.1f} s")
    # the latency run uses the faster of streams and the merged grid
    batch = burst_batch["aggregate_mlups"] > burst_streams["aggregate_mlups"]
    burst = burst_batch if batch else burst_streams
    capacity = n_cases / burst["wall_s"]
    rate = 0.8 * capacity
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_cases))
    srv = EnsembleServer(budget, max_active=64, chunk_steps=16, n_workers=4, batch=batch)
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
    note_errors(srv)
    srv.close()
    progress(f"serving {storage}: Poisson stream done")
    return {
        "cases": n_cases,
        "storage": storage,
        "lattice_updates": lups,
        "sequential_mlups": round(lups / seq_wall / 1e6, 1),
        "mode": "merged launch grid" if batch else "streams",
        "ensemble_mlups": round(burst["aggregate_mlups"], 1),
        "ensemble_mlups_streams": round(burst_streams["aggregate_mlups"], 1),
        "ensemble_mlups_graphs": round(burst_graphs["aggregate_mlups"], 1),
        "ensemble_mlups_merged_grid": round(burst_batch["aggregate_mlups"], 1),
        "speedup_vs_sequential": round(burst["aggregate_mlups"] / (lups / seq_wall / 1e6), 3),
        "poisson_rate_per_s": round(rate, 3),
        "poisson_mlups": round(poisson["aggregate_mlups"], 1),
        "latency_p50_s": round(poisson["latency_p50_s"], 4),
        "latency_p99_s": round(poisson["latency_p99_s"], 4),
        "peak_active": max(burst["peak_active"], poisson["peak_active"]),
        "failed": burst_streams["failed"] + burst_graphs["failed"] + burst_batch["failed"] + poisson["failed"],
        "failure_messages": errors,
    }


WORKLOADS = (
    # label, preset, overrides, algorithmic bytes per lattice update
    ("c1_plain_64x32x32_fp32", "plain_64x32x32", {"LBM_STORAGE": "fp32"}, 152),
    ("c3_plain_512x256x192_fp32", "plain_512x256x192", {"LBM_STORAGE": "fp32"}, 152),
    ("c3_plain_512x256x192_fp16", "plain_512x256x192", {"LBM_STORAGE": "fp16"}, 76),
    ("c3_plain_512x256x192_bf16", "plain_512x256x192", {"LBM_STORAGE": "bf16"}, 76),
    ("c4_vof_192x96x160_fp32", "vof_192x96x160", {}, 170),
    ("c4_vof_192x96x160_fp16", "vof_192x96x160", {"LBM_STORAGE": "fp16"}, 94),
)


def workloads(steps: int):
    """The other benchmark configurations: MLUPS and the fraction of the HBM
    roofline at their algorithmic bytes per update (VOF: populations +
    mass, fill and type)."""
    from lbm_config import preset_env

    out = {}
    for label, name, extra, bpl in WORKLOADS:
        # (20 warmup steps: the fused step's schedule is chosen by timing on
        # the first call of at least 18 steps)
        r = run_device(dict(preset_env(name), **extra), steps, 20)
        progress(f"workload {label}: {r['mlups']:.1f} MLUPS")
        out[label] = {
            "mlups": round(r["mlups"], 1),
            "kernel_ms": round(r["ms_kernel"], 5),
            "hbm_frac": round(r["cells"] * bpl / (r["ms_kernel"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "path": "fused" if r["path"] == 2 else "staged",
            "schedule": r["schedule"],
        }
    ref = out["c3_plain_512x256x192_fp32"]
    for st in ("fp16", "bf16"):
        r = out[f"c3_plain_512x256x192_{st}"]
        # whole-step MLUPS ratio (the metric), and the step-kernel time ratio beside it
        out[f"c3_{st}_speedup_vs_fp32"] = round(r["mlups"] / ref["mlups"], 3)
        out[f"c3_{st}_kernel_speedup_vs_fp32"] = round(ref["kernel_ms"] / r["kernel_ms"], 3)
    c4, c4h = out["c4_vof_192x96x160_fp32"], out["c4_vof_192x96x160_fp16"]
    out["c4_fp16_speedup_vs_fp32"] = round(c4h["mlups"] / c4["mlups"], 3)
    return out


def graph_launch(chunks: int = 100, chunk: int = 16):
    """One small case alone, `chunks` calls of `chunk` steps: direct launches
    against each call captured into a HIP graph (MLUPS from the wall clock,
    the same case and steps)."""
    from lbm_config import parse_env, preset_env
    from lbm_solver import Solver

    cases = {
        "plain_64x32x32": preset_env("plain_64x32x32"),
        "waves_32x32x32": dict(preset_env("waves_256x128x128"), LBM_LX_M="32", LBM_LY_M="32", LBM_LZ_M="32",
                               LBM_BED_AMP_M="2", LBM_BED_WAVELENGTH_M="8"),
        "waves_128x64x64": dict(preset_env("waves_256x128x128"), LBM_LX_M="128", LBM_LY_M="64", LBM_LZ_M="64",
                                LBM_BED_AMP_M="3", LBM_BED_WAVELENGTH_M="32"),
    }
    out = {"chunk_steps": chunk, "chunks": chunks}
    for label, env in cases.items():
        cfg = parse_env(env)
        r = {}
        for mode in (0, 1):
            solver = Solver(config=cfg)
            solver.initialize()
            case = solver.case
            case.set_option("graph", mode)
            case.step(2 * chunk)
            case.sync()
            t0 = time.perf_counter()
            for _ in range(chunks):
                case.step(chunk)
            case.sync()
            wall = time.perf_counter() - t0
            r["graph_mlups" if mode else "direct_mlups"] = round(cfg.Nx * cfg.Ny * cfg.Nz * chunks * chunk / wall / 1e6, 1)
            solver.close()
        r["graph_speedup"] = round(r["graph_mlups"] / r["direct_mlups"], 3)
        out[label] = r
        progress(f"graph launch {label}: {r}")
    return out


def accuracy():
    """Reduced storage against the FP64 run of the same case, three cases:
      * plain channel physics at 256x128x128 and the air-over-waves case
        (256x128x128), seed 12345, after 2000 steps: max |u - u64| /
        max |u64|, max |dp| with dp = (rho - rho64)/3, and the relative drift
        of the total mass;
      * the VOF box 192x96x160 after 1500 steps: water volume error, the
        fraction of cells of another type, and over the cells liquid in both
        runs the velocity error (max and rms, relative to the FP64 field's)
        and the pressure perturbation error in units of the pool's
        hydrostatic head rho0 g H."""
    import numpy as np

    import lbm_abi as abi
    from lbm_config import parse_env, preset_env
    from lbm_solver import Solver

    def air_case(env, steps):
        def run(storage):
            solver = Solver(config=parse_env(dict(env, LBM_STORAGE=storage)))
            solver.initialize()
            case = solver.case
            m0 = case.reduce(abi.REDUCE_MASS)
            case.step(steps)
            u = np.stack([case.download(x).astype(np.float64) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)])
            rho = case.download(abi.FIELD_RHO).astype(np.float64)
            m1 = case.reduce(abi.REDUCE_MASS)
            case.close()
            return u, rho, m0, m1

        u64, r64, m0, m1 = run("fp64")
        out = {"steps": steps, "fp64_mass_drift": abs(m1 - m0) / m0}
        umax = float(np.max(np.linalg.norm(u64, axis=0)))
        for st in ("fp32", "fp16", "bf16"):
            u, r, a0, a1 = run(st)
            out[st] = {
                "velocity_rel_error": float(np.max(np.linalg.norm(u - u64, axis=0)) / umax),
                "pressure_abs_error": float(np.max(np.abs(r - r64)) / 3.0),
                "mass_drift": abs(a1 - a0) / a0,
            }
        return out

    def water_case(env, steps):
        cfg = parse_env(env)
        p_scale = float(cfg.g_lb) * float(cfg.free_surface_base_m) / float(cfg.dx)

        def run(storage):
            solver = Solver(config=parse_env(dict(env, LBM_STORAGE=storage)))
            solver.initialize()
            case = solver.case
            case.step(steps)
            r = {"type": case.download(abi.FIELD_FS_TYPE), "rho": case.download(abi.FIELD_RHO).astype(np.float64),
                 "u": np.stack([case.download(x).astype(np.float64) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]),
                 "volume": case.reduce(abi.REDUCE_VOF_VOLUME)}
            case.close()
            return r

        ref = run("fp64")
        rms = lambda x: float(np.sqrt(np.mean(np.square(x))))
        out = {"steps": steps, "pressure_scale": p_scale}
        for st in ("fp32", "fp16", "bf16"):
            r = run(st)
            liq = (ref["type"] == abi.FS_LIQUID) & (r["type"] == abi.FS_LIQUID)
            un = np.linalg.norm(ref["u"], axis=0)[liq]
            dun = np.linalg.norm(r["u"] - ref["u"], axis=0)[liq]
            dp = ((r["rho"] - ref["rho"]) / 3.0)[liq]
            out[st] = {
                "volume_rel_error": abs(r["volume"] - ref["volume"]) / ref["volume"],
                "type_mismatch_frac": float(np.mean(r["type"] != ref["type"])),
                "velocity_rel_error_max": float(np.max(dun) / np.max(un)),
                "velocity_rel_error_rms": rms(dun) / rms(un),
                "pressure_error_max_over_head": float(np.max(np.abs(dp))) / p_scale,
            }
        return out

    out = {}
    out["plain_256x128x128"] = air_case(dict(preset_env("plain_64x32x32"), LBM_LX_M="256", LBM_LY_M="128",
                                             LBM_LZ_M="128"), 2000)
    progress("accuracy: plain channel done")
    out["waves_256x128x128"] = air_case(preset_env("waves_256x128x128"), 2000)
    progress("accuracy: waves done")
    out["vof_192x96x160"] = water_case(preset_env("vof_192x96x160"), 1500)
    progress("accuracy: vof done")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="waves_256x128x128")
    ap.add_argument("--storage", default=None, choices=("fp32", "fp16", "bf16", "fp64"),
                    help="population storage of the main run (default: the workload's)")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-cases", type=int, default=256, help="requests per precision of the serving stream (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (3 launches): the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
            "(several GPUs are not implemented)",
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
    if args.storage:
        env["LBM_STORAGE"] = args.storage
    cfg = parse_env(env)
    # (a warmup below 18 steps leaves the fused step's choice of schedule,
    # timed on the first call of at least 18 steps, to the timed call)
    main_run = run_device(env, args.steps, args.warmup)
    progress(f"main run {args.workload}: {main_run['mlups']:.1f} MLUPS")
    bpl = BYTES_PER_LU[cfg.storage]
    if cfg.water_free_surface_enabled and cfg.free_surface_tracking_mode == "vof":
        bpl = VOF_BYTES_PER_LU[cfg.storage]  # populations + mass, fill and type
    achieved = main_run["cells"] * bpl / (main_run["ms_kernel"] * 1e-3) / 1e9
    achieved_wall = main_run["cells"] * bpl / (main_run["ms_per_step"] * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.workload, cfg.storage)
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
        "dtype": "f64" if cfg.storage == "fp64" else "f32",  # arithmetic type (16-bit storage computes in FP32)
        "data": "synthetic (seeded log-law inflow + 2% uniform noise, seed 12345)",
        "config": {
            "workload": args.workload,
            "grid": [cfg.Nx, cfg.Ny, cfg.Nz],
            "collision": cfg.collision_model,
            "les": bool(cfg.use_LES),
            "storage": cfg.storage,
            "path": "fused" if main_run["path"] == 2 else "staged",
            "schedule": main_run["schedule"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel_ms": round(main_run["ms_kernel"], 5),
            "bytes_per_lu": bpl,
            "scope": "step kernels only (HIP events on the library stream around each step's launches; "
                     "the 2-D bed kernels and launch gaps excluded)",
            "achieved_wall": round(achieved_wall, 1),
            "frac_wall": round(achieved_wall / HBM_PEAK_GBS, 4),
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
            r = run_device(dict(env, **extra), vsteps, 20)
            b = BYTES_PER_LU[parse_env(dict(env, **extra)).storage]
            variants[label] = {
                "mlups": round(r["mlups"], 1),
                "hbm_frac": round(r["cells"] * b / (r["ms_kernel"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            }
        result["variants"] = variants
    if not args.no_variants:
        result["workloads"] = workloads(max(20, args.steps // 4))
        result["graph_launch"] = graph_launch()
        result["accuracy"] = accuracy()
    if args.serving_cases > 0 and not args.no_variants:
        result["serving"] = {st: serving(args.serving_cases, st) for st in ("fp32", "fp16")}
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(env)
        progress("cpu baseline done")
    print(json.dumps(result))


if __name__ == "__main__":
    main()
