"""Throughput of the CA step on one MI355X.

Workload: directional solidification with multiple grains on a
256 x 256 x 512 block (dx = 1 um), frozen linear temperature field
T = T_liq - 5 K + G (z - V t) with G = 1e4 K/m and V = 1e-3 m/s, surface
nucleation 1e10 m^-2 (mean 1 K, sigma 0.2 K) on every padded face, bulk
nucleation 1e13 m^-3 (mean 5 K, sigma 0.5 K), random orientations from seed
1, c0 = 4 wt%, k = 0.14, m = -2.6 K/wt%, T_liq = 933 K, D_l = 3e-9, D_s = 1e-12
m^2/s, dt = 0.9 dx^2 / (6 D_l) = 5e-5 s, FSGrow 0.66, plain FP64.

A "step" is one doiteration() (timestep++, temperature, nucleation, grain
activation, diffusion, partitioning, growth, interface list, capture) over
the whole block, with the state resident in HBM.  At this dt the reference
CPU build stops at the first negative ce (growth.c:163-165, step ~25); the
timed run uses the reference GPU build's behaviour, which has no such check.

Prints one JSON line on stdout.  --gpus N with N > 1 is refused: this build
runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mumatscale_amd import abi  # noqa: E402
from mumatscale_amd.ctrl import Config  # noqa: E402

METRIC = "Mcell-updates/s & % HBM roofline; batched/FP32 vs plain FP64 at stated error"
HBM_PEAK_GBS = 8000.0
B_ALG_FP64 = 56.0      # {ce, cl, fs, gr} read + written per cell-update


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def directional_config(nx=256, ny=256, nz=512):
    c = Config()
    c.nsub = [1, 1, 1]
    c.sbdim = [nx, ny, nz]
    c.pad_bndy = [1, 1, 1]
    c.cell_size = 1.0e-6
    c.fsgrow = 0.66
    c.dt = 0.9 * c.cell_size ** 2 / (6 * 3.0e-9)
    c.random_seed = 1
    c.liquidus_temp = 933.0
    c.solidus_temp = 850.0
    c.c_init = 4.0
    c.m_solute = [0.0, 0.0, -2.6]
    c.part_coef = 0.14
    c.d_sol = 1.0e-12
    c.d_liq = 3.0e-9
    c.t_init = 933.0 - 5.0
    c.gradient = 1.0e4
    c.velocity = 1.0e-3
    c.bulk_density = 1.0e13
    c.bulk_mean = 5.0
    c.bulk_sigma = 0.5
    c.surf_density = 1.0e10
    c.surf_mean = 1.0
    c.surf_sigma = 0.2
    c.oriented = 1
    c.max_total_grains = 1000000
    return c


def ensemble_configs(n=96, seed=3):
    """96 independent domains: sides drawn from {32,48,64,96,128} per axis,
    cooling rate log-uniform in [1, 1000] K/s, initial undercooling uniform
    in [1, 5] K, bulk nucleation 1e14 m^-3 (mean 3 K, sigma 0.5 K) with a
    distinct seed per domain; alloy and dt of the single-grain physics."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = rng.choice([32, 48, 64, 96, 128], size=(n, 3))
    rates = 10.0 ** rng.uniform(0.0, 3.0, size=n)
    under = rng.uniform(1.0, 5.0, size=n)
    out = []
    for i in range(n):
        c = directional_config(*(int(v) for v in sides[i]))
        c.gradient = 0.0
        c.velocity = 0.0
        c.cool_rate = float(rates[i])
        c.t_init = 933.0 - float(under[i])
        c.surf_density = 0.0
        c.bulk_density = 1.0e14
        c.bulk_mean = 3.0
        c.bulk_sigma = 0.5
        c.random_seed = 1000 + i
        c.max_total_grains = 100000
        out.append(c)
    return out


def run_ensemble(steps, baseline_steps):
    """Batched FP32-deviation step of the ensemble against the same domains
    stepped one at a time in plain FP64."""
    from mumatscale_amd.lib import Batch, Subblock, autotune_batch
    cfgs = ensemble_configs()
    cells = sum(c.sbdim[0] * c.sbdim[1] * c.sbdim[2] for c in cfgs)
    p32 = [c.params(check_negative_ce=False, precision=abi.PREC_FP32DEV) for c in cfgs]
    nuc = [c.nucleation() for c in cfgs]
    t0 = time.time()
    planes, rows, tms = autotune_batch(p32, nuc, warmup=10, trial_steps=20)
    log("ensemble: autotuner picked planes=%d rows=%d (%.3f ms) in %.1f s" % (planes, rows, tms, time.time() - t0))
    with Batch(p32, nuc) as bt:
        bt.set_tuning(abi.TUNE_ZM_PLANES, planes)
        bt.set_tuning(abi.TUNE_ZM_ROWS, rows)
        bt.set_tuning(abi.TUNE_GRAPH, 1)
        bt.step(10)
        t0 = time.perf_counter()
        reps = bt.step(steps, want_sums=True)
        wall = time.perf_counter() - t0
        drift = max(abs(r.solute_sum + r.solute_ledger - (c.sbdim[0] * c.sbdim[1] * c.sbdim[2]) * c.c_init)
                    / ((c.sbdim[0] * c.sbdim[1] * c.sbdim[2]) * c.c_init) for r, c in zip(reps, cfgs))
    batched = cells * steps / wall / 1e6
    # the same domains one at a time, plain FP64
    t_one = 0.0
    for c in cfgs:
        with Subblock(c.params(check_negative_ce=False), c.nucleation()) as sb:
            sb.step(2)
            t0 = time.perf_counter()
            sb.step(baseline_steps)
            t_one += time.perf_counter() - t0
    one = cells * baseline_steps / t_one / 1e6
    return {"domains": len(cfgs), "cells": cells, "steps": steps, "dtype": "f32-deviation storage, f64 arithmetic",
            "value": round(batched, 1), "unit": "Mcell-updates/s", "ms_per_step": round(wall / steps * 1e3, 4),
            "tile": {"zm_planes": planes, "zm_rows": rows}, "hip_graph": True,
            "one_at_a_time_fp64": round(one, 1), "one_at_a_time_steps": baseline_steps,
            "speedup_vs_one_at_a_time_fp64": round(batched / one, 2),
            "max_solute_drift_with_ledger": drift}


def run_fp32_variant(cfg, steps, warmup, zm, rows, xcd):
    """The headline workload with FP32-deviation storage; errors against the
    FP64 run after the same steps."""
    import numpy as np
    from mumatscale_amd.lib import Subblock
    n = cfg.sbdim[0] * cfg.sbdim[1] * cfg.sbdim[2]
    res = {}
    fields = {}
    for prec in (abi.PREC_FP32DEV, abi.PREC_FP64):
        with Subblock(cfg.params(check_negative_ce=False, precision=prec), cfg.nucleation()) as sb:
            sb.set_tuning(abi.TUNE_ZM_PLANES, zm)
            sb.set_tuning(abi.TUNE_ZM_ROWS, rows)
            sb.set_tuning(abi.TUNE_XCD_REMAP, xcd)
            sb.step(warmup)
            sb.set_kernel_timing(prec == abi.PREC_FP32DEV)
            t0 = time.perf_counter()
            rep = sb.step(steps, want_sums=True)
            wall = time.perf_counter() - t0
            k_ms, _ = sb.kernel_time_ms()
            fields[prec] = (sb.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1], rep)
            if prec == abi.PREC_FP32DEV:
                res = {"value": round(n * steps / wall / 1e6, 1), "ms_per_step": round(wall / steps * 1e3, 4),
                       "k_solute_ms": round(k_ms, 4),
                       "roofline_frac": round(n * 40.0 / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "bytes_per_cell": 40.0}
    ce32, r32 = fields[abi.PREC_FP32DEV]
    ce64, r64 = fields[abi.PREC_FP64]
    res["error_vs_fp64"] = {
        "solute_l2_rel": float(np.linalg.norm(ce32 - ce64) / np.linalg.norm(ce64)),
        "solid_fraction_abs": abs(r32.solid_sum - r64.solid_sum) / n,
        "grains_fp32": r32.num_grains - 1, "grains_fp64": r64.num_grains - 1,
        "solute_drift_stored": (r32.solute_sum - n * cfg.c_init) / (n * cfg.c_init),
        "solute_drift_with_ledger": (r32.solute_sum + r32.solute_ledger - n * cfg.c_init) / (n * cfg.c_init)}
    return res


def cpu_baseline(cfg, target_s):
    """Oracle (scalar C port of the reference path), one thread, bounded sample."""
    from oracle.oracle import OracleSubblock
    t0 = time.time()
    orc = OracleSubblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    log("cpu baseline: oracle set-up %.1f s" % (time.time() - t0))
    n = cfg.sbdim[0] * cfg.sbdim[1] * cfg.sbdim[2]
    t0 = time.time()
    orc.step(1)
    one = time.time() - t0
    steps = max(1, min(50, int(target_s / max(one, 1e-3))))
    t0 = time.time()
    orc.step(steps)
    dt = time.time() - t0
    orc.close()
    return {"value": round(n * steps / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": "same workload at full size, steps 2-%d after set-up (%.1f s), one thread"
                      % (steps + 1, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=str, default="256x256x512")
    ap.add_argument("--zm", type=int, default=64,
                    help="z-marching solute kernel, planes per workgroup (0: cell-parallel)")
    ap.add_argument("--zm-rows", type=int, default=16, help="rows per z-marching workgroup (4, 8, 16)")
    ap.add_argument("--xcd", type=int, default=1, help="1: contiguous tile runs per XCD")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the FP32-deviation and batched-ensemble measurements")
    ap.add_argument("--extra-steps", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the CA step on one GPU; --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    nx, ny, nz = (int(v) for v in args.size.split("x"))
    cfg = directional_config(nx, ny, nz)

    from mumatscale_amd.lib import Subblock
    t0 = time.time()
    gpu = Subblock(cfg.params(check_negative_ce=False), cfg.nucleation(), device=0)
    gpu.set_tuning(abi.TUNE_ZM_PLANES, args.zm)
    gpu.set_tuning(abi.TUNE_ZM_ROWS, args.zm_rows)
    gpu.set_tuning(abi.TUNE_XCD_REMAP, args.xcd)
    log("gpu set-up %.1f s" % (time.time() - t0))
    n = nx * ny * nz
    if args.warmup:
        gpu.step(args.warmup)
    gpu.set_kernel_timing(True)
    t0 = time.perf_counter()
    rep = gpu.step(args.steps)        # synchronises on return
    t1 = time.perf_counter()
    gpu.set_kernel_timing(False)
    k_ms, k_n = gpu.kernel_time_ms()
    wall = t1 - t0
    value = n * args.steps / wall / 1e6
    log("gpu: %d steps in %.3f s, grains %d, gindex %d, k_solute avg %.4f ms over %d launches"
        % (args.steps, wall, rep.num_grains, rep.gindex, k_ms, k_n))
    achieved = n * B_ALG_FP64 / (k_ms * 1e-3) / 1e9 if k_ms > 0 else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": None,
            "kernel": ("k_solute_zm<%d>" % args.zm_rows) if args.zm > 0 else "k_solute<true,true>",
            "bytes_per_cell": B_ALG_FP64}
    gpu.close()
    extra = {}
    if not args.no_extra:
        extra["fp32dev"] = run_fp32_variant(cfg, args.extra_steps, args.warmup, args.zm, args.zm_rows, args.xcd)
        log("fp32dev:", extra["fp32dev"])
        extra["batched_ensemble"] = run_ensemble(3000, 200)
        log("ensemble:", extra["batched_ensemble"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, args.cpu_seconds)
        log("cpu baseline:", cpu)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "directional solidification %dx%dx%d, plain FP64" % (nx, ny, nz),
                   "cells": n, "dt": cfg.dt, "zm_planes": args.zm, "zm_rows": args.zm_rows,
                   "xcd_remap": args.xcd, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
