"""Throughput of the CA step on one MI355X.

Workload: directional solidification with multiple grains on a
256 x 256 x 512 block (dx = 1 um), frozen linear temperature field
T = T_liq - 5 K + G (z - V t) with G = 1e4 K/m and V = 1e-3 m/s, surface
nucleation 1e10 m^-2 (mean 1 K, sigma 0.2 K) on every padded face, bulk
nucleation 1e13 m^-3 (mean 5 K, sigma 0.5 K), random orientations from seed
1, c0 = 4 wt%, k = 0.14, m = -2.6 K/wt%, T_liq = 933 K, D_l = 3e-9, D_s = 1e-12
m^2/s, FSGrow 0.66, plain FP64, dt = 0.45 dx^2 / (6 D_l) = 2.5e-5 s: half
the explicit diffusion limit.  At the limit itself (5e-5 s) the partitioning
update is unstable with k = 0.14 and the reference CPU build stops at the
first negative ce (growth.c:163-165, step ~25); the headline runs the stable
dt with that check on, and the bench line reports the literal-limit run (the
step at which ce turns negative, and its throughput without the check) as
`literal_dt`.

A "step" is one doiteration() (timestep++, temperature, nucleation, grain
activation, diffusion, partitioning, growth, interface list, capture) over
the whole block, with the state resident in HBM.

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
from mumatscale_amd.workloads import directional_config, ensemble_configs  # noqa: E402,F401

METRIC = "Mcell-updates/s & % HBM roofline; batched/FP32 vs plain FP64 at stated error"
HBM_PEAK_GBS = 8000.0
B_ALG_FP64 = 56.0      # {ce, cl, fs, gr} read + written per cell-update
B_ALG_FP32 = 40.0      # ce, cl as FP32 deviations; fs FP64; gr int32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILES = {"fp64": os.path.join(ROOT, "profiles", "k_solute_zl_directional_256x256x512_fp64.md"),
                    "fp32": os.path.join(ROOT, "profiles", "k_solute_zl_directional_256x256x512_fp32dev.md")}


def bandwidth_probe(cells, reps=20):
    """Achievable HBM bandwidth on this GPU for the solute kernel's field
    mixes (mmca_bandwidth_probe: streaming kernels with no arithmetic), at
    `cells` cells per launch: GB/s of each mix and its bytes per cell."""
    import ctypes as C
    from mumatscale_amd.lib import lib
    out = {}
    for mix, name in ((0, "copy_f64"), (1, "solute_mix_fp64"), (2, "solute_mix_fp32dev")):
        ms, bpc = C.c_double(0), C.c_double(0)
        st = lib().mmca_bandwidth_probe(0, mix, cells, reps, C.byref(ms), C.byref(bpc))
        if st != abi.OK:
            return None
        out[name] = {"GB/s": round(cells * bpc.value / (ms.value * 1e-3) / 1e9, 1),
                     "bytes_per_cell": bpc.value, "ms": round(ms.value, 4)}
    out["cells"] = cells
    return out


def apply_shape(sb, shape):
    """Solute-kernel tile shape (planes, rows, XCD remap) of a subblock."""
    sb.set_tuning(abi.TUNE_ZM_PLANES, shape["zm"])
    sb.set_tuning(abi.TUNE_ZM_ROWS, shape["rows"])
    sb.set_tuning(abi.TUNE_XCD_REMAP, shape["xcd"])
    sb.set_tuning(abi.TUNE_ZM_LDS, shape.get("lds", 0))
    sb.set_tuning(abi.TUNE_CAPTURE_CHUNK, shape.get("chunk", 0))
    sb.set_tuning(abi.TUNE_CAPTURE_BAND, shape.get("band", 0))


def measured_traffic(size, shape, prec="fp64"):
    """HBM bytes per launch of the solute kernel from the committed
    FETCH_SIZE/WRITE_SIZE summary, when this run has the shape it was
    measured on (None otherwise)."""
    import re
    path = TRAFFIC_PROFILES[prec]
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        text = fh.read()
    rec = re.search(r"shape = (\S+) zm_planes (\d+) zm_rows (\d+) xcd_remap (\d+)(?: zm_lds (\d+))?", text)
    m = re.search(r"traffic_bytes_per_launch = ([0-9.eE+]+)", text)
    want = (size, str(shape["zm"]), str(shape["rows"]), str(shape["xcd"]), str(shape.get("lds", 0)))
    if not rec or not m or rec.groups()[:4] + (rec.group(5) or "0",) != want:
        return None
    return float(m.group(1))


def run_ensemble(steps, baseline_steps):
    """Batched FP32-deviation step of the 96-domain ensemble (graph replay,
    autotuned tile) against the same domains stepped one at a time in plain
    FP64 (throughput), and against an FP64 batch of the same steps (grain
    counts and solute balance per domain).  dt is half the explicit limit of
    the single-grain physics, where the partitioning update is stable."""
    from mumatscale_amd.lib import Batch, Subblock, autotune_batch
    cfgs = ensemble_configs()
    for c in cfgs:
        c.dt *= 0.5
    cells = sum(c.sbdim[0] * c.sbdim[1] * c.sbdim[2] for c in cfgs)
    p32 = [c.params(check_negative_ce=False, precision=abi.PREC_FP32DEV) for c in cfgs]
    p64 = [c.params(check_negative_ce=False) for c in cfgs]
    nuc = [c.nucleation() for c in cfgs]
    t0 = time.time()
    planes, rows, tms = autotune_batch(p32, nuc, warmup=10, trial_steps=20)
    log("ensemble: autotuner picked planes=%d rows=%d (%.3f ms) in %.1f s" % (planes, rows, tms, time.time() - t0))
    reps = {}
    walls = {}
    kms = {}
    for prec, params in (("fp32", p32), ("fp64", p64)):
        with Batch(params, nuc) as bt:
            bt.set_tuning(abi.TUNE_ZM_PLANES, planes)
            bt.set_tuning(abi.TUNE_ZM_ROWS, rows)
            bt.set_tuning(abi.TUNE_GRAPH, 1)
            bt.set_kernel_timing(True)
            t0 = time.perf_counter()
            reps[prec] = bt.step(steps, want_sums=True)
            walls[prec] = time.perf_counter() - t0
            kms[prec], _ = bt.kernel_time_ms()
    ncell = [c.sbdim[0] * c.sbdim[1] * c.sbdim[2] for c in cfgs]
    drift = max(abs(r.solute_sum + r.solute_ledger - n * c.c_init) / (n * c.c_init)
                for r, c, n in zip(reps["fp32"], cfgs, ncell))
    grains_equal = sum(a.num_grains == b.num_grains for a, b in zip(reps["fp32"], reps["fp64"]))
    fs_err = max(abs(a.solid_sum - b.solid_sum) / n for a, b, n in zip(reps["fp32"], reps["fp64"], ncell))
    batched = cells * steps / walls["fp32"] / 1e6
    # the same domains one at a time, plain FP64
    t_one = 0.0
    for c in cfgs:
        with Subblock(c.params(check_negative_ce=False), c.nucleation()) as sb:
            sb.step(2)
            t0 = time.perf_counter()
            sb.step(baseline_steps)
            t_one += time.perf_counter() - t0
    one = cells * baseline_steps / t_one / 1e6
    return {"domains": len(cfgs), "cells": cells, "steps": steps, "dt": cfgs[0].dt,
            "dtype": "f32-deviation storage, f64 arithmetic",
            "value": round(batched, 1), "unit": "Mcell-updates/s", "ms_per_step": round(walls["fp32"] / steps * 1e3, 4),
            "tile": {"zm_planes": planes, "zm_rows": rows}, "hip_graph": True,
            "k_solute_ms": round(kms["fp32"], 4),
            "roofline_frac": round(cells * B_ALG_FP32 / (kms["fp32"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            if kms["fp32"] > 0 else None,
            "step_over_kernel": round(walls["fp32"] / steps * 1e3 / kms["fp32"], 3) if kms["fp32"] > 0 else None,
            "one_at_a_time_fp64": round(one, 1), "one_at_a_time_steps": baseline_steps,
            "speedup_vs_one_at_a_time_fp64": round(batched / one, 2),
            "batched_fp64": round(cells * steps / walls["fp64"] / 1e6, 1),
            "max_solute_drift_with_ledger": drift,
            "domains_with_equal_grain_count_vs_fp64": grains_equal,
            "grains_fp32": sum(r.num_grains - 1 for r in reps["fp32"]),
            "grains_fp64": sum(r.num_grains - 1 for r in reps["fp64"]),
            "max_solid_fraction_err_vs_fp64": fs_err}


def run_fp32_variant(cfg, steps, warmup, shape):
    """The headline workload with FP32-deviation storage: throughput, the
    solute kernel's roofline fraction at 40 B per cell, and the error against
    the FP64 run of the same steps."""
    import numpy as np
    from mumatscale_amd.lib import Subblock
    n = cfg.sbdim[0] * cfg.sbdim[1] * cfg.sbdim[2]
    fields = {}
    res = {}
    for prec in (abi.PREC_FP32DEV, abi.PREC_FP64):
        with Subblock(cfg.params(check_negative_ce=True, precision=prec), cfg.nucleation()) as sb:
            apply_shape(sb, shape)
            sb.step(warmup)
            sb.set_kernel_timing(True)
            t0 = time.perf_counter()
            rep = sb.step(steps, want_sums=True)
            wall = time.perf_counter() - t0
            k_ms, _ = sb.kernel_time_ms()
            fields[prec] = (sb.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1], rep)
            if prec == abi.PREC_FP32DEV:
                res = {"value": round(n * steps / wall / 1e6, 1), "ms_per_step": round(wall / steps * 1e3, 4),
                       "k_solute_ms": round(k_ms, 4),
                       "roofline_frac": round(n * B_ALG_FP32 / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "bytes_per_cell": B_ALG_FP32, "steps": steps, "warmup": warmup}
    ce32, r32 = fields[abi.PREC_FP32DEV]
    ce64, r64 = fields[abi.PREC_FP64]
    total0 = n * cfg.c_init
    res["error_vs_fp64"] = {
        "dt": cfg.dt, "after_steps": warmup + steps,
        "solute_l2_rel": float(np.linalg.norm(ce32 - ce64) / np.linalg.norm(ce64)),
        "solid_fraction_abs": abs(r32.solid_sum - r64.solid_sum) / n,
        "grains_fp32": r32.num_grains - 1, "grains_fp64": r64.num_grains - 1,
        "solute_drift_stored": (r32.solute_sum - total0) / total0,
        "solute_drift_with_ledger": (r32.solute_sum + r32.solute_ledger - total0) / total0}
    return res


def run_literal_dt(dims, steps, warmup, shape):
    """The headline workload at the explicit limit dt = 0.9 dx^2 / (6 D_l)
    itself: the step at which ce first turns negative (where the reference
    CPU build stops, growth.c:163-165), and the throughput of the same steps
    run on without that check (the reference GPU build's behaviour)."""
    from mumatscale_amd.lib import Subblock
    cfg = directional_config(*dims)
    n = dims[0] * dims[1] * dims[2]
    with Subblock(cfg.params(check_negative_ce=True), cfg.nucleation()) as sb:
        apply_shape(sb, shape)
        rep = sb.step(warmup + steps, raise_on_error=False)
        failed = rep.failed_step if rep.status != abi.OK else None
    with Subblock(cfg.params(check_negative_ce=False), cfg.nucleation()) as sb:
        apply_shape(sb, shape)
        sb.step(warmup)
        t0 = time.perf_counter()
        sb.step(steps)
        wall = time.perf_counter() - t0
    return {"dt": cfg.dt, "negative_ce_at_step": failed,
            "value_without_check": round(n * steps / wall / 1e6, 1), "steps": steps, "warmup": warmup}


def run_coupled(dims, steps, warmup, shape, nsub=(2, 2, 2)):
    """The headline domain run as the reference's MPI ranks on this GPU (a
    coupled batch: face-halo exchanges and cross-rank grain ids each step),
    against the same domain as one block over the same steps.  The ranks
    sample their own sites and streams, so the two runs are different
    simulations of the same size."""
    from mumatscale_amd.lib import Batch, Subblock
    nx, ny, nz = dims
    n = nx * ny * nz
    one = directional_config(nx, ny, nz)
    with Subblock(one.params(check_negative_ce=False), one.nucleation()) as sb:
        apply_shape(sb, shape)
        sb.step(warmup)
        t0 = time.perf_counter()
        sb.step(steps)
        wall_one = time.perf_counter() - t0
    cfg = directional_config(nx, ny, nz)
    cfg.nsub = list(nsub)
    cfg.sbdim = [nx // nsub[0], ny // nsub[1], nz // nsub[2]]
    nr = nsub[0] * nsub[1] * nsub[2]
    coords = [(r % nsub[0], r // nsub[0] % nsub[1], r // (nsub[0] * nsub[1])) for r in range(nr)]
    params = [cfg.params(check_negative_ce=False, whole_domain=False, sb_coords=coords[r]) for r in range(nr)]
    nucs = [cfg.nucleation(rank=r) for r in range(nr)]
    with Batch(params, nucs) as bt:
        bt.couple(nsub)
        bt.set_tuning(abi.TUNE_ZM_PLANES, shape["zm"])
        bt.set_tuning(abi.TUNE_ZM_ROWS, shape["rows"])
        bt.set_tuning(abi.TUNE_XCD_REMAP, shape["xcd"])
        bt.step(warmup)
        t0 = time.perf_counter()
        reps = bt.step(steps)
        wall = time.perf_counter() - t0
    return {"ranks": nr, "nsub": list(nsub), "subblock": cfg.sbdim, "steps": steps,
            "value": round(n * steps / wall / 1e6, 1), "ms_per_step": round(wall / steps * 1e3, 4),
            "one_block_value": round(n * steps / wall_one / 1e6, 1),
            "one_block_ms_per_step": round(wall_one / steps * 1e3, 4),
            "grains": reps[0].num_grains - 1}


def run_adjoint(n=128, nseeds=16, steps=4000):
    """Gradients of the final-ce L2 mismatch with respect to D_l, D_s, k and
    the initial ce field along a recorded trajectory of the seeded-grain
    workload (n^3, `nseeds` seeds from seed 4, half the explicit dt), against
    FP64 central finite differences of the frozen-branch replay at
    h = 1e-6, 1e-5 and 1e-4 relative, each numerator formed without
    cancellation (adjoint.l2_loss_difference).  The target is a free run
    with D_l, D_s, k scaled by 1.1, 0.8, 1.05."""
    import numpy as np
    import torch
    from mumatscale_amd import adjoint as ADJ
    from mumatscale_amd.lib import Subblock
    from mumatscale_amd.workloads import seeded_grains_config
    cfg = seeded_grains_config(n, nseeds, seed=4)
    t0 = time.perf_counter()
    with Subblock(cfg.params(check_negative_ce=False), cfg.nucleation()) as sb:
        traj = ADJ.record(sb, steps)
        grains = sb.num_grains - 1
    t_rec = time.perf_counter() - t0
    cfg2 = seeded_grains_config(n, nseeds, seed=4)
    cfg2.d_liq *= 1.1
    cfg2.d_sol *= 0.8
    cfg2.part_coef *= 1.05
    with Subblock(cfg2.params(check_negative_ce=False), cfg2.nucleation()) as sb2:
        sb2.step(steps)
        target = torch.from_numpy(sb2.download(abi.FIELD_CE).reshape(-1)).cuda()
    base = (cfg.d_liq, cfg.d_sol, cfg.part_coef)
    x = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in base]
    ce0 = traj.ce0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = ADJ.l2_loss(traj, ADJ.solute_map(traj, *x, ce0), target)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0

    def ce_T(p, c0=None):
        (ce, _, _), _ = ADJ.replay(traj, *p, ce0=c0)
        return ce

    def dloss(pa, pb, ca=None, cb=None):
        return ADJ.l2_loss_difference(traj, ce_T(pa, ca), ce_T(pb, cb), target)

    names = ("d_liq", "d_sol", "part_coef")
    grad, fd, rel = {}, {}, {}
    hs = (1e-6, 1e-5, 1e-4)
    for i, v in enumerate(base):
        grad[names[i]] = g = x[i].grad.item()
        fd[names[i]], rel[names[i]] = {}, {}
        for h in hs:
            a, b = list(base), list(base)
            a[i] = v * (1 + h)
            b[i] = v * (1 - h)
            f = dloss(a, b) / (a[i] - b[i])
            fd[names[i]]["%g" % h], rel[names[i]]["%g" % h] = f, abs(g - f) / abs(f)
    rng = np.random.default_rng(11)
    v = torch.zeros_like(traj.ce0)
    traj.interior(v)[...] = torch.from_numpy(rng.normal(size=traj.interior(v).shape)).cuda()
    grad["ce0_random_direction"] = g = (ce0.grad * v).sum().item()
    fd["ce0_random_direction"], rel["ce0_random_direction"] = {}, {}
    for h in hs:
        f = dloss(base, base, traj.ce0 + h * v, traj.ce0 - h * v) / (2 * h)
        fd["ce0_random_direction"]["%g" % h] = f
        rel["ce0_random_direction"]["%g" % h] = abs(g - f) / abs(f)
    K = ADJ.checkpoint_interval(steps)
    P = traj.padded_cells
    return {"grid": "%d^3" % n, "seeds": nseeds, "grains": grains, "steps": steps, "dt": cfg.dt,
            "loss": L.item(), "checkpoint_interval": K,
            "fd_numerator": "sum of (a - b)(a + b - 2 t) over the interior, math.fsum",
            "device_bytes": {"branch_record": steps * ADJ.record_bytes(traj.params),
                             "checkpoints": ((steps + K - 1) // K) * 3 * 8 * P,
                             "segment_states": K * 3 * 8 * P},
            "record_s": round(t_rec, 3), "forward_s": round(t_fwd, 3), "backward_s": round(t_bwd, 3),
            "grad": grad, "central_fd": fd, "rel_err": rel,
            "max_rel_err_h1e-6": max(r["1e-06"] for r in rel.values()),
            "max_rel_err_best_h": max(min(r.values()) for r in rel.values())}


def run_am_track(steps=60):
    """The AM single track over a 2000-grain Voronoi substrate (1024 x 512 x
    256, dx = 0.5 um, Gaussian-beam temperature), its error against the FP64
    run of the same steps, in three forms:
      fp32dev            FP32-deviation storage, the beam temperature from its
                         separable factors (a 400-deep product over the block,
                         computed once per step);
      fp32dev_quadrature FP32-deviation storage, the 400-term quadrature per
                         cell (temperature.c:88-118);
      fp64               FP64 storage and the quadrature, bit for bit the
                         reference's.
    The quadrature forms are bound by FP64 VALU work, not HBM."""
    import numpy as np
    from mumatscale_amd.lib import Subblock
    from mumatscale_amd.workloads import am_track_config, apply_substrate, voronoi_substrate
    cfg = am_track_config()
    t0 = time.time()
    gr, recs = voronoi_substrate(cfg, 2000, 2)
    log("am track: substrate built in %.1f s" % (time.time() - t0))
    n = cfg.sbdim[0] * cfg.sbdim[1] * cfg.sbdim[2]
    out, fields = {}, {}
    legs = (("fp32dev", abi.PREC_FP32DEV, 1), ("fp32dev_quadrature", abi.PREC_FP32DEV, 0),
            ("fp64", abi.PREC_FP64, None))
    for key, prec, sep in legs:
        with Subblock(cfg.params(check_negative_ce=False, precision=prec), cfg.nucleation()) as sb:
            if sep is not None:
                sb.set_tuning(abi.TUNE_BEAM_SEP, sep)
            apply_substrate(sb, gr, recs)
            sb.set_kernel_timing(True)
            t0 = time.perf_counter()
            rep = sb.step(steps, want_sums=True)
            wall = time.perf_counter() - t0
            k_ms, _ = sb.kernel_time_ms()
            ce = sb.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1]
            temp = sb.download(abi.FIELD_TEMP)[1:-1, 1:-1, 1:-1] if key != "fp32dev_quadrature" else None
            fields[key] = (ce, rep, temp)
            out[key] = {"value": round(n * steps / wall / 1e6, 1), "unit": "Mcell-updates/s",
                        "ms_per_step": round(wall / steps * 1e3, 3), "k_solute_ms": round(k_ms, 3),
                        "grains": rep.num_grains - 1, "solid_fraction": rep.solid_sum / n}
            if key != "fp32dev":
                out[key]["beam_terms_per_s"] = round(n * 400 / (k_ms * 1e-3), 1) if k_ms > 0 else None
            log("am track %s: %.3f ms/step" % (key, wall / steps * 1e3))
    c64, r64, t64 = fields["fp64"]
    for key in ("fp32dev", "fp32dev_quadrature"):
        c32, r32, t32 = fields[key]
        err = {"solute_l2_rel": float(np.linalg.norm(c32 - c64) / np.linalg.norm(c64)),
               "solid_fraction_abs": abs(r32.solid_sum - r64.solid_sum) / n,
               "grains_fp32": r32.num_grains - 1, "grains_fp64": r64.num_grains - 1,
               "solute_drift_with_ledger": (r32.solute_sum + r32.solute_ledger - n * cfg.c_init) / (n * cfg.c_init)}
        if t32 is not None:
            err["temperature_rel_max"] = float(np.max(np.abs(t32 - t64) / np.maximum(np.abs(t64), 1.0)))
        out[key]["error_vs_fp64"] = err
    # The compiled quadrature loop issues 51 FP64 VALU instructions per term
    # (two divisions and an exp each); MI355X issues 78.6e12 / 2 FP64 lane
    # operations per second.
    valu_per_term = 51
    for key in ("fp32dev_quadrature", "fp64"):
        t = out[key]["beam_terms_per_s"]
        out[key]["valu_frac"] = round(t * valu_per_term / (78.6e12 / 2), 3) if t else None
    out.update({"cells": n, "steps": steps, "dt": cfg.dt, "substrate_grains": len(recs) - 1,
                "valu_ops_per_term": valu_per_term})
    return out


def _split(dims, ranks):
    """Subblock counts per axis for `ranks` MPI ranks (z first, then y, x)."""
    nsub = [1, 1, 1]
    left = ranks
    for a in (2, 1, 0):
        for f in range(left, 0, -1):
            if left % f == 0 and dims[a] % (nsub[a] * f) == 0 and dims[a] // (nsub[a] * f) >= 16:
                nsub[a] *= f
                left //= f
                break
    return nsub if left == 1 else None


def _run_in_group(cmd, timeout, **kw):
    """subprocess.run in a process group of its own.  The whole group
    (mpiexec and the proxies it starts) is killed once the run returns or
    times out, so no helper outlives it."""
    import signal
    import subprocess
    p = subprocess.Popen(cmd, start_new_session=True, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, **kw)
    try:
        out, err = p.communicate(timeout=timeout)
    finally:
        try:
            os.killpg(p.pid, signal.SIGKILL)
        except ProcessLookupError:
            pass
        if p.returncode is None:
            p.communicate()
    return subprocess.CompletedProcess(cmd, p.returncode, out, err)


def reference_baseline(cfg, ranks, lo=2, hi=14, repeats=2):
    """The reference's own CPU build (oracle/_ref, built from its sources) on
    `ranks` MPI ranks of the host, one OpenMP thread each, the domain split
    into as many subblocks.
    Per-step time = (elapsed(hi steps) - elapsed(lo steps)) / (hi - lo), which
    cancels the set-up and the end-of-run grain statistics; each elapsed time
    is the least of `repeats` runs.  The runs use cfg's dt.  None when the
    build or mpiexec is missing or a run fails."""
    import re
    import shutil
    import subprocess
    import tempfile
    from mumatscale_amd.ctrl import write_ctrl
    ref = os.path.join(ROOT, "oracle", "_ref", "mumatscale_ref")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    nsub = _split(cfg.gdim, ranks)
    if not (os.path.exists(ref) and os.path.exists(mpiexec) and nsub):
        return None
    times = {}
    with tempfile.TemporaryDirectory() as tmp:
        for steps in (lo, hi) * repeats:
            c = directional_config(*cfg.gdim)
            c.dt = cfg.dt
            c.finish_time = (steps - 0.5) * c.dt      # done after exactly `steps` steps
            c.screenpfreq = 100000
            c.basefilename = "refbase"
            path = os.path.join(tmp, "run%d.in" % steps)
            write_ctrl(c, path, nsub=nsub)
            try:
                r = _run_in_group([mpiexec, "-hosts", "127.0.0.1", "-n", str(ranks), ref, "-c", path], 600,
                                  cwd=tmp, env=dict(os.environ, OMP_NUM_THREADS="1"))
            except (OSError, subprocess.SubprocessError) as e:
                log("reference baseline: %s" % e)
                return None
            m = re.search(r"Elapsed time:\s+(\d+) iterations, ([0-9.eE+-]+) sec", r.stdout)
            if r.returncode != 0 or not m or int(m.group(1)) != steps:
                log("reference baseline: run of %d steps failed (rc %d)" % (steps, r.returncode))
                return None
            times[steps] = min(times.get(steps, float("inf")), float(m.group(2)))
    n = cfg.gdim[0] * cfg.gdim[1] * cfg.gdim[2]
    per_step = (times[hi] - times[lo]) / (hi - lo)
    if per_step <= 0:
        log("reference baseline: elapsed times %s do not grow with the step count" % times)
        return None
    return {"value": round(n / per_step / 1e6, 3), "unit": "Mcell-updates/s", "cores": ranks, "kind": "reference",
            "sample": "reference CPU build, %d MPI ranks (subblocks %dx%dx%d), full-size workload, steps %d-%d: "
                      "elapsed(%d steps) %.2f s - elapsed(%d steps) %.2f s, least of %d runs each"
                      % (ranks, nsub[0], nsub[1], nsub[2], lo + 1, hi, hi, times[hi], lo, times[lo], repeats)}


def cpu_baseline(cfg, target_s):
    """The oracle (C restatement of the reference path) on the host cores:
    its per-cell loops run with OpenMP (OMP_NUM_THREADS threads), list
    building and nucleation serially; a bounded sample of the same workload."""
    from oracle.oracle import OracleSubblock, threads
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
    nthr = threads()
    return {"value": round(n * steps / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": nthr,
            "kind": "port",
            "sample": "same workload at full size, steps 2-%d after set-up (%.1f s), %d OpenMP threads"
                      % (steps + 1, dt, nthr)}


def capture_footprint(shape, steps=(250, 500, 750, 1000)):
    """The capture kernel's read floor on the directional workload: at each
    of `steps`, the interface-list length and the distinct 128-byte lines of
    gr, dc and fs that its solid face neighbours occupy, plus the list's own
    streamed bytes (slot id, dc_tmp, neighbour bits).  Under rocprofv3 with
    FETCH_SIZE this gives the re-read factor of the kernel."""
    import numpy as np
    from mumatscale_amd.lib import Subblock
    cfg = directional_config()
    sb = Subblock(cfg.params(check_negative_ce=False), cfg.nucleation(), device=0)
    apply_shape(sb, shape)
    done = 0
    for t in steps:
        rep = sb.step(t - done)
        done = t
        gr = sb.download(abi.FIELD_GR).reshape(-1)
        lst = sb.interface_list().astype(np.int64)
        sx = cfg.sbdim[0] + 2
        sxy = sx * (cfg.sbdim[1] + 2)
        refs = []
        for off in (1, sx, -1, -sx, sxy, -sxy):
            n = lst + off
            refs.append(n[gr[n] > 0])
        nref = sum(r.size for r in refs)
        sol = np.unique(np.concatenate(refs))
        lines = (np.unique(sol * 4 // 128).size
                 + np.unique(np.concatenate([sol * 24 // 128, (sol * 24 + 23) // 128])).size
                 + np.unique(sol * 8 // 128).size)
        floor = lines * 128 + lst.size * (4 + 24 + 1)
        print(json.dumps({"step": rep.timestep, "members": int(lst.size), "solid_neighbour_refs": int(nref),
                          "distinct_solid_neighbours": int(sol.size), "read_floor_bytes": int(floor)}),
              flush=True)
    sb.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=str, default="256x256x512")
    ap.add_argument("--precision", choices=("fp64", "fp32dev"), default="fp64",
                    help="storage of the timed headline run (the line's metric is quoted on fp64)")
    ap.add_argument("--zm", type=int, default=None,
                    help="z-marching solute kernel, planes per workgroup (0: cell-parallel; default 128 for FP64, "
                         "8 for FP32-deviation storage)")
    ap.add_argument("--zm-rows", type=int, default=None,
                    help="rows per z-marching workgroup (4, 8, 16; default 16 for FP64, 4 for FP32-deviation)")
    ap.add_argument("--xcd", type=int, default=None,
                    help="1: contiguous tile runs per XCD (default 1 for FP64, 0 for FP32-deviation)")
    ap.add_argument("--lds", type=int, default=1,
                    help="1: z-marching kernel with its planes staged in LDS by direct loads (k_solute_zl)")
    ap.add_argument("--capture-chunk", type=int, default=0,
                    help=">0: capture deals its list slots in chunks of this many, pinned to XCDs")
    ap.add_argument("--capture-band", type=int, default=0,
                    help=">0: capture works on units of this many rows of one plane, bands pinned to XCDs")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the FP32-deviation, literal-dt, coupled, ensemble and adjoint measurements")
    ap.add_argument("--extra-steps", type=int, default=1000)
    ap.add_argument("--no-am", action="store_true",
                    help="skip the AM single track (FP32 against FP64, 1024x512x256)")
    ap.add_argument("--am-steps", type=int, default=60)
    ap.add_argument("--capture-footprint", action="store_true",
                    help="print the capture kernel's read floor at steps 250-1000 and exit")
    ap.add_argument("--coupled-only", action="store_true",
                    help="run only the 2x2x2 coupled-ranks leg (and its one-block comparison), print it and exit")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched-ensemble leg (--steps steps), print it and exit")
    ap.add_argument("--am-only", action="store_true",
                    help="run only the AM single-track leg (--am-steps steps), print it and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the CA step on one GPU; --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    dims = tuple(int(v) for v in args.size.split("x"))
    # tile shapes: long 16-row tiles for FP64 storage (the fastest over the
    # 5000-step run), short 8 x 4 tiles for FP32-deviation storage
    fp32_tile = {"zm": 8, "rows": 4, "xcd": 0}
    fp64_tile = {"zm": 128, "rows": 16, "xcd": 1}
    dflt = fp32_tile if args.precision == "fp32dev" else fp64_tile
    args.zm = dflt["zm"] if args.zm is None else args.zm
    args.zm_rows = dflt["rows"] if args.zm_rows is None else args.zm_rows
    args.xcd = dflt["xcd"] if args.xcd is None else args.xcd
    shape = {"zm": args.zm, "rows": args.zm_rows, "xcd": args.xcd, "lds": args.lds, "chunk": args.capture_chunk,
             "band": args.capture_band}
    if args.capture_footprint:
        capture_footprint(shape)
        return
    if args.coupled_only:
        print(json.dumps(run_coupled(dims, args.extra_steps, args.warmup, shape)), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps(run_ensemble(args.steps, 20)), flush=True)
        return
    if args.am_only:
        print(json.dumps(run_am_track(args.am_steps)), flush=True)
        return
    nx, ny, nz = dims
    cfg = directional_config(nx, ny, nz)
    cfg.dt *= 0.5                     # the stable dt (half the explicit limit)
    fp32 = args.precision == "fp32dev"
    prec = abi.PREC_FP32DEV if fp32 else abi.PREC_FP64
    b_alg = B_ALG_FP32 if fp32 else B_ALG_FP64

    from mumatscale_amd.lib import Subblock
    t0 = time.time()
    gpu = Subblock(cfg.params(check_negative_ce=True, precision=prec), cfg.nucleation(), device=0)
    apply_shape(gpu, shape)
    log("gpu set-up %.1f s" % (time.time() - t0))
    n = nx * ny * nz
    if args.warmup:
        gpu.step(args.warmup)
    gpu.set_kernel_timing(True)
    t0 = time.perf_counter()
    rep = gpu.step(args.steps)        # synchronises on return; raises on a negative ce
    t1 = time.perf_counter()
    gpu.set_kernel_timing(False)
    k_ms, k_n = gpu.kernel_time_ms()
    wall = t1 - t0
    value = n * args.steps / wall / 1e6
    log("gpu: %d steps in %.3f s, grains %d, gindex %d, k_solute avg %.4f ms over %d launches"
        % (args.steps, wall, rep.num_grains, rep.gindex, k_ms, k_n))
    gpu.close()
    probe = bandwidth_probe(n)
    log("bandwidth probe:", probe)
    achieved = n * b_alg / (k_ms * 1e-3) / 1e9 if k_ms > 0 else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(args.size, shape, "fp32" if fp32 else "fp64"),
            "kernel": ("k_solute_%s<%d>" % ("zl" if args.lds else "zm", args.zm_rows)) if args.zm > 0
            else "k_solute<true,true>",
            "bytes_per_cell": b_alg}
    if probe and achieved:
        mix = probe["solute_mix_fp32dev" if fp32 else "solute_mix_fp64"]["GB/s"]
        roof["probe_same_mix_gbs"] = mix
        roof["frac_of_probe"] = round(achieved / mix, 4)
    extra = {"bandwidth_probe": probe}
    if not args.no_extra:
        if not fp32:
            extra["fp32dev"] = run_fp32_variant(cfg, args.extra_steps, args.warmup, dict(shape, **fp32_tile))
            log("fp32dev:", extra["fp32dev"])
        extra["literal_dt"] = run_literal_dt(dims, args.extra_steps, args.warmup, shape)
        log("literal dt:", extra["literal_dt"])
        extra["coupled_ranks"] = run_coupled(dims, args.extra_steps, args.warmup, shape)
        log("coupled ranks:", extra["coupled_ranks"])
        extra["batched_ensemble"] = run_ensemble(3000, 200)
        log("ensemble:", extra["batched_ensemble"])
        extra["adjoint"] = run_adjoint()
        log("adjoint:", extra["adjoint"])
        if not args.no_am:
            extra["am_track"] = run_am_track(args.am_steps)
            log("am track:", extra["am_track"])
    cpu = None
    if not args.no_cpu:
        ranks = min(16, os.cpu_count() or 1)
        cpu = reference_baseline(cfg, ranks)
        log("reference cpu baseline:", cpu)
        port = cpu_baseline(cfg, args.cpu_seconds)
        log("oracle cpu baseline:", port)
        if cpu is None:
            cpu = port
        else:
            cpu["oracle_port"] = port
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
        "config": {"workload": "directional solidification %dx%dx%d, %s storage" % (nx, ny, nz,
                                                                                  "FP32-deviation" if fp32
                                                                                  else "plain FP64"),
                   "cells": n, "dt": cfg.dt, "zm_planes": args.zm, "zm_rows": args.zm_rows,
                   "xcd_remap": args.xcd, "zm_lds": args.lds, "capture_chunk": args.capture_chunk,
                   "capture_band": args.capture_band,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
