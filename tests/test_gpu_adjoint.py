"""Branch recording, frozen-branch replay and the adjoint on the GPU.

- The branches recorded by the HIP step equal the oracle's, and replaying
  them on the GPU (mmca_replay_step) reproduces the oracle's state bit for
  bit (uniform cooling with remelting and wrapped faces, frozen gradient).
- The adjoint (mmca_adjoint_step behind mumatscale_amd.adjoint.SoluteMap)
  equals torch.autograd through the torch FP64 step on the CPU.
- It agrees with FP64 central finite differences of the GPU replay
  (h = 1e-6 relative) for D_l, D_s, k and the initial ce field.
- sqrt(T) checkpointing gives the same gradients as keeping every state.
"""
import math

import numpy as np
import pytest
import torch

import adjoint_torch as AT
from golden_cases import ctrl_path
from mumatscale_amd import abi
from mumatscale_amd import adjoint as ADJ
from mumatscale_amd.ctrl import read_ctrl
from mumatscale_amd.lib import Subblock
from mumatscale_amd.workloads import seeded_grains_config
from oracle.oracle import OracleSubblock

pytestmark = pytest.mark.gpu

FIELDS = (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS)


def gpu_record(cfg, pre, steps):
    sb = Subblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    if pre:
        sb.step(pre)
    return sb, ADJ.record(sb, steps)


CASES = [
    ("c4_remelt_renucleate", lambda: read_ctrl(ctrl_path("c4_remelt_renucleate")), 0, 1200),
    ("seeded_grains_20x16x12", lambda: seeded_grains_config(n=20, nseeds=3, seed=4, ny=16, nz=12), 0, 400),
    ("c1_directional", lambda: read_ctrl(ctrl_path("c1_directional")), 0, 40),
]


@pytest.mark.parametrize("name,make,pre,steps", CASES, ids=[c[0] for c in CASES])
def test_record_and_replay_match_oracle(name, make, pre, steps):
    cfg = make()
    print("case", name, "steps", steps, flush=True)
    o = OracleSubblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    if pre:
        o.step(pre)
    rec, st = o.step_recording(steps)
    assert st == abi.OK
    final = [o.download(f) for f in FIELDS]
    sb, traj = gpu_record(cfg, pre, steps)
    codes = traj.unpacked_codes().cpu().numpy().reshape(rec.shape)
    assert np.array_equal(codes, rec)
    for f, ref in zip(FIELDS, final):
        assert np.array_equal(sb.download(f), ref)
    out, _ = ADJ.replay(traj, cfg.d_liq, cfg.d_sol, cfg.part_coef)
    torch.cuda.synchronize()
    for t, ref, f in zip(out, final, FIELDS):
        assert np.array_equal(t.cpu().numpy().reshape(ref.shape), ref), "replayed field %d differs" % f
    sb.close()


def test_record_refuses_fp32_and_overrun():
    cfg = seeded_grains_config(n=16, nseeds=1, seed=4)
    codes = torch.zeros((4, ADJ.record_bytes(cfg.params(check_negative_ce=False))), dtype=torch.uint8, device="cuda")
    with Subblock(cfg.params(check_negative_ce=False, precision=abi.PREC_FP32DEV), cfg.nucleation()) as sb:
        with pytest.raises(abi.MmcaError):
            sb.set_record(codes.data_ptr(), 4)
    with Subblock(cfg.params(check_negative_ce=False), cfg.nucleation()) as sb:
        torch.cuda.synchronize()
        sb.set_record(codes.data_ptr(), 4)
        sb.step(3)
        with pytest.raises(abi.MmcaError):
            sb.step(2)                          # past the end of the record
        sb.step(1)
        sb.set_record(None, 0)
        sb.step(5)


def test_adjoint_matches_torch_autograd():
    cfg = read_ctrl(ctrl_path("c4_remelt_renucleate"))
    sb, traj = gpu_record(cfg, 560, 40)         # every branch active in this window
    rng = np.random.default_rng(7)
    tgt_cpu = traj.ce0.cpu() + torch.from_numpy(rng.normal(scale=1e-3, size=traj.padded_cells))
    target = tgt_cpu.cuda()
    # GPU adjoint
    x = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (cfg.d_liq, cfg.d_sol, cfg.part_coef)]
    ce0 = traj.ce0.clone().requires_grad_(True)
    ceT = ADJ.solute_map(traj, *x, ce0)
    L = ADJ.l2_loss(traj, ceT, target)
    L.backward()
    # torch FP64 on the CPU, same branches
    shp = traj.shape
    y = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (cfg.d_liq, cfg.d_sol, cfg.part_coef)]
    ce0c = traj.ce0.cpu().view(shp).clone().requires_grad_(True)
    out = AT.run(cfg.params(), traj.unpacked_codes().cpu().view((traj.nsteps,) + shp), traj.t0, ce0c,
                 traj.cl0.cpu().view(shp), traj.fs0.cpu().view(shp), *y)
    assert torch.equal(out[0].detach().reshape(-1), ceT.detach().cpu())
    d = (out[0] - tgt_cpu.view(shp))[1:-1, 1:-1, 1:-1]
    Lc = (d * d).sum()
    Lc.backward()
    for a, b in zip(x, y):
        ga, gb = a.grad.item(), b.grad.item()
        assert abs(ga - gb) <= 1e-10 * abs(gb), (ga, gb)
    gc = ce0.grad.cpu().view(shp)
    ref = ce0c.grad
    assert torch.max(torch.abs(gc - ref)).item() <= 1e-10 * torch.max(torch.abs(ref)).item()
    sb.close()


@pytest.mark.timeout(600)
def test_adjoint_matches_central_differences():
    """The GPU adjoint against central finite differences at h = 1e-6
    (relative) for D_l, D_s, k and a random direction of the initial ce field.

    The finite differences are of the same frozen-branch map, evaluated in
    x87 extended precision (numpy.longdouble, 64-bit significand) from the
    recorded branches: the loss is ~4e4 times less sensitive to D_s than to
    D_l, so the two perturbed FP64 replays differ by less than their own
    rounding noise allows at this h (their quotient is printed beside).  The
    map itself is pinned to the GPU replay and the oracle elsewhere
    (test_record_and_replay_match_oracle; the FP64 numpy evaluation of
    adjoint_torch.run_numpy equals the oracle bit for bit)."""
    cfg = seeded_grains_config(n=32, nseeds=4, seed=4)
    steps = 600
    sb, traj = gpu_record(cfg, 0, steps)
    # target: a free run with D_l, D_s, k scaled by 1.1, 0.8, 1.05
    cfg2 = seeded_grains_config(n=32, nseeds=4, seed=4)
    cfg2.d_liq *= 1.1
    cfg2.d_sol *= 0.8
    cfg2.part_coef *= 1.05
    with Subblock(cfg2.params(check_negative_ce=False), cfg2.nucleation()) as sb2:
        sb2.step(steps)
        target = torch.from_numpy(sb2.download(abi.FIELD_CE).reshape(-1)).cuda()
    base = (cfg.d_liq, cfg.d_sol, cfg.part_coef)
    x = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in base]
    ce0 = traj.ce0.clone().requires_grad_(True)
    L = ADJ.l2_loss(traj, ADJ.solute_map(traj, *x, ce0), target)
    L.backward()

    shp = traj.shape
    LD = np.longdouble
    codes = traj.unpacked_codes().cpu().numpy().reshape((steps,) + shp)
    init = [t.cpu().numpy().reshape(shp) for t in (traj.ce0, traj.cl0, traj.fs0)]
    tgt = target.cpu().numpy().reshape(shp)[1:-1, 1:-1, 1:-1].astype(LD)

    def ld_dloss(pa, pb, ca=None, cb=None):
        """L(a) - L(b) of the map in extended precision, as sum (a-b)(a+b-2t)."""
        a = AT.run_numpy(traj.params, codes, traj.t0, init[0] if ca is None else ca, init[1], init[2], *pa)[0]
        b = AT.run_numpy(traj.params, codes, traj.t0, init[0] if cb is None else cb, init[1], init[2], *pb)[0]
        a, b = a[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1]
        return np.sum((a - b) * (a + b - 2 * tgt), dtype=LD)

    def ce_T(p, c0=None):
        (ce, _, _), _ = ADJ.replay(traj, *p, ce0=c0)
        return ce

    h = LD("1e-6")
    for i, v in enumerate(base):
        a, b = [LD(u) for u in base], [LD(u) for u in base]
        a[i] = LD(v) * (1 + h)
        b[i] = LD(v) * (1 - h)
        fd = float(ld_dloss(a, b) / (a[i] - b[i]))
        a64, b64 = list(base), list(base)
        a64[i], b64[i] = v * (1 + 1e-6), v * (1 - 1e-6)
        fd64 = ADJ.l2_loss_difference(traj, ce_T(a64), ce_T(b64), target) / (a64[i] - b64[i])
        g = x[i].grad.item()
        print("param", i, "h 1e-6 adjoint", g, "fd(extended)", fd, "rel", abs(g - fd) / abs(fd),
              "| fd(fp64)", fd64, "rel", abs(g - fd64) / abs(fd64), flush=True)
        assert abs(g - fd) <= 1e-6 * abs(fd), (i, g, fd)
    rng = np.random.default_rng(11)
    v = np.zeros(shp)
    v[1:-1, 1:-1, 1:-1] = rng.normal(size=(shp[0] - 2, shp[1] - 2, shp[2] - 2))
    eps = LD("1e-6")
    c0 = init[0].astype(LD)
    fd = float(ld_dloss(base, base, c0 + eps * v, c0 - eps * v) / (2 * eps))
    g = (ce0.grad * torch.from_numpy(v.reshape(-1)).cuda()).sum().item()
    print("ce0 direction adjoint", g, "fd(extended)", fd, "rel", abs(g - fd) / abs(fd), flush=True)
    assert abs(g - fd) <= 1e-6 * abs(fd), (g, fd)
    sb.close()


def test_checkpointing_is_exact():
    cfg = seeded_grains_config(n=24, nseeds=3, seed=5)
    steps = 150
    sb, traj = gpu_record(cfg, 0, steps)
    params = ADJ._params_with(traj.params, cfg.d_liq, cfg.d_sol, cfg.part_coef)
    rng = np.random.default_rng(3)
    g = torch.from_numpy(rng.normal(size=traj.padded_cells)).cuda()
    traj_g = torch.zeros_like(g)
    traj.interior(traj_g)[...] = traj.interior(g)
    res = []
    for K in (ADJ.checkpoint_interval(steps), 1, steps):
        _, keep = ADJ.replay(traj, cfg.d_liq, cfg.d_sol, cfg.part_coef, keep_every=K)
        gs, gp = ADJ.backward(traj, params, keep, K, traj_g)
        res.append((gs[0].cpu(), gp.cpu()))
    assert ADJ.checkpoint_interval(steps) == int(math.ceil(math.sqrt(steps)))
    for r in res[1:]:
        assert torch.equal(r[0], res[0][0]) and torch.equal(r[1], res[0][1])
    sb.close()
