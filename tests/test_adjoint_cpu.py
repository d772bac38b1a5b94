"""Frozen-branch map on the CPU: the branch record of the oracle and the
torch FP64 step (tests/adjoint_torch.py) that the HIP replay and adjoint
kernels are checked against.

- Replaying the oracle's recorded branches with the torch step reproduces
  the oracle's (ce, cl, fs) bit for bit: the frozen-branch map is the CA's
  own solute update, including remelting and wrapped faces.
- The branch codes agree with what each branch does to the state.
- torch.autograd through the step agrees with central finite differences
  (a check of the map's smoothness with frozen branches, not of autograd).
"""
import numpy as np
import pytest
import torch

import adjoint_torch as AT
from golden_cases import ctrl_path
from mumatscale_amd import abi
from mumatscale_amd.ctrl import read_ctrl
from mumatscale_amd.workloads import seeded_grains_config
from oracle.oracle import OracleSubblock

FIELDS = (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS)


def oracle_record(cfg, steps):
    o = OracleSubblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    init = [o.download(f) for f in FIELDS]
    rec, st = o.step_recording(steps)
    assert st == abi.OK
    final = [o.download(f) for f in FIELDS]
    return init, rec, final


CASES = [
    ("seeded_grains_20x16x12", lambda: seeded_grains_config(n=20, nseeds=3, seed=4, ny=16, nz=12), 400),
    ("c4_remelt_renucleate", lambda: read_ctrl(ctrl_path("c4_remelt_renucleate")), 1200),
]


@pytest.mark.parametrize("name,make,steps", CASES, ids=[c[0] for c in CASES])
def test_replay_reproduces_oracle(name, make, steps):
    cfg = make()
    init, rec, final = oracle_record(cfg, steps)
    counts = np.bincount(rec[:, 1:-1, 1:-1, 1:-1].ravel(), minlength=5)
    assert counts[abi.BRANCH_LIQUID] > 0 and counts[abi.BRANCH_GROW] > 0 and counts[abi.BRANCH_SOLID] > 0
    if name.startswith("c4"):
        assert counts[abi.BRANCH_REMELT] > 0 and counts[abi.BRANCH_FULL] > 0
    out = AT.run(cfg.params(), torch.from_numpy(rec), 0, *(torch.from_numpy(a) for a in init),
                 cfg.d_liq, cfg.d_sol, cfg.part_coef)
    for a, b, f in zip(out, final, FIELDS):
        assert np.array_equal(a.numpy(), b), "field %d differs" % f


def test_branch_codes_match_their_branch():
    """LIQUID/REMELT cells end with fs = 0 and cl = ce; SOLID cells keep
    fs = 1 with cl = ce * (1/k); GROW cells end with 0 <= fs <= 1 and FULL
    cells with fs = 1 (growth.c:228-272)."""
    cfg = read_ctrl(ctrl_path("c4_remelt_renucleate"))
    o = OracleSubblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    k_inv = 1.0 / cfg.part_coef
    seen = set()
    for _ in range(700):
        rec, st = o.step_recording(1)
        assert st == abi.OK
        # after capture: captured cells were LIQUID/REMELT with cl = ce, fs = 0 already
        ce, cl, fs = (o.download(f)[1:-1, 1:-1, 1:-1] for f in FIELDS)
        b = rec[0][1:-1, 1:-1, 1:-1]
        liq = (b == abi.BRANCH_LIQUID) | (b == abi.BRANCH_REMELT)
        assert np.all(fs[liq] == 0.0) and np.array_equal(cl[liq], ce[liq])
        sol = b == abi.BRANCH_SOLID
        assert np.all(fs[sol] == 1.0) and np.array_equal(cl[sol], ce[sol] * k_inv)
        grow = b == abi.BRANCH_GROW
        assert np.all((fs[grow] >= 0.0) & (fs[grow] <= 1.0))
        assert np.all(fs[b == abi.BRANCH_FULL] == 1.0)
        seen |= set(np.unique(b).tolist())
    assert seen == {0, 1, 2, 3, 4}


def test_torch_gradient_matches_central_differences():
    cfg = read_ctrl(ctrl_path("c4_remelt_renucleate"))
    steps = 40
    o = OracleSubblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    o.step(560)                                   # a state with every branch active
    t0 = o.timestep
    init = [torch.from_numpy(o.download(f)) for f in FIELDS]
    rec, st = o.step_recording(steps)
    assert st == abi.OK
    codes = torch.from_numpy(rec)
    p = cfg.params()
    rng = np.random.default_rng(7)
    target = init[0] + torch.from_numpy(rng.normal(scale=1e-3, size=init[0].shape))

    def loss(dl, ds, k, ce0):
        ce, _, _ = AT.run(p, codes, t0, ce0, init[1], init[2], dl, ds, k)
        d = (ce - target)[1:-1, 1:-1, 1:-1]
        return (d * d).sum()

    x = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (cfg.d_liq, cfg.d_sol, cfg.part_coef)]
    ce0 = init[0].clone().requires_grad_(True)
    L = loss(*x, ce0)
    L.backward()
    h = 1e-6
    for i, v in enumerate((cfg.d_liq, cfg.d_sol, cfg.part_coef)):
        a = [cfg.d_liq, cfg.d_sol, cfg.part_coef]
        b = list(a)
        a[i] = v * (1 + h)
        b[i] = v * (1 - h)
        fd = (loss(*a, init[0]) - loss(*b, init[0])).item() / (2 * h * v)
        g = x[i].grad.item()
        assert abs(g - fd) <= 1e-6 * abs(fd) + 1e-300, (i, g, fd)
    # directional derivative along a random interior field
    v = torch.zeros_like(init[0])
    v[1:-1, 1:-1, 1:-1] = torch.from_numpy(rng.normal(size=tuple(s - 2 for s in v.shape)))
    eps = 1e-6
    fd = (loss(cfg.d_liq, cfg.d_sol, cfg.part_coef, init[0] + eps * v) -
          loss(cfg.d_liq, cfg.d_sol, cfg.part_coef, init[0] - eps * v)).item() / (2 * eps)
    g = (ce0.grad * v).sum().item()
    assert abs(g - fd) <= 1e-6 * abs(fd), (g, fd)


def test_unpack_codes_follows_the_record_layout():
    """mumatscale_amd.adjoint.unpack_codes reads the 3-bit record layout that
    mmca.h states (three uint64 bit planes per 64-cell row segment), packed
    here independently from per-cell codes."""
    from mumatscale_amd import adjoint as ADJ
    from mumatscale_amd.ctrl import read_ctrl
    from golden_cases import ctrl_path
    for nx, ny, nz in ((70, 5, 3), (64, 2, 2), (17, 3, 4)):
        p = read_ctrl(ctrl_path("c4_remelt_renucleate")).params()
        p.nx, p.ny, p.nz = nx, ny, nz
        rng = np.random.default_rng(nx)
        T = 3
        codes = np.zeros((T, nz + 2, ny + 2, nx + 2), np.uint8)
        codes[:, 1:-1, 1:-1, 1:-1] = rng.integers(0, 5, size=(T, nz, ny, nx))
        w64 = (nx + 63) // 64
        words = np.zeros((T, nz + 2, ny + 2, w64, 3), np.uint64)
        for t, k, j, x in zip(*np.nonzero(codes)):
            seg, bit = (x - 1) // 64, (x - 1) % 64
            for pl in range(3):
                if (int(codes[t, k, j, x]) >> pl) & 1:
                    words[t, k, j, seg, pl] |= np.uint64(1) << np.uint64(bit)
        packed = torch.from_numpy(words.view(np.uint8).reshape(T, -1).copy())
        assert np.array_equal(ADJ.unpack_codes(p, packed).numpy(), codes)
