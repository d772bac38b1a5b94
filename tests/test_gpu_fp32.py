"""FP32-deviation storage against the FP64 path (both on the GPU).

ce and cl are stored as FP32 deviations from c0; fs, dc, the octahedron
half-diagonal, capture geometry and all arithmetic stay FP64.  The solute
lost to FP32 rounding of ce is kept in a compensated FP64 ledger: stored sum
+ ledger must equal the initial solute to FP64 rounding (every face is zero
flux in these cases), and the errors against the FP64 run are reported.
"""
import numpy as np
import pytest

from mumatscale_amd import abi
from mumatscale_amd.lib import Subblock
from golden_cases import load_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case,steps", [("c1_directional", 1000), ("c0_single_grain_half_dt", 2000)])
def test_fp32dev_against_fp64(case, steps):
    cfg, _, _, _ = load_case(case)
    p64 = cfg.params(precision=abi.PREC_FP64)
    p32 = cfg.params(precision=abi.PREC_FP32DEV)
    n = p64.nx * p64.ny * p64.nz
    with Subblock(p64, cfg.nucleation()) as a, Subblock(p32, cfg.nucleation()) as b:
        r0 = b.step(0, want_sums=True)
        total0 = n * cfg.c_init
        ra = a.step(steps, want_sums=True)
        rb = b.step(steps, want_sums=True)
        ce_a = a.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1]
        ce_b = b.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1]
        l2 = np.linalg.norm(ce_b - ce_a) / np.linalg.norm(ce_a)
        fs_err = abs(rb.solid_sum - ra.solid_sum) / n
        drift_stored = (rb.solute_sum - total0) / total0
        drift_ledger = (rb.solute_sum + rb.solute_ledger - total0) / total0
        print("%s: grains fp64 %d fp32 %d, solute L2 %.3e, fs err %.3e, stored drift %.3e, "
              "ledger drift %.3e" % (case, ra.num_grains, rb.num_grains, l2, fs_err, drift_stored,
                                     drift_ledger))
        assert abs(drift_ledger) <= 1e-12
        assert abs(drift_stored) <= 1e-6
        assert rb.num_grains == ra.num_grains
        assert fs_err <= 1e-3
        assert l2 <= 1e-3


def test_fp32dev_fields_do_not_depend_on_tile_shape():
    """The FP32-deviation fields are the same bits for every solute-kernel
    shape (only the ledger's summation order follows the workgroups)."""
    cfg, _, _, _ = load_case("c1_directional")
    p32 = cfg.params(precision=abi.PREC_FP32DEV)
    out = []
    for planes, rows, xcd, lds in ((16, 16, 1, 0), (7, 4, 0, 0), (32, 8, 1, 0), (0, 4, 0, 0), (16, 16, 1, 1),
                                   (7, 4, 0, 1), (32, 8, 1, 1)):
        print("fp32 solute kernel shape", planes, rows, xcd, lds, flush=True)
        with Subblock(p32, cfg.nucleation()) as sb:
            sb.set_tuning(abi.TUNE_ZM_PLANES, planes)
            sb.set_tuning(abi.TUNE_ZM_ROWS, rows)
            sb.set_tuning(abi.TUNE_XCD_REMAP, xcd)
            sb.set_tuning(abi.TUNE_ZM_LDS, lds)
            rep = sb.step(300, want_sums=True)
            out.append(([sb.download(f) for f in (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS, abi.FIELD_GR)], rep))
    f0, r0 = out[0]
    for fields, rep in out[1:]:
        for a, b in zip(f0, fields):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
        assert rep.num_grains == r0.num_grains
        assert rep.solute_sum == r0.solute_sum
        total, total0 = rep.solute_sum + rep.solute_ledger, r0.solute_sum + r0.solute_ledger
        assert abs(total - total0) <= 1e-12 * abs(total0)


def _dilate(m, r):
    """Cells within r face steps of a cell of m."""
    out = m.copy()
    for _ in range(r):
        o = out.copy()
        for ax in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax], hi[ax] = slice(1, None), slice(None, -1)
            o[tuple(lo)] |= out[tuple(hi)]
            o[tuple(hi)] |= out[tuple(lo)]
        out = o
    return out


@pytest.mark.timeout(600)
def test_fp32dev_error_at_full_size_comes_from_flipped_decisions():
    """Where the FP32-deviation error of the 256 x 256 x 512 directional
    block (stable dt, the bench's FP32 line) comes from.  The two runs are
    the same in every discrete decision up to step 300: grain counts and
    every cell's grain equal, solute L2 below 1e-6.  From then on a few
    cells are captured at another step than in FP64 (a capture test or a
    partitioning branch decided on values that differ in the last bits);
    each such cell carries an O(1) difference of its solute and solid
    fraction, which diffusion then spreads.  At step 1050 the L2 of the
    whole field reaches the 5e-3 class, but under 0.1 % of the solid cells
    have another grain, the grain counts are equal, and away from those cells
    (more than 8 cells) the L2 is about 1e-4."""
    from mumatscale_amd.workloads import directional_config
    cfg = directional_config(256, 256, 512)
    cfg.dt *= 0.5
    with Subblock(cfg.params(), cfg.nucleation()) as a, \
            Subblock(cfg.params(precision=abi.PREC_FP32DEV), cfg.nucleation()) as b:
        done = 0
        for t in (300, 1050):
            ra = a.step(t - done, want_sums=True)
            rb = b.step(t - done, want_sums=True)
            done = t
            ga = a.download(abi.FIELD_GR)[1:-1, 1:-1, 1:-1]
            gb = b.download(abi.FIELD_GR)[1:-1, 1:-1, 1:-1]
            ca = a.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1]
            cb = b.download(abi.FIELD_CE)[1:-1, 1:-1, 1:-1]
            diff = ga != gb
            nrm = np.linalg.norm(ca)
            l2 = float(np.linalg.norm(cb - ca) / nrm)
            away = ~_dilate(diff, 8)
            l2_away = float(np.linalg.norm((cb - ca)[away]) / nrm)
            solid = int(np.sum(ga > 0))
            print("step %d: grains %d/%d, cells with another grain %d of %d solid, L2 %.3e, L2 away %.3e"
                  % (t, ra.num_grains, rb.num_grains, int(diff.sum()), solid, l2, l2_away), flush=True)
            n = ca.size
            assert abs(rb.solute_sum + rb.solute_ledger - n * cfg.c_init) <= 1e-12 * n * cfg.c_init
            assert rb.num_grains == ra.num_grains
            if t == 300:
                assert not diff.any()
                assert l2 <= 1e-6
            else:
                assert diff.sum() <= 1e-3 * solid
                assert l2 <= 1e-2
                assert l2_away <= 5e-4
