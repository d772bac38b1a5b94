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
