"""The reference's own regression tests, run through this build's loop.

tests/testOneGrain.py:25-39 of the reference runs tests/one_grain.in on 6 MPI
ranks and expects "1 grains" with an average size of 0.000737875 mm^3
(tolerance 1e-6).  With one fixed nucleus and no random-orientation grain
before termination, the result does not depend on the decomposition or on
the RNG streams, so the 60x60x90 domain runs here as one block.
"""
import os

import pytest

from mumatscale_amd import driver
from mumatscale_amd.ctrl import read_ctrl
from golden_cases import GOLDEN

ONE_GRAIN = os.path.join(GOLDEN, "reference_inputs", "one_grain.in")
EXPECTED_GRAINS = 1                 # tests/testOneGrain.py:29-30
EXPECTED_AVERAGE = 0.000737875      # mm^3
TOLERANCE = 1e-6


def _check(res):
    words = res["average_line"].split()
    assert int(words[0]) == EXPECTED_GRAINS
    assert abs(float(words[9]) - EXPECTED_AVERAGE) <= TOLERANCE
    assert res["average_line"] == (
        "1 grains.  Min: 0.000737875 mm^3   Max: 0.000737875 mm^3   Average:  0.000737875 mm^3")


def test_one_grain_oracle():
    from oracle.oracle import OracleStepper, OracleSubblock
    cfg = read_ctrl(ONE_GRAIN)
    orc = OracleSubblock(cfg.params(), cfg.nucleation())
    res = driver.run(cfg, OracleStepper(orc), out=lambda s: None)
    assert res["timestep"] == 9000
    _check(res)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32dev"])
def test_one_grain_gpu(precision):
    from mumatscale_amd import abi
    from mumatscale_amd.lib import Subblock
    cfg = read_ctrl(ONE_GRAIN)
    prec = abi.PREC_FP64 if precision == "fp64" else abi.PREC_FP32DEV
    with Subblock(cfg.params(precision=prec), cfg.nucleation()) as sb:
        res = driver.run(cfg, sb, out=print)
    assert res["timestep"] == 9000
    if precision == "fp64":
        _check(res)
    else:
        # FP32 storage: same grain, size within the reference's tolerance
        words = res["average_line"].split()
        assert int(words[0]) == EXPECTED_GRAINS
        assert abs(float(words[9]) - EXPECTED_AVERAGE) <= TOLERANCE
