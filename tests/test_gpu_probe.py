"""The streaming bandwidth probe the bench reports beside the solute kernel
(mmca_bandwidth_probe): every mix runs, its rate lies below MI355X's 8 TB/s
HBM peak, and bad arguments are refused before any allocation."""
import ctypes as C

import pytest

from mumatscale_amd import abi
from mumatscale_amd.lib import lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mix,bpc", [(0, 16.0), (1, 52.0), (2, 36.0)])
def test_probe_mixes(mix, bpc):
    # over 1 GB per mix, well beyond the 256 MB Infinity Cache (a smaller copy
    # is partly served from it and reads above the HBM peak); a partial last
    # run of 1024 cells
    cells = (1 << 26) + 1000
    ms, b = C.c_double(0), C.c_double(0)
    assert lib().mmca_bandwidth_probe(0, mix, cells, 5, C.byref(ms), C.byref(b)) == abi.OK
    assert b.value == bpc
    gbs = cells * bpc / (ms.value * 1e-3) / 1e9
    print("mix", mix, "%.1f GB/s" % gbs)
    assert 100.0 < gbs < 8000.0


def test_probe_rejects_bad_arguments():
    ms = C.c_double(0)
    for mix, cells, reps in ((3, 1000, 1), (-1, 1000, 1), (0, 0, 1), (0, 1 << 41, 1), (0, 1000, 0)):
        assert lib().mmca_bandwidth_probe(0, mix, cells, reps, C.byref(ms), None) == abi.ERR_ARG
