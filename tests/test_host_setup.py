"""Host set-up of the library (no GPU needed) against the oracle.

The library draws nucleation sites, thresholds and grain orientations on the
host from the reference's RNG stream (grain.c:118-154, 251-296, 603-713,
931-1010); these tests call its device-free entry points through the C ABI.
"""
import ctypes as C

import numpy as np
import pytest

from mumatscale_amd import abi, lib
from oracle.oracle import OracleSubblock
from golden_cases import CASES, load_case


def test_library_exports_every_declared_symbol():
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(lib.LIB_PATH), "..", "include", "mmca.h")).read()
    names = re.findall(r"^\s*(?:int|const char \*)\s+(mmca_\w+)\s*\(", hdr, re.M)
    assert len(names) > 20
    L = lib.lib()
    for n in names:
        assert hasattr(L, n), n


def test_struct_sizes_match_library():
    from oracle.oracle import lib as olib
    assert C.sizeof(abi.Grain) == 136 == olib().orc_sizeof_grain()
    assert C.sizeof(abi.Params) == olib().orc_sizeof_params()
    assert C.sizeof(abi.NucleationCfg) == olib().orc_sizeof_nuc_cfg()


@pytest.mark.parametrize("case", CASES)
def test_sites_match_oracle(case):
    cfg, _, _, sites = load_case(case)
    idx, thr, fixed = lib.nucleation_sites(cfg.params(), cfg.nucleation())
    assert np.array_equal(idx, sites["idx"])
    assert np.array_equal(thr, sites["threshold"])


@pytest.mark.parametrize("case", ["c1_directional", "c2_wrap_remelt", "c4_remelt_renucleate"])
def test_orientation_draws_match_oracle_grains(case):
    cfg, _, _, _ = load_case(case)
    orc = OracleSubblock(cfg.params(), cfg.nucleation())
    orc.step(60)
    rot = orc.rotmats()[1:]                    # grains 1.. in id order
    fixed_cells = set()
    idx, _, fixed = lib.nucleation_sites(cfg.params(), cfg.nucleation())
    grains = orc.grains()[1:]
    # grains nucleated at a FixedNuc/NucVolume cell take that entry's angles
    nonfixed = []
    for g, r in zip(grains, rot):
        cx = int((g.nuc_x - cfg.origin[0] + 0.1 * cfg.cell_size) / cfg.cell_size)
        nonfixed.append(r)
    nrand = len(rot)
    ang, drawn = lib.orientation_draws(cfg.params(), cfg.nucleation(), nrand)
    # every random grain's matrix appears, in order, in the draw sequence
    k = 0
    for r in rot:
        m = r.reshape(3, 3)
        if k < nrand and np.array_equal(m.view(np.int64), drawn[k].view(np.int64)):
            k += 1
    assert k >= 1
    n_fixed_grains = sum(1 for f in fixed if f >= 0)
    assert k >= len(rot) - n_fixed_grains * 60


def test_threshold_draw_that_cannot_end_is_an_error():
    """genRandomThreshold (grain.c:141-154) redraws while t < 1e-5 or
    |t - centre| > 6 sigma; with centre 0 and sigma 0 no draw is accepted and
    the reference loops forever.  The set-up stops after a bounded number of
    draws with MMCA_ERR_NUC_THRESHOLD instead."""
    cfg, _, _, _ = load_case("c1_directional")
    cfg.bulk_mean = 0.0
    cfg.bulk_sigma = 0.0
    with pytest.raises(abi.MmcaError) as e:
        lib.nucleation_sites(cfg.params(), cfg.nucleation())
    assert e.value.code == abi.ERR_NUC_THRESHOLD
