"""Entry points the reference's output path calls between steps.

- mmca_io_fields (cache_io_data, loop.c:74-113): FP32 copies converted on
  the device equal the FP64 downloads converted on the host.
- mmca_rand: raw rand() values drawn from the block's stream between steps,
  as the XDMF writer's grain renumbering does (io_xdmf.c:345-376), equal the
  oracle's, and grains created afterwards get the orientations the oracle
  draws after those values (the library's pre-drawn orientations are
  redrawn).
"""
import numpy as np
import pytest

from golden_cases import grain_table_bytes, load_case
from mumatscale_amd import abi
from mumatscale_amd.lib import Subblock
from oracle.oracle import OracleSubblock

pytestmark = pytest.mark.gpu


def test_io_fields_match_downloads():
    cfg, _, _, _ = load_case("c1_directional")
    for prec in (abi.PREC_FP64, abi.PREC_FP32DEV):
        with Subblock(cfg.params(check_negative_ce=False, precision=prec), cfg.nucleation()) as sb:
            sb.step(25)
            io = sb.io_fields()
            assert np.array_equal(io["fs"], sb.download(abi.FIELD_FS).astype(np.float32))
            assert np.array_equal(io["ce"], sb.download(abi.FIELD_CE).astype(np.float32))
            assert np.array_equal(io["temperature"], sb.download(abi.FIELD_TEMP).astype(np.float32))
            assert np.array_equal(io["gr"], sb.download(abi.FIELD_GR))


def test_rand_between_steps_keeps_the_stream():
    cfg, _, _, _ = load_case("c4_remelt_renucleate")      # random orientations, renucleation
    orc = OracleSubblock(cfg.params(check_negative_ce=False), cfg.nucleation())
    with Subblock(cfg.params(check_negative_ce=False), cfg.nucleation()) as sb:
        done = 0
        for t, n in ((150, 37), (400, 5), (600, 0), (900, 1), (1300, 11)):
            assert orc.step(t - done) == abi.OK
            sb.step(t - done)
            done = t
            ng = orc.num_grains
            assert sb.num_grains == ng, t
            assert grain_table_bytes(sb.grains(), ng) == grain_table_bytes(orc.grains(), ng), t
            assert np.array_equal(sb.download(abi.FIELD_GR), orc.field(abi.FIELD_GR)), t
            if n:
                assert np.array_equal(sb.rand(n), orc.rand(n)), t
        assert orc.num_grains > 10                         # grains were created after the draws


@pytest.mark.parametrize("case,prec,pre,post", [
    ("c4_remelt_renucleate", abi.PREC_FP64, 700, 500),
    ("c1_directional", abi.PREC_FP64, 30, 30),
    ("c4_remelt_renucleate", abi.PREC_FP32DEV, 400, 300),
], ids=["c4_fp64", "c1_fp64", "c4_fp32dev"])
def test_restart_image_continues_bit_for_bit(case, prec, pre, post):
    """mmca_save_state / mmca_load_state: a block restored into a fresh
    context continues exactly as the uninterrupted one, including the
    orientations of grains created after the restart and a raw draw taken
    just before the save."""
    cfg, _, _, _ = load_case(case)
    params = cfg.params(check_negative_ce=False, precision=prec)
    with Subblock(params, cfg.nucleation()) as a:
        a.step(pre)
        a.rand(3)
        image = a.save_state()
        a.step(post)
        with Subblock(params, cfg.nucleation()) as b:
            b.load_state(image)
            assert b.timestep == pre
            b.step(post)
            for f in (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS, abi.FIELD_GR, abi.FIELD_DC):
                assert np.array_equal(a.download(f), b.download(f)), f
            ng = a.num_grains
            assert b.num_grains == ng
            assert grain_table_bytes(a.grains(), ng) == grain_table_bytes(b.grains(), ng)
            assert np.array_equal(a.interface_list(), b.interface_list())
            assert np.array_equal(a.rand(4), b.rand(4))
        with Subblock(cfg.params(check_negative_ce=False, precision=prec, sb_coords=(0, 0, 0)),
                      cfg.nucleation()) as c:
            with pytest.raises(abi.MmcaError):
                c.load_state(image[:100])            # truncated image
