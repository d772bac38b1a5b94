"""Batched stepping: golden cases of different sizes and boundaries advanced
together (one set of launches per step for all of them) must each match
their own oracle run bit for bit, with and without HIP-graph replay."""
import numpy as np
import pytest

from mumatscale_amd import abi
from mumatscale_amd.lib import Batch
from oracle.oracle import OracleSubblock
from golden_cases import load_case

pytestmark = pytest.mark.gpu

CASES = ["c0_single_grain_half_dt", "c1_directional", "c2_wrap_remelt", "c4_remelt_renucleate"]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int64) if a.dtype == np.float64 else a


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("zm,lds", [(0, 0), (16, 0), (16, 1)])
def test_batch_matches_oracle(graph, zm, lds):
    cfgs = [load_case(c)[0] for c in CASES]
    orcs = [OracleSubblock(c.params(), c.nucleation()) for c in cfgs]
    with Batch([c.params() for c in cfgs], [c.nucleation() for c in cfgs]) as bt:
        bt.set_tuning(abi.TUNE_GRAPH, graph)
        bt.set_tuning(abi.TUNE_ZM_PLANES, zm)
        bt.set_tuning(abi.TUNE_ZM_LDS, lds)
        done = 0
        for t in (1, 2, 7, 40, 121, 400, 1000):
            reps = bt.step(t - done)
            for o in orcs:
                assert o.step(t - done) == abi.OK
            done = t
            for name, o, d, r in zip(CASES, orcs, bt.domains, reps):
                assert r.timestep == t
                assert r.num_grains == o.num_grains, (name, t)
                for f in (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS, abi.FIELD_GR):
                    assert np.array_equal(bits(d.download(f)), bits(o.field(f))), (name, t, f)
                ok = o.dc_taint() == 0
                assert np.array_equal(bits(d.download(abi.FIELD_DC))[ok], bits(o.field(abi.FIELD_DC))[ok])
                assert np.array_equal(d.interface_list(), o.diff_id()), (name, t)


@pytest.mark.parametrize("band", [8])
def test_batch_capture_row_bands_match_oracle(band):
    """Capture on row bands in a batch: two directional blocks of different
    sizes (their capture workgroups start at different XCDs) and a golden
    case too small for the bands, each against its oracle bit for bit."""
    import bench
    cfgs = [bench.directional_config(100, 72, 64), bench.directional_config(64, 96, 40), load_case("c1_directional")[0]]
    for c in cfgs[:2]:
        c.dt *= 0.5
    orcs = [OracleSubblock(c.params(), c.nucleation()) for c in cfgs]
    with Batch([c.params() for c in cfgs], [c.nucleation() for c in cfgs]) as bt:
        bt.set_tuning(abi.TUNE_CAPTURE_BAND, band)
        done = 0
        for t in (1, 60, 200, 300):
            reps = bt.step(t - done)
            for o in orcs:
                assert o.step(t - done) == abi.OK
            done = t
            for i, (o, d, r) in enumerate(zip(orcs, bt.domains, reps)):
                assert r.num_grains == o.num_grains, (i, t)
                for f in (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS, abi.FIELD_GR):
                    assert np.array_equal(bits(d.download(f)), bits(o.field(f))), (i, t, f)
                ok = o.dc_taint() == 0
                assert np.array_equal(bits(d.download(abi.FIELD_DC))[ok], bits(o.field(abi.FIELD_DC))[ok]), (i, t)
                assert np.array_equal(d.interface_list(), o.diff_id()), (i, t)
        assert min(o.gindex for o in orcs[:2]) > 1000


def test_batch_restart_image_continues_bit_for_bit():
    """mmca_batch_save_state / mmca_batch_load_state on a batch of separate
    blocks: the loaded batch continues exactly as the uninterrupted one."""
    cfgs = [load_case(c)[0] for c in CASES]
    params = [c.params() for c in cfgs]
    nucs = [c.nucleation() for c in cfgs]
    with Batch(params, nucs) as a, Batch(params, nucs) as b:
        a.step(121)
        b.load_state(a.save_state())
        a.step(279)
        b.step(279)
        for name, da, db in zip(CASES, a.domains, b.domains):
            for f in (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS, abi.FIELD_GR, abi.FIELD_DC):
                assert np.array_equal(bits(da.download(f)), bits(db.download(f))), (name, f)
            assert np.array_equal(da.interface_list(), db.interface_list()), name
            assert da.num_grains == db.num_grains, name


def test_corrupted_restart_image_is_refused():
    """A restart image whose interface list names a cell outside the
    interior, or whose gr names a grain beyond the table, is refused (capture
    would write, the kernels read, through it), and a block whose load failed
    part way refuses to step.  An uploaded gr field is checked the same way."""
    from mumatscale_amd.lib import Subblock
    cfg = load_case("c1_directional")[0]
    with Subblock(cfg.params(), cfg.nucleation()) as a, Subblock(cfg.params(), cfg.nucleation()) as b:
        a.step(120)
        lst = a.interface_list()
        assert len(lst) > 8
        img = bytearray(a.save_state())
        pat = np.asarray(lst[:8], dtype=np.int32).tobytes()
        at = bytes(img).find(pat)
        assert at > 0
        img[at:at + 4] = np.int32(-5).tobytes()                 # slot 1 -> outside the box
        with pytest.raises(abi.MmcaError):
            b.load_state(bytes(img))
        with pytest.raises(abi.MmcaError):
            b.step(1)
        gr = a.download(abi.FIELD_GR)
        img = bytearray(a.save_state())
        at = bytes(img).find(gr.tobytes())
        assert at > 0
        g = np.frombuffer(img, dtype=np.int32, count=gr.size, offset=at).copy()
        g[gr.size // 2] = cfg.max_total_grains + 1             # one past the grain table
        img[at:at + 4 * gr.size] = g.tobytes()
        with Subblock(cfg.params(), cfg.nucleation()) as c:
            with pytest.raises(abi.MmcaError):
                c.load_state(bytes(img))
            with pytest.raises(abi.MmcaError):
                c.upload(abi.FIELD_GR, g.reshape(gr.shape))
            c.upload(abi.FIELD_GR, gr)                            # in range: accepted
