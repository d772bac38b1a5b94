"""HIP path (through the C ABI) against the oracle, bit for bit.

Every golden case whose fixtures pin the oracle to the reference
(tests/golden/) is advanced on the GPU and on the oracle from the same
control file; fields are compared bit for bit at checkpoints: ce, cl, fs,
gr, the interface list, the grain table and dc (dc where the oracle marks no
cell as read from the reference's out-of-bounds dc_tmp slot).  Error cases
must stop at the same step with the same error.
"""
import numpy as np
import pytest

from mumatscale_amd import abi
from mumatscale_amd.lib import Subblock
from oracle.oracle import OracleSubblock
from golden_cases import CASES, grain_table_bytes, load_case

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int64) if a.dtype == np.float64 else a


def assert_same_state(gpu, orc, where, check_d=False):
    for f, name in ((abi.FIELD_CE, "ce"), (abi.FIELD_CL, "cl"), (abi.FIELD_FS, "fs"),
                    (abi.FIELD_GR, "gr")):
        a = gpu.download(f)
        b = orc.field(f)
        if not np.array_equal(bits(a), bits(b)):
            bad = np.argwhere(bits(a) != bits(b))
            z, y, x = bad[0]
            pytest.fail("%s: %s differs at %d cells, first (z,y,x)=(%d,%d,%d): gpu %r oracle %r"
                        % (where, name, len(bad), z, y, x, a[z, y, x], b[z, y, x]))
    dc_g = gpu.download(abi.FIELD_DC)
    dc_o = orc.field(abi.FIELD_DC)
    ok = orc.dc_taint() == 0
    if not np.array_equal(bits(dc_g)[ok], bits(dc_o)[ok]):
        pytest.fail("%s: dc differs" % where)
    assert gpu.num_grains == orc.num_grains, where
    assert np.array_equal(gpu.interface_list(), orc.diff_id()), where + ": interface list"
    ng = orc.num_grains
    assert grain_table_bytes(gpu.grains(), ng) == grain_table_bytes(orc.grains(), ng), where + ": grains"
    if check_d:
        # d of cells solid before the last capture (the others hold stale values
        # in the reference and are never read)
        gr = orc.field(abi.FIELD_GR)
        listed = np.zeros(gr.size, bool)
        listed[orc.diff_id()] = True
        sel = (gr.ravel() > 0) & ~listed
        assert np.array_equal(bits(gpu.download(abi.FIELD_D)).ravel()[sel],
                              bits(orc.field(abi.FIELD_D)).ravel()[sel]), where + ": d"


def assert_bits_equal(a, b, where):
    a, b = bits(a), bits(b)
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        pytest.fail("%s: %d cells differ, first at %s" % (where, len(bad), tuple(bad[0])))


def checkpoints(last):
    pts = list(range(1, min(last, 20) + 1))
    pts += list(range(40, last + 1, 40))
    if last > 0 and pts[-1] != last:
        pts.append(last)
    return pts


# every golden case, the Gaussian-beam one included (its exp is glibc's, mmca_exp.h)
BIT_EXACT_CASES = list(CASES)


# solute kernel variants (results must not depend on them): planes per
# z-marching workgroup (0 = cell-parallel kernel), rows per workgroup, XCD
# remap, planes staged in LDS by direct loads (k_solute_zl)
VARIANTS = [(0, 4, 0, 0), (8, 8, 1, 0), (5, 16, 0, 0), (16, 4, 1, 0), (8, 8, 1, 1), (5, 16, 0, 1), (16, 4, 1, 1),
            (1, 16, 1, 1)]


@pytest.mark.parametrize("variant", VARIANTS, ids=["cell", "zm8x8", "zm5x16", "zm16x4", "zl8x8", "zl5x16", "zl16x4",
                                                   "zl1x16"])
@pytest.mark.parametrize("case", BIT_EXACT_CASES)
def test_gpu_matches_oracle(case, variant):
    cfg, hashes, expected, _ = load_case(case)
    last = expected["last_step"]
    zm, rows, xcd, lds = variant
    print("case", case, "steps", last, "variant", variant, flush=True)
    orc = OracleSubblock(cfg.params(check_negative_ce=True), cfg.nucleation())
    with Subblock(cfg.params(check_negative_ce=True), cfg.nucleation()) as gpu:
        gpu.set_tuning(abi.TUNE_ZM_PLANES, zm)
        gpu.set_tuning(abi.TUNE_ZM_ROWS, rows)
        gpu.set_tuning(abi.TUNE_XCD_REMAP, xcd)
        gpu.set_tuning(abi.TUNE_ZM_LDS, lds)
        assert_bits_equal(gpu.download(abi.FIELD_THRESHOLD), orc.field(abi.FIELD_THRESHOLD), "thresholds")
        done = 0
        for t in checkpoints(last):
            assert orc.step(t - done) == abi.OK
            rep = gpu.step(t - done)
            done = t
            assert rep.timestep == t
            assert_same_state(gpu, orc, "%s step %d" % (case, t), check_d=(t % 200 == 0))
        if expected["returncode"] != 0:
            st_o = orc.step(1)
            rep = gpu.step(1, raise_on_error=False)
            assert rep.status == st_o != abi.OK
            assert rep.failed_step == orc.failed_step == last + 1
            with pytest.raises(abi.MmcaError):
                gpu.step(1)


def test_gpu_single_operators_match_oracle():
    """Each operator of doiteration() called on its own (calculate.c:247-343)."""
    cfg, _, _, _ = load_case("c1_directional")
    orc = OracleSubblock(cfg.params(), cfg.nucleation())
    with Subblock(cfg.params(), cfg.nucleation()) as gpu:
        for t in range(1, 41):
            where = "step %d" % t
            orc.timestep = t
            gpu.timestep = t
            orc.temp_update()
            n_o = orc.cell_nucleation()
            assert gpu.cell_nucleation() == n_o, where
            orc.activate_new_grains()
            gpu.activate_new_grains()
            assert_bits_equal(gpu.download(abi.FIELD_GR), orc.field(abi.FIELD_GR), where + " gr")
            orc.copy_oce()
            assert orc.diffuse() == abi.OK
            gpu.sb_diffuse_alloy_decentered()
            assert_bits_equal(gpu.download(abi.FIELD_CE), orc.field(abi.FIELD_CE), where + " ce")
            assert_bits_equal(gpu.download(abi.FIELD_OCE), orc.field(abi.FIELD_OCE), where + " oce")
            orc.fs_change()
            gpu.fs_change_diffuse()
            for f in (abi.FIELD_CL, abi.FIELD_FS, abi.FIELD_GR):
                assert_bits_equal(gpu.download(f), orc.field(f), where + " field %d" % f)
            orc.grow_octahedra()
            gpu.grow_octahedra()
            gr = orc.field(abi.FIELD_GR)
            sel = (gr > 0) & (gr < cfg.max_total_grains)
            assert_bits_equal(gpu.download(abi.FIELD_D)[sel], orc.field(abi.FIELD_D)[sel], where + " d")
            assert gpu.grow_cell_reduction() == orc.grow_cell_reduction(), where
            assert np.array_equal(gpu.interface_list(), orc.diff_id()), where
            orc.capture()
            gpu.capture_octahedra_diffuse()
            assert_same_state(gpu, orc, where)


def test_gpu_temperature_field_matches_oracle():
    """tempUpdate (temperature.c:219-272) for the three temperature models."""
    for case in ("c0_single_grain_half_dt", "c1_directional", "c3_am_beam"):
        cfg, _, _, _ = load_case(case)
        orc = OracleSubblock(cfg.params())
        with Subblock(cfg.params()) as gpu:
            for t in (0, 1, 57, 150, 299):
                orc.timestep = t
                gpu.timestep = t
                orc.temp_update()
                a = gpu.download(abi.FIELD_TEMP)
                b = orc.field(abi.FIELD_TEMP)
                assert np.array_equal(bits(a), bits(b)), (case, t)


def test_gpu_solid_volume_and_sums():
    cfg, _, _, _ = load_case("c1_directional")
    orc = OracleSubblock(cfg.params(), cfg.nucleation())
    with Subblock(cfg.params(), cfg.nucleation()) as gpu:
        orc.step(300)
        rep = gpu.step(300, want_sums=True)
        assert rep.solid_sum == pytest.approx(orc.solid_sum(), rel=1e-13)
        assert rep.solute_sum == pytest.approx(orc.solute_sum(), rel=1e-13)
        assert gpu.solid_volume() == pytest.approx(orc.solid_sum() * cfg.cell_size ** 3, rel=1e-13)


@pytest.mark.parametrize("variant", [(64, 16, 1, 1), (8, 4, 0, 1), (16, 8, 1, 0), (0, 4, 0, 0)],
                         ids=["zl64x16", "zl8x4", "zm16x8", "cell"])
def test_gpu_two_dimensional_block_matches_oracle(variant):
    """A one-plane directional block, 128x96x1: with a global z extent of 1 the
    frozen gradient runs along y (temperature.c:41-61), which the LDS-staged
    kernel evaluates once per row instead of once per plane.  Bit for bit
    against the oracle at checkpoints over 800 steps (123 grains, an interface
    of 2 372 cells at the end), for the LDS-staged kernel at two tile shapes,
    the register z-marching kernel and the cell-parallel kernel."""
    import bench
    zm, rows, xcd, lds = variant
    cfg = bench.directional_config(128, 96, 1)
    cfg.dt *= 0.5
    orc = OracleSubblock(cfg.params(check_negative_ce=True), cfg.nucleation())
    with Subblock(cfg.params(check_negative_ce=True), cfg.nucleation()) as gpu:
        gpu.set_tuning(abi.TUNE_ZM_PLANES, zm)
        gpu.set_tuning(abi.TUNE_ZM_ROWS, rows)
        gpu.set_tuning(abi.TUNE_XCD_REMAP, xcd)
        gpu.set_tuning(abi.TUNE_ZM_LDS, lds)
        orc.temp_update()
        assert_bits_equal(gpu.download(abi.FIELD_TEMP), orc.field(abi.FIELD_TEMP), "2-D temperature")
        done = 0
        for t in (1, 50, 200, 400, 800):
            assert orc.step(t - done) == abi.OK
            gpu.step(t - done)
            done = t
            assert_same_state(gpu, orc, "2-D step %d" % t, check_d=(t == 800))
        assert orc.num_grains > 100 and orc.gindex > 2000
        orc.close()


@pytest.mark.parametrize("band", [4, 12])
def test_gpu_capture_row_bands_match_oracle(band):
    """Capture on row bands pinned to XCDs (MMCA_TUNE_CAPTURE_BAND): a
    120x100x80 directional block (100 rows: the last band is partial for
    both band heights; 2 mask words per row, the second partial) against the
    oracle bit for bit for 400 steps."""
    import bench
    cfg = bench.directional_config(120, 100, 80)
    cfg.dt *= 0.5
    orc = OracleSubblock(cfg.params(check_negative_ce=True), cfg.nucleation())
    with Subblock(cfg.params(check_negative_ce=True), cfg.nucleation()) as gpu:
        gpu.set_tuning(abi.TUNE_CAPTURE_BAND, band)
        assert gpu.get_tuning(abi.TUNE_CAPTURE_BAND) == band
        done = 0
        for t in (1, 50, 150, 300, 400):
            assert orc.step(t - done) == abi.OK
            gpu.step(t - done)
            done = t
            print("bands", band, "step", t, "list", orc.gindex, flush=True)
            assert_same_state(gpu, orc, "band %d step %d" % (band, t))
        assert orc.gindex > 5000
        orc.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("lds,chunk", [(1, 16384)], ids=["zl_capture16k"])
def test_gpu_full_size_directional_block(lds, chunk):
    """The 256x256x512 directional workload at full size and the stable dt
    (half the explicit limit; at the literal limit the reference stops with a
    negative ce within ~25 steps): bit for bit against the oracle at every
    checkpoint up to step 1000 (ce, cl, fs, gr, dc, the interface list, the
    grain table), and total solute conserved to 1e-10 relative (zero flux on
    every face).  At this size the list grid has enough workgroups for the
    XCD-pinned capture schedule, which the small golden blocks do not reach."""
    import bench
    cfg = bench.directional_config()
    cfg.dt *= 0.5
    print("full size: oracle set-up", flush=True)
    orc = OracleSubblock(cfg.params(check_negative_ce=True), cfg.nucleation())
    with Subblock(cfg.params(check_negative_ce=True), cfg.nucleation()) as gpu:
        gpu.set_tuning(abi.TUNE_ZM_PLANES, 128)
        gpu.set_tuning(abi.TUNE_ZM_LDS, lds)
        gpu.set_tuning(abi.TUNE_CAPTURE_CHUNK, chunk)
        s0 = orc.solute_sum()
        done = 0
        for t in (1, 2, 3, 50, 200, 400, 700, 1000):
            assert orc.step(t - done) == abi.OK
            rep = gpu.step(t - done, want_sums=True)
            done = t
            print("full size: step", t, "grains", orc.num_grains, "list", orc.gindex, flush=True)
            assert_same_state(gpu, orc, "full size step %d" % t, check_d=(t in (200, 1000)))
            assert abs(rep.solute_sum - s0) <= 1e-10 * abs(s0), (t, rep.solute_sum, s0)
        assert orc.num_grains > 100
        orc.close()
