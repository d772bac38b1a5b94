"""A polycrystalline substrate given through mmca_set_grains and field
uploads (the inputs of a restart, addGrainCache checkpoint.c:676-745), under
the Gaussian-beam AM model: the GPU against the oracle given the same
substrate, bit for bit (the beam temperature's exp gives the host C
library's bits)."""
import numpy as np
import pytest

from mumatscale_amd import abi
from mumatscale_amd.lib import Subblock
from mumatscale_amd.workloads import am_track_config, apply_substrate, seeded_grains_config, voronoi_substrate
from oracle.oracle import OracleSubblock

pytestmark = pytest.mark.gpu


def _records(n, rng):
    recs = [abi.Grain()]
    for i in range(1, n):
        g = abi.Grain()
        g.nuc_x, g.nuc_y, g.nuc_z = rng.uniform(size=3) * 1e-5
        g.nuc_timestep = int(rng.integers(0, 100))
        g.nuc_threshold = float(rng.uniform())
        for a in range(3):
            g.rotang[a] = float(rng.uniform())
            for b in range(3):
                g.rotmat[a][b] = float(rng.normal())
        recs.append(g)
    return recs


def test_set_grains_roundtrip_and_next_ids():
    cfg = seeded_grains_config(n=16, nseeds=2, seed=4)
    recs = _records(7, np.random.default_rng(1))
    with Subblock(cfg.params(check_negative_ce=False), cfg.nucleation()) as sb:
        sb.set_grains(recs)
        assert sb.num_grains == 7
        got = sb.grains()
        for a, b in zip(got[1:], recs[1:]):
            assert bytes(a) == bytes(b)
        sb.step(3)                                  # the two FixedNuc seeds nucleate at step 1
        assert sb.num_grains == 9
        ids = set(np.unique(sb.download(abi.FIELD_GR)).tolist()) - {0}
        assert ids == {7, 8}
        too_many = _records(cfg.max_total_grains + 2, np.random.default_rng(2))
        with pytest.raises(abi.MmcaError):
            sb.set_grains(too_many)


def test_am_substrate_matches_oracle():
    cfg = am_track_config(nx=48, ny=24, nz=16)
    cfg.laser["power"] = 24.0                      # melts about half of the block
    gr, recs = voronoi_substrate(cfg, 40, 2)
    params = cfg.params(check_negative_ce=False)
    orc = OracleSubblock(params, cfg.nucleation())
    orc.set_grains(recs)
    orc.upload(abi.FIELD_GR, gr)
    fs = np.zeros(gr.shape)
    fs[1:-1, 1:-1, 1:-1] = 1.0
    orc.upload(abi.FIELD_FS, fs)
    melted = 0
    with Subblock(params, cfg.nucleation()) as sb:
        apply_substrate(sb, gr, recs)
        done = 0
        for t in (5, 17, 25, 35, 50, 70, 90):
            print("am substrate step", t, flush=True)
            assert orc.step(t - done) == abi.OK
            sb.step(t - done)
            done = t
            g_gpu = sb.download(abi.FIELD_GR)
            assert np.array_equal(g_gpu, orc.field(abi.FIELD_GR)), t
            assert sb.num_grains == orc.num_grains, t
            assert np.array_equal(sb.interface_list(), orc.diff_id()), t
            for f in (abi.FIELD_CE, abi.FIELD_CL, abi.FIELD_FS):
                a, b = sb.download(f), orc.field(f)
                assert np.array_equal(a.view(np.int64), b.view(np.int64)), (t, f, np.max(np.abs(a - b)))
            melted = max(melted, int(np.sum(g_gpu[1:-1, 1:-1, 1:-1] <= 0)))
        final = sb.download(abi.FIELD_GR)[1:-1, 1:-1, 1:-1]
    assert melted > final.size // 4                # the beam remelted part of the substrate
    assert np.sum(final <= 0) < melted // 2        # which regrows from the grains that survived
    assert set(np.unique(final[final > 0]).tolist()) <= set(range(1, 41))

