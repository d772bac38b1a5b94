"""The C host driver (mumatscale_amd/mmca_run): its control-file reader
against the Python reader, and the reference's OneGrain regression run end
to end through it (tests/testOneGrain.py:25-39 of the reference: "1 grains",
average 0.000737875 mm^3 within 1e-6)."""
import glob
import math
import os
import subprocess

import pytest

from golden_cases import GOLDEN
from mumatscale_amd import abi
from mumatscale_amd.ctrl import read_ctrl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mumatscale_amd", "mmca_run")
CTRL_FILES = sorted(glob.glob(os.path.join(GOLDEN, "cases", "*.in")) +
                    glob.glob(os.path.join(GOLDEN, "cases_mr", "*.in")) +
                    glob.glob(os.path.join(GOLDEN, "cases_xdmf", "*.in")) +
                    glob.glob(os.path.join(GOLDEN, "reference_inputs", "*.in")))


def dump(path):
    out = subprocess.run([RUN, path, "--dump-ctrl"], check=True, capture_output=True, text=True).stdout
    d = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        d.setdefault(k, []).append(v.split())
    return d


def _f(v):
    return [float(x) for x in v]


@pytest.mark.parametrize("path", CTRL_FILES, ids=[os.path.basename(p) for p in CTRL_FILES])
def test_ctrl_reader_matches_python(path):
    if not os.path.exists(RUN):
        pytest.fail("mmca_run is not built: run __graft_entry__.build()")
    d = dump(path)
    cfg = read_ctrl(path)
    p = cfg.params()
    n = cfg.nucleation()
    assert d["basefilename"][0][0] == cfg.basefilename
    assert [int(x) for x in d["nsub"][0]] == list(cfg.nsub)
    assert float(d["finish_time"][0][0]) == cfg.finish_time
    assert float(d["fs_finish"][0][0]) == cfg.fs_finish
    assert int(d["screenpfreq"][0][0]) == cfg.screenpfreq
    assert int(d["data_write_freq"][0][0]) == cfg.data_write_freq
    assert int(d["data_write_start"][0][0]) == cfg.data_write_start
    assert d["data_format"][0][0] == cfg.data_format
    assert int(d["randomize_grains"][0][0]) == cfg.randomize_grains
    for k in ("nx", "ny", "nz", "temp_model", "am_layers", "am_tracks", "gnsbx", "max_total_grains"):
        assert int(d[k][0][0]) == getattr(p, k), k
    for k in ("cell_size", "dt", "courant_limit", "c_init", "liquidus_temp", "solidus_temp", "part_coef",
              "d_sol", "d_liq", "fsgrow", "t_init", "cool_rate", "gradient", "grad_coef", "velocity",
              "velo_coef"):
        assert float(d[k][0][0]) == getattr(p, k), k
    assert [int(x) for x in d["pad_bndy"][0]] == list(p.pad_bndy)
    assert [int(x) for x in d["gdim"][0]] == list(p.gdim)
    assert _f(d["m_solute"][0]) == list(p.m_solute)
    assert _f(d["laser"][0]) == [p.laser_power, p.laser_speed, p.laser_spot, p.laser_absorp, p.laser_hatch,
                                 p.sub_init_temp, p.th_conductivity, p.density, p.specific_heat]
    assert _f(d["laser_start"][0]) == list(p.laser_start)
    with open(path) as fh:
        if any(l.split()[:1] and l.split()[0].lower() == "randseedval" for l in fh):
            assert int(d["random_seed"][0][0]) == n.random_seed
    assert int(d["oriented"][0][0]) == n.oriented
    assert int(d["do_surface"][0][0]) == n.do_surface
    assert _f(d["bulk"][0]) == [n.bulk_density, n.bulk_mean, n.bulk_sigma]
    assert _f(d["surf"][0]) == [n.surf_density, n.surf_mean, n.surf_sigma]
    fixed = [_f(v) for v in d.get("fixed", [])]
    assert fixed == [list(t) for t in cfg.fixed]
    vols = [_f(v) for v in d.get("volume", [])]
    assert vols == [list(t) for t in cfg.volumes]


def test_ctrl_reader_rejects_mismatched_faces(tmp_path):
    f = tmp_path / "bad.in"
    f.write_text("NCellsPerSB 8 8 8\nFaceCtrl 1 0 1 1 1 1\n")
    r = subprocess.run([RUN, str(f), "--dump-ctrl"], capture_output=True, text=True)
    assert r.returncode != 0 and "FaceCtrl" in r.stderr


@pytest.mark.gpu
def test_one_grain_through_c_driver(tmp_path):
    path = os.path.join(GOLDEN, "reference_inputs", "one_grain.in")
    r = subprocess.run([RUN, path], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if "Average" in l]
    assert lines == ["1 grains.  Min: 0.000737875 mm^3   Max: 0.000737875 mm^3   Average:  0.000737875 mm^3"]
    # FsFinish 0.01 ends the run at step 9000 (t = 9 s), as in the reference
    assert "Solid Fraction termination condition met." in r.stdout
    screen = [l.split() for l in r.stdout.splitlines() if "%" in l]
    assert float(screen[-1][1]) == 9.0
    cfg = read_ctrl(path)
    csv = (tmp_path / ("%s_grains.csv" % cfg.basefilename)).read_text().splitlines()
    assert csv[0].startswith("Number,Nuc Timestep")
    assert len(csv) == 2
    size = float(csv[1].split(",")[-1])
    assert math.isclose(size, 0.000737875, rel_tol=1e-6)


@pytest.mark.parametrize("path", CTRL_FILES, ids=[os.path.basename(p) for p in CTRL_FILES])
def test_write_ctrl_round_trips(path, tmp_path):
    """write_ctrl (used to hand a workload to the reference's CPU build) reads
    back to the same parameters, also under another subblock split."""
    from mumatscale_amd.ctrl import write_ctrl
    cfg = read_ctrl(path)
    splits = [list(cfg.nsub)]
    if cfg.gdim[2] % (2 * cfg.nsub[2]) == 0:
        splits.append([cfg.nsub[0], cfg.nsub[1], 2 * cfg.nsub[2]])
    for nsub in splits:
        out = tmp_path / "w.in"
        write_ctrl(cfg, str(out), nsub=nsub)
        back = read_ctrl(str(out))
        assert list(back.nsub) == nsub
        assert list(back.gdim) == list(cfg.gdim)
        a, b = cfg.params(), back.params()
        for name, _ in a._fields_:
            if name in ("nx", "ny", "nz", "gnsbx"):
                continue
            va, vb = getattr(a, name), getattr(b, name)
            if hasattr(va, "__len__"):
                va, vb = list(va), list(vb)
            assert va == vb, name
        na, nb = cfg.nucleation(), back.nucleation()
        for name, _ in na._fields_:
            if name not in ("fixed", "volumes"):
                assert getattr(na, name) == getattr(nb, name), name
        assert [list(t) for t in back.fixed] == [list(t) for t in cfg.fixed]
        assert [list(t) for t in back.volumes] == [list(t) for t in cfg.volumes]
        assert (back.finish_time, back.fs_finish, back.basefilename) == (cfg.finish_time, cfg.fs_finish,
                                                                          cfg.basefilename)


def _run_reference_input(name, tmp_path, timeout=900):
    path = os.path.join(GOLDEN, "reference_inputs", name)
    r = subprocess.run([RUN, path], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    avg = [l for l in r.stdout.splitlines() if "Average" in l]
    assert len(avg) == 1
    w = avg[0].split()
    screen = [l.split() for l in r.stdout.splitlines() if len(l.split()) == 4 and l.split()[2].endswith("%")]
    return int(w[0]), float(w[9]), screen


@pytest.mark.gpu
def test_multigrains_through_c_driver(tmp_path):
    """The reference's MultiGrains regression (tests/testMultiGrains.py:25-37,
    ca6mpi.in on 6 MPI ranks): 336 grains, average 7.14e-5 mm^3 within 1e-6.
    The driver runs the 1 x 2 x 3 subblocks as 6 coupled ranks on one GPU."""
    ngrains, avg, _ = _run_reference_input("ca6mpi.in", tmp_path)
    assert ngrains == 336
    assert abs(avg - 7.14e-5) <= 1e-6


@pytest.mark.gpu
def test_amtrack_through_c_driver(tmp_path):
    """The reference's AMtrack regression (tests/testAMtrack.py:25-56,
    am_singletrack.in on 6 MPI ranks): 1889 grains, average 9.1e-7 mm^3
    within 1e-8, last solid fraction 86 % within 0.2."""
    ngrains, avg, screen = _run_reference_input("am_singletrack.in", tmp_path)
    fs = float(screen[-1][2][:-1])
    print("grains", ngrains, "average", avg, "fs", fs)
    assert ngrains == 1889
    assert abs(avg - 9.1e-7) <= 1e-8
    assert abs(fs - 86.0) <= 0.2


@pytest.mark.gpu
def test_courant_abort_through_c_driver(tmp_path):
    """A dt beyond the Courant limit stops the run at the first step with the
    reference's message (growth.c:43-50) and a failing exit status."""
    path = os.path.join(GOLDEN, "cases", "c5_courant.in")
    r = subprocess.run([RUN, path], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode != 0
    assert "step 1: Instability by Courant criterion" in r.stderr
