"""XDMF + HDF5 output of the C host driver (DataFormat XDMF) against the
reference's own files for the same control file (tests/golden/cases_xdmf/,
fixtures from tests/golden/make_golden.py, compared through the digests of
tests/output_digest.py).

Every HDF5 file must have the reference's structure (datasets, types, shapes,
chunking, shuffle + deflate 9, fill time, and so the same stored size) and
the same bytes in every dataset: the fields as cache_io_data converts them
(loop.c:74-113), the grain angles and the renumbering of io_xdmf.c:345-376,
whose rand() draws come from rank 0's stream between the orientations that
rank 0 draws for its new grains.  The time series must be the same text
apart from its wall-clock date line.

  x1  6 ranks (1 x 2 x 3), files every 150 steps and after the last step
  x2  one block, RandomizeGrains 0
  x3  2 ranks, grains nucleating while files are written every 20 steps
  x4  x3 with DataWriteStart 100; the reference ends in its checkpoint
      writer after step 100, so the run stops there"""
import gzip
import json
import os
import subprocess

import pytest

import output_digest
from golden_cases import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mumatscale_amd", "mmca_run")
CASES = sorted(f[:-3] for f in os.listdir(os.path.join(GOLDEN, "cases_xdmf")) if f.endswith(".in"))


def load(case):
    ctrl = os.path.join(GOLDEN, "cases_xdmf", case + ".in")
    with gzip.open(os.path.join(GOLDEN, case, "digest.json.gz"), "rt") as fh:
        digest = json.load(fh)
    expected = json.load(open(os.path.join(GOLDEN, case, "expected.json")))
    base = next(l.split()[1] for l in open(ctrl) if l.split()[:1] == ["BaseFileName"])
    return ctrl, base, digest, expected


def write_steps(case):
    """Steps the reference wrote, from the rank-0 file names of the fixture."""
    _, base, digest, _ = load(case)
    return sorted(int(f[-8:-3]) for f in digest["files"] if f.startswith(base + "_r0/"))


def test_writer_library_exports_its_entry_points():
    """libmmca_io.so (loaded by mmca_run for DataFormat XDMF or VTK) is built
    and exports what mmca_xdmf.h and mmca_vtk.h declare."""
    import ctypes
    import re
    so = os.path.join(ROOT, "mumatscale_amd", "libmmca_io.so")
    if not os.path.exists(so):
        pytest.fail("libmmca_io.so is not built: run __graft_entry__.build()")
    names = []
    for h in ("mmca_xdmf.h", "mmca_vtk.h"):
        hdr = open(os.path.join(ROOT, "mumatscale_amd", "csrc", "host", h)).read()
        names += re.findall(r"(mmca_(?:xdmf|vtk)_\w+)\s*\(", hdr)
    assert len(set(names)) == 9
    L = ctypes.CDLL(so)
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("case", CASES)
def test_fixture_is_a_complete_output(case):
    """Every rank has a file for every written step, the time series names
    exactly those files, and the step set follows loop.c:201-254 (step 0,
    multiples of DataWriteFreq from DataWriteStart on, the last step)."""
    ctrl, base, digest, expected = load(case)
    keys = {l.split()[0]: l.split()[1:] for l in open(ctrl) if l.split() and not l.startswith("#")}
    freq = int(keys["DataWriteFreq"][0])
    start = int(keys.get("DataWriteStart", ["0"])[0])
    last = expected["last_step"]
    steps = [0] + [t for t in range(1, last + 1) if t >= start and t % freq == 0]
    if expected["returncode"] == 0 and last % freq:
        steps.append(last)
    assert write_steps(case) == steps
    local = os.path.basename(base)
    for r in range(expected["ranks"]):
        for t in steps:
            f = "%s_r%d/%05d.h5" % (base, r, t)
            assert f in digest["files"], f
            assert "SERIAL:%s_r%d/%05d.h5:" % (local, r, t) in digest["xmf"]
    assert digest["xmf"].count("<Grid Name=\"Full Volume") == len(steps)
    gbl = digest["files"][base + "_gbl.h5"]["data"]
    assert ("/GrainRenumber" in gbl) == (keys.get("RandomizeGrains", ["1"])[0] != "0")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_xdmf_output_matches_reference(case, tmp_path):
    if not os.path.exists(RUN):
        pytest.fail("mmca_run is not built: run __graft_entry__.build()")
    if not output_digest.available():
        pytest.skip("h5dump is not installed")
    ctrl, base, want, expected = load(case)
    if os.path.dirname(base):
        os.makedirs(tmp_path / os.path.dirname(base))
    cmd = [RUN, ctrl]
    if expected["returncode"] != 0:
        cmd += ["--max-steps", str(expected["last_step"])]
    print("case", case, " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    if expected["average_line"]:
        assert expected["average_line"] in r.stdout
    got = output_digest.digest_output(str(tmp_path), base)
    bad = output_digest.compare(got, want)
    assert not bad, bad[:20]


@pytest.mark.gpu
@pytest.mark.parametrize("case,stop", [("x2_directional_one_block", 200), ("x3_cooling_x2", 120)])
def test_restarted_run_writes_the_same_files(case, stop, tmp_path):
    """writeCheckpoint / restart through the checkpoint files (the
    reference's layout, with this build's restart extras in /MMCA): a run
    stopped at `stop` and restarted with --restart writes the reference's
    files for every step, byte for byte (the later ones from the restarted
    run), and ends with its grain summary.  The time series of the restarted run starts afresh, as the
    reference's does after a restart, and holds the later entries only."""
    if not os.path.exists(RUN):
        pytest.fail("mmca_run is not built: run __graft_entry__.build()")
    if not output_digest.available():
        pytest.skip("h5dump is not installed")
    ctrl, base, want, expected = load(case)
    if os.path.dirname(base):
        os.makedirs(tmp_path / os.path.dirname(base))
    half = stop // 2
    for cmd in ([RUN, ctrl, "--max-steps", str(stop), "--checkpoint-freq", str(half)],
                [RUN, ctrl, "--restart", "--checkpoint-freq", "0"]):
        print("case", case, " ".join(cmd), flush=True)
        r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=600)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
    assert "Restarting from timestep %d" % stop in r.stdout
    assert expected["average_line"] in r.stdout
    got = output_digest.digest_output(str(tmp_path), base)
    later = [t for t in write_steps(case) if t > stop]
    bad = output_digest.compare(dict(got, xmf=""), dict(want, xmf=""))   # every HDF5 file
    assert not bad, bad[:20]
    head = got["xmf"].splitlines()
    n0 = next(i for i, l in enumerate(head) if 'Name="Time Series"' in l) + 1
    assert head[n0:] == want["xmf"].splitlines()[-(len(head) - n0):]
    assert got["xmf"].count("<Grid Name=\"Full Volume") == len(later)


def test_digest_reads_dataset_sizes_from_the_header():
    """The digest splits h5dump's binary dump by the datasets its header
    lists, in order; the parser must find nested groups, types and shapes."""
    header = "\n".join([
        'HDF5 "f.h5" {',
        'GROUP "/" {',
        '   DATASET "GrainAngleX" {',
        '      DATATYPE  H5T_IEEE_F32LE',
        '      DATASPACE  SIMPLE { ( 29 ) / ( 29 ) }',
        '      STORAGE_LAYOUT {',
        '         CHUNKED ( 29 )',
        '      }',
        '   }',
        '   GROUP "SubBlock_0x1x2" {',
        '      DATASET "Grain" {',
        '         DATATYPE  H5T_STD_I32LE',
        '         DATASPACE  SIMPLE { ( 12, 12, 24 ) / ( 12, 12, 24 ) }',
        '         FILTERS {',
        '            COMPRESSION DEFLATE { LEVEL 9 }',
        '         }',
        '      }',
        '   }',
        '   DATASET "GrainRenumber" {',
        '      DATATYPE  H5T_STD_U32LE',
        '      DATASPACE  SIMPLE { ( 29 ) / ( 29 ) }',
        '   }',
        '}',
        '}'])
    assert output_digest._datasets(header) == [("/GrainAngleX", 4 * 29), ("/SubBlock_0x1x2/Grain", 4 * 12 * 12 * 24),
                                               ("/GrainRenumber", 4 * 29)]
