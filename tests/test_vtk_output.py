"""VTK output of the C host driver (DataFormat VTK) against the reference's
own files for the same control file (tests/golden/cases_vtk/, fixtures from
tests/golden/make_golden.py): every .pvtu and .vtu file must be
byte-identical (io_vtk.c, vtkCompress.c: FP64 temperature and grain ids as
zlib blocks per x slab behind a block-size header, base64 encoded, plus the
point coordinates and the hexahedral cells of the subblock).

  v1  2 ranks along x, files every 50 steps and after the last step
  v2  the 1 x 2 x 3 MultiGrains split, files every 100 steps"""
import json
import os
import subprocess

import pytest

import output_digest
from golden_cases import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mumatscale_amd", "mmca_run")
CASES = sorted(f[:-3] for f in os.listdir(os.path.join(GOLDEN, "cases_vtk")) if f.endswith(".in"))


def load(case):
    ctrl = os.path.join(GOLDEN, "cases_vtk", case + ".in")
    files = json.load(open(os.path.join(GOLDEN, case, "vtk.json")))
    expected = json.load(open(os.path.join(GOLDEN, case, "expected.json")))
    base = next(l.split()[1] for l in open(ctrl) if l.split()[:1] == ["BaseFileName"])
    return ctrl, base, files, expected


@pytest.mark.parametrize("case", CASES)
def test_fixture_lists_every_rank_and_step(case):
    """One .pvtu per written step and one .vtu per rank and step
    (vtk_writeMain / vtk_writeSubblocks, io_vtk.c:455-492)."""
    ctrl, base, files, expected = load(case)
    keys = {l.split()[0]: l.split()[1:] for l in open(ctrl) if l.split() and not l.startswith("#")}
    freq = int(keys["DataWriteFreq"][0])
    last = expected["last_step"]
    steps = sorted({0, last} | set(range(freq, last + 1, freq)))
    want = {"%s_%d.pvtu" % (base, t) for t in steps}
    want |= {"%s_%04d/%d.vtu" % (base, r, t) for r in range(expected["ranks"]) for t in steps}
    assert set(files) == {os.path.normpath(f) for f in want}


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_vtk_output_matches_reference(case, tmp_path):
    if not os.path.exists(RUN):
        pytest.fail("mmca_run is not built: run __graft_entry__.build()")
    ctrl, base, want, expected = load(case)
    if os.path.dirname(base):
        os.makedirs(tmp_path / os.path.dirname(base))
    print("case", case, RUN, ctrl, flush=True)
    r = subprocess.run([RUN, ctrl], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert expected["average_line"] in r.stdout
    got = output_digest.vtk_digest(str(tmp_path), base)
    bad = sorted(f for f in set(got) | set(want) if got.get(f) != want.get(f))
    assert not bad, bad[:20]
