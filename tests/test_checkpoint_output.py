"""Checkpoint files of the C host driver against the reference's own
(tests/golden/cases_ckpt/, fixtures from tests/golden/make_golden.py, compared
through the digests of tests/output_digest.py).

The reference's writeTaskCheckpoint (checkpoint.c:955-1238) leaves, per rank,
cp_main.h5 (rank 0) or cp_task_<rank>.h5 with /curTimeStep, /g_ts_idx,
/grainCache, /numGrain and /SUBBLOCK/SubBlock_XxYxZ/{CoordTemp, temperature,
gr, fs, ce, oce, cl, diff_id, mold, d, nuc_threshold, dc, nuc_pt}.  Every file
must have the reference's structure (datasets, types, shapes, chunking,
shuffle + deflate 9, stored sizes, offsets of the unchunked datasets) and the
same values in every dataset: the fields after the last checkpoint step, oce
(ce before that step's diffusion), d as the reference's array holds it
(cells that melted keep their last value), the interface-list array's first
padded-cell-count bytes, the grain table of MaxTotGrains + 1 slots.  The
driver's cp_main.h5 also has a group /MMCA with what the reference does not
store (random streams, dc_tmp slots, the list); it is left out of the
comparison.  A run stopped at an earlier checkpoint and restarted from these
files must leave the reference's final files.

  k1  one block, directional, checkpoints at 150 and 300
  k2  2 ranks along x, cooling, checkpoints at 100 and 200
  k3  3 ranks around a wrapped x axis with a fixed nucleus (rank 0 holds
      nuc_pt), checkpoints at 100, 200, 300"""
import gzip
import json
import os
import subprocess

import pytest

import output_digest
from golden_cases import GOLDEN

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "mumatscale_amd", "mmca_run")
CASES = sorted(f[:-3] for f in os.listdir(os.path.join(GOLDEN, "cases_ckpt")) if f.endswith(".in"))


def load(case):
    ctrl = os.path.join(GOLDEN, "cases_ckpt", case + ".in")
    with gzip.open(os.path.join(GOLDEN, case, "digest.json.gz"), "rt") as fh:
        digest = json.load(fh)
    expected = json.load(open(os.path.join(GOLDEN, case, "expected.json")))
    base = next(l.split()[1] for l in open(ctrl) if l.split()[:1] == ["BaseFileName"])
    return ctrl, base, digest, expected


def run(cmd, cwd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return r


def check_files(tmp_path, want):
    got = output_digest.checkpoint_digest(str(tmp_path), drop=("MMCA",))
    bad = output_digest.compare_files(got, want)
    assert not bad, bad[:20]


@pytest.mark.parametrize("case", CASES)
def test_checkpoint_files_match_reference(case, tmp_path):
    if not output_digest.available():
        pytest.skip("h5dump is not installed")
    ctrl, base, want, expected = load(case)
    run([RUN, ctrl], tmp_path)
    check_files(tmp_path, want)


@pytest.mark.parametrize("case", CASES)
def test_restart_from_checkpoint_files(case, tmp_path):
    """Stopped at the second-last checkpoint, restarted from its files (the
    fields read from the reference's datasets, the rest from /MMCA): the
    run's final checkpoint files are the reference's."""
    if not output_digest.available():
        pytest.skip("h5dump is not installed")
    ctrl, base, want, expected = load(case)
    steps = expected["checkpoint_steps"]
    run([RUN, ctrl, "--max-steps", str(steps[-2])], tmp_path)
    r = run([RUN, ctrl, "--restart"], tmp_path)
    assert "Restarting from timestep %d" % steps[-2] in r.stdout
    assert "Checkpoint at timestep %d" % steps[-1] in r.stdout
    check_files(tmp_path, want)
