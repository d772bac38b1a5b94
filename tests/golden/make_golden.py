"""Write the golden fixtures of tests/golden/ from the reference itself.

Builds the reference CPU code with oracle/ref.mk (into oracle/_ref/), runs it
on one rank for every control file in tests/golden/cases/, and keeps what its
dump hook (oracle/ref_hook.c) writes:

  <case>/hashes.txt.gz     per-step FNV-1a 64 of ce, cl, fs, gr, dc, d, the
                           interface list and the grain table
  <case>/init_sites.npz    nucleation thresholds after initialisation
                           (padded index and float32 value of every site)
  <case>/expected.json     exit status, last completed step, the reference's
                           "N grains ... Average" line

Run in the container that holds the reference sources:
    python tests/golden/make_golden.py
"""
import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CASES = os.path.join(HERE, "cases")
REF_BIN = os.path.join(REPO, "oracle", "_ref", "mumatscale_ref")


def read_state0(path):
    b = open(path, "rb").read()
    assert b[:8] == b"MMCAREF1"
    hdr = np.frombuffer(b, np.int32, 8, 8)
    nx, ny, nz = (int(v) for v in hdr[:3])
    P = (nx + 2) * (ny + 2) * (nz + 2)
    off = 48 + P * (8 + 8 + 8 + 4 + 24 + 8 + 8)     # ce cl fs gr dc d temperature
    thr = np.frombuffer(b, np.float32, P, off)
    idx = np.nonzero(np.isfinite(thr))[0].astype(np.int64)
    return idx, thr[idx].copy()


def run_case(name, out_dir):
    ctrl = os.path.join(CASES, name + ".in")
    with tempfile.TemporaryDirectory() as tmp:
        env = dict(os.environ, MMCA_REF_DUMP_DIR=tmp, MMCA_REF_DUMP_STEPS="0",
                   OMP_NUM_THREADS="1")
        env.pop("OMP_NUM_THREADS")
        proc = subprocess.run([REF_BIN, "-c", ctrl], cwd=tmp, env=env,
                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=1800)
        text = proc.stdout.decode(errors="replace")
        lines = open(os.path.join(tmp, "hashes.txt")).read().splitlines()
        os.makedirs(out_dir, exist_ok=True)
        with gzip.open(os.path.join(out_dir, "hashes.txt.gz"), "wt") as fh:
            fh.write("\n".join(lines) + "\n")
        idx, thr = read_state0(os.path.join(tmp, "state_000000.bin"))
        np.savez_compressed(os.path.join(out_dir, "init_sites.npz"), idx=idx, threshold=thr)
        avg = [l for l in text.splitlines() if "Average" in l]
        err = [l.strip() for l in text.splitlines() if "ERROR" in l]
        exp = {
            "returncode": proc.returncode,
            "last_step": int(lines[-1].split()[0]),
            "error_line": err[0] if err else None,
            "average_line": avg[0].strip() if avg else None,
        }
        with open(os.path.join(out_dir, "expected.json"), "w") as fh:
            json.dump(exp, fh, indent=1)
        return exp


def main():
    subprocess.check_call(["make", "-s", "-f", os.path.join(REPO, "oracle", "ref.mk")], cwd=REPO)
    names = sorted(f[:-3] for f in os.listdir(CASES) if f.endswith(".in"))
    if len(sys.argv) > 1:
        names = [n for n in names if n in sys.argv[1:]]
    for n in names:
        out = os.path.join(HERE, n)
        if os.path.isdir(out):
            shutil.rmtree(out)
        exp = run_case(n, out)
        print(n, exp)


if __name__ == "__main__":
    main()
