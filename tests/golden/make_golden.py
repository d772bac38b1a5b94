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

Control files in tests/golden/cases_mr/ split the block into several
subblocks; each runs on as many MPI ranks (MPICH's mpiexec from the image)
and keeps, per rank and step, the hook's interior hashes:

  <case>/ihash.txt.gz      "rank ts num_grains gindex" and word hashes of the
                           interior ce, cl, fs, gr, dc, of the rank's
                           interface list and of the (shared) grain table
  <case>/expected.json     as above

Control files in tests/golden/cases_xdmf/ ask for the reference's XDMF
output (DataFormat XDMF); they run on their NSubblocks ranks, with the hook's
per-rank hashes left out, and keep a digest of the files written
(tests/output_digest.py):

  <case>/digest.json.gz    the time series text (without its date line), the
                           structure of every HDF5 file and a hash of each
                           dataset
  <case>/expected.json     as above

Control files in tests/golden/cases_vtk/ ask for DataFormat VTK:

  <case>/vtk.json          a hash of every .pvtu and .vtu file written
  <case>/expected.json     as above

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

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import output_digest  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CASES = os.path.join(HERE, "cases")
CASES_MR = os.path.join(HERE, "cases_mr")
CASES_XDMF = os.path.join(HERE, "cases_xdmf")
CASES_VTK = os.path.join(HERE, "cases_vtk")
CASES_CKPT = os.path.join(HERE, "cases_ckpt")
MPIEXEC = "/opt/conda/bin/mpiexec"
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


def nranks(ctrl):
    for line in open(ctrl):
        w = line.split()
        if w and w[0] == "NSubblocks":
            return int(w[1]) * int(w[2]) * int(w[3])
    return 1


def run_case_mr(name, out_dir):
    ctrl = os.path.join(CASES_MR, name + ".in")
    n = nranks(ctrl)
    with tempfile.TemporaryDirectory() as tmp:
        env = dict(os.environ, MMCA_REF_DUMP_DIR=tmp, OMP_NUM_THREADS="1")
        proc = subprocess.run([MPIEXEC, "-hosts", "127.0.0.1", "-n", str(n), REF_BIN, "-c", ctrl], cwd=tmp,
                              env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=3600)
        text = proc.stdout.decode(errors="replace")
        rows = []
        for r in range(n):
            for line in open(os.path.join(tmp, "ihash_r%02d.txt" % r)).read().splitlines():
                rows.append("%d %s" % (r, line))
        os.makedirs(out_dir, exist_ok=True)
        with gzip.open(os.path.join(out_dir, "ihash.txt.gz"), "wt") as fh:
            fh.write("\n".join(rows) + "\n")
        avg = [l for l in text.splitlines() if "Average" in l]
        err = [l.strip() for l in text.splitlines() if "ERROR" in l]
        exp = {
            "returncode": proc.returncode,
            "ranks": n,
            "last_step": max(int(r.split()[1]) for r in rows),
            "error_line": err[0] if err else None,
            "average_line": avg[0].strip() if avg else None,
        }
        with open(os.path.join(out_dir, "expected.json"), "w") as fh:
            json.dump(exp, fh, indent=1)
        return exp


def base_name(ctrl):
    for line in open(ctrl):
        w = line.split()
        if w and w[0] == "BaseFileName":
            return w[1]
    raise ValueError(ctrl)


def run_case_xdmf(name, out_dir):
    ctrl = os.path.join(CASES_XDMF, name + ".in")
    n, base = nranks(ctrl), base_name(ctrl)
    with tempfile.TemporaryDirectory() as tmp:
        if os.path.dirname(base):
            os.makedirs(os.path.join(tmp, os.path.dirname(base)))
        env = dict(os.environ, MMCA_REF_DUMP_DIR=os.path.join(tmp, "hook"), OMP_NUM_THREADS="1")
        os.makedirs(env["MMCA_REF_DUMP_DIR"])
        proc = subprocess.run([MPIEXEC, "-hosts", "127.0.0.1", "-n", str(n), REF_BIN, "-c", ctrl], cwd=tmp,
                              env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=3600)
        text = proc.stdout.decode(errors="replace")
        digest = output_digest.digest_output(tmp, base)
        os.makedirs(out_dir, exist_ok=True)
        with gzip.open(os.path.join(out_dir, "digest.json.gz"), "wt") as fh:
            json.dump(digest, fh, indent=0, sort_keys=True)
        avg = [l for l in text.splitlines() if "Average" in l]
        err = [l.strip() for l in text.splitlines() if "ERROR" in l]
        steps = [int(os.path.basename(f)[:-3]) for f in digest["files"] if "_r0/" in f]
        exp = {
            "returncode": proc.returncode,
            "ranks": n,
            "last_step": max(steps),
            "error_line": err[0] if err else None,
            "average_line": avg[0].strip() if avg else None,
        }
        with open(os.path.join(out_dir, "expected.json"), "w") as fh:
            json.dump(exp, fh, indent=1)
        return exp


def run_case_ckpt(name, out_dir):
    """The reference's checkpoint files for a control file of cases_ckpt/.
    Its main-file writer fails on every run that writes checkpoints (it
    reads gtasks.task, which nothing allocates, checkpoint.c:1391) before
    any file is complete; MMCA_REF_TASK_CHECKPOINT makes the link-time hook
    skip it, so each rank's writeTaskCheckpoint (checkpoint.c:955-1238)
    writes its file, rank 0's under the main file's name, which a working
    main writer's file would carry too (rank 0 writes its task file to the
    same path right after).  The digest covers the files left after the
    last checkpoint step."""
    ctrl = os.path.join(CASES_CKPT, name + ".in")
    n, base = nranks(ctrl), base_name(ctrl)
    with tempfile.TemporaryDirectory() as tmp:
        env = dict(os.environ, MMCA_REF_TASK_CHECKPOINT="1", OMP_NUM_THREADS="1")
        proc = subprocess.run([MPIEXEC, "-hosts", "127.0.0.1", "-n", str(n), REF_BIN, "-c", ctrl], cwd=tmp,
                              env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=3600)
        text = proc.stdout.decode(errors="replace")
        digest = output_digest.checkpoint_digest(tmp)
        os.makedirs(out_dir, exist_ok=True)
        with gzip.open(os.path.join(out_dir, "digest.json.gz"), "wt") as fh:
            json.dump(digest, fh, indent=0, sort_keys=True)
        steps = [int(l.split()[-1]) for l in text.splitlines() if l.startswith("Checkpoint at timestep")]
        exp = {"returncode": proc.returncode, "ranks": n, "checkpoint_steps": sorted(set(steps)),
               "files": sorted(digest)}
        with open(os.path.join(out_dir, "expected.json"), "w") as fh:
            json.dump(exp, fh, indent=1)
        return exp


def run_case_vtk(name, out_dir):
    ctrl = os.path.join(CASES_VTK, name + ".in")
    n, base = nranks(ctrl), base_name(ctrl)
    with tempfile.TemporaryDirectory() as tmp:
        if os.path.dirname(base):
            os.makedirs(os.path.join(tmp, os.path.dirname(base)))
        env = dict(os.environ, MMCA_REF_DUMP_DIR=os.path.join(tmp, "hook"), OMP_NUM_THREADS="1")
        os.makedirs(env["MMCA_REF_DUMP_DIR"])
        proc = subprocess.run([MPIEXEC, "-hosts", "127.0.0.1", "-n", str(n), REF_BIN, "-c", ctrl], cwd=tmp,
                              env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=3600)
        text = proc.stdout.decode(errors="replace")
        files = output_digest.vtk_digest(tmp, base)
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "vtk.json"), "w") as fh:
            json.dump(files, fh, indent=0, sort_keys=True)
        avg = [l for l in text.splitlines() if "Average" in l]
        err = [l.strip() for l in text.splitlines() if "ERROR" in l]
        steps = [int(f.rsplit("_", 1)[1][:-5]) for f in files if f.endswith(".pvtu")]
        exp = {
            "returncode": proc.returncode,
            "ranks": n,
            "last_step": max(steps),
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
    names = sorted(f[:-3] for f in os.listdir(CASES_MR) if f.endswith(".in"))
    if len(sys.argv) > 1:
        names = [n for n in names if n in sys.argv[1:]]
    for n in names:
        out = os.path.join(HERE, n)
        if os.path.isdir(out):
            shutil.rmtree(out)
        exp = run_case_mr(n, out)
        print(n, exp)
    names = sorted(f[:-3] for f in os.listdir(CASES_XDMF) if f.endswith(".in"))
    if len(sys.argv) > 1:
        names = [n for n in names if n in sys.argv[1:]]
    for n in names:
        out = os.path.join(HERE, n)
        if os.path.isdir(out):
            shutil.rmtree(out)
        exp = run_case_xdmf(n, out)
        print(n, exp)
    names = sorted(f[:-3] for f in os.listdir(CASES_VTK) if f.endswith(".in"))
    if len(sys.argv) > 1:
        names = [n for n in names if n in sys.argv[1:]]
    for n in names:
        out = os.path.join(HERE, n)
        if os.path.isdir(out):
            shutil.rmtree(out)
        exp = run_case_vtk(n, out)
        print(n, exp)
    names = sorted(f[:-3] for f in os.listdir(CASES_CKPT) if f.endswith(".in"))
    if len(sys.argv) > 1:
        names = [n for n in names if n in sys.argv[1:]]
    for n in names:
        out = os.path.join(HERE, n)
        if os.path.isdir(out):
            shutil.rmtree(out)
        exp = run_case_ckpt(n, out)
        print(n, exp)


if __name__ == "__main__":
    main()
