"""Helpers shared by the tests: golden-case loading and per-step hashing."""
import ctypes as C
import gzip
import json
import os

import numpy as np

from mumatscale_amd import abi
from mumatscale_amd.ctrl import read_ctrl

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-3] for f in os.listdir(os.path.join(GOLDEN, "cases")) if f.endswith(".in"))
HASH_NAMES = ["ts", "num_grains", "gindex", "ce", "cl", "fs", "gr", "dc", "d", "diff_id", "grains"]
FNV_EMPTY = 1469598103934665603


def ctrl_path(case):
    return os.path.join(GOLDEN, "cases", case + ".in")


def load_case(case):
    cfg = read_ctrl(ctrl_path(case))
    with gzip.open(os.path.join(GOLDEN, case, "hashes.txt.gz"), "rt") as fh:
        hashes = [l.split() for l in fh if l.strip()]
    with open(os.path.join(GOLDEN, case, "expected.json")) as fh:
        expected = json.load(fh)
    sites = np.load(os.path.join(GOLDEN, case, "init_sites.npz"))
    return cfg, hashes, expected, sites


def grain_table_bytes(grains, num_grains):
    """Bytes of grains 1..num_grains-1 as grain_t records (136 B each)."""
    if num_grains <= 1:
        return b""
    arr = (abi.Grain * (num_grains - 1))(*grains[1:num_grains])
    return bytes(C.string_at(C.addressof(arr), C.sizeof(arr)))


def hash_line(fnv, ts, num_grains, gindex, fields, diff_id, grain_bytes):
    """The per-step line oracle/ref_hook.c writes, from numpy state."""
    out = [str(ts), str(num_grains), str(gindex)]
    for name in ("ce", "cl", "fs", "gr", "dc", "d"):
        out.append("%016x" % fnv(fields[name]))
    out.append("%016x" % fnv(np.asarray(diff_id, dtype=np.int32)))
    out.append("%016x" % (fnv(np.frombuffer(grain_bytes, np.uint8)) if grain_bytes else FNV_EMPTY))
    return out
