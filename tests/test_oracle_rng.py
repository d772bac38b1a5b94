"""The oracle's glibc-rand restatement against glibc rand() itself.

The reference draws every random number with rand() after
srand(iproc + 1 + seed) (ca_main.c:215, functions.c:377-445)."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from oracle.oracle import rng_sequence

libc = ctypes.CDLL(ctypes.util.find_library("c"))
libc.srand.argtypes = [ctypes.c_uint]
libc.rand.restype = ctypes.c_int


@pytest.mark.parametrize("seed", [0, 1, 2, 42, 3823798713, 2**31 - 1, 2**31, 2**32 - 1])
def test_rand_stream_matches_glibc(seed):
    n = 5000
    libc.srand(seed)
    want = np.array([libc.rand() for _ in range(n)], dtype=np.int32)
    got = rng_sequence(seed, n)
    assert np.array_equal(got, want)


def test_seed_zero_equals_seed_one():
    # glibc srandom_r maps seed 0 to 1
    assert np.array_equal(rng_sequence(0, 100), rng_sequence(1, 100))
