"""The device's exp (the Gaussian-beam quadrature's, mmca_exp.h) against the
host C library's exp, bit for bit, over 1.1e8 arguments.

The reference's beam temperature sums 400 glibc exp terms per cell
(temperature.c:104-112), so the device's exp must give glibc's bits for the
beam-on trajectory to stay bit-identical.  Arguments: the quadrature's own
range (every exponent it forms is <= 0, down to results that underflow),
log-spaced magnitudes up to 1e6, a band around 0, arguments next to the
round-to-integer boundaries of x*128/ln2, and special values.
"""
import numpy as np
import pytest

from mumatscale_amd import lib
from oracle.oracle import host_exp

pytestmark = pytest.mark.gpu

CHUNK = 10_000_000


def _boundary_args(n, rng):
    k = rng.integers(-95000, 95000, n).astype(np.float64)
    x0 = (k + 0.5) / (float.fromhex("0x1.71547652b82fep0") * 128)
    d = rng.integers(-6, 7, n)
    return (x0.view(np.int64) + d).view(np.float64)


def _chunks(rng):
    for _ in range(6):                                       # the quadrature's range
        yield "uniform [-746, 0]", rng.uniform(-746.0, 0.0, CHUNK)
    for _ in range(2):
        yield "-10^U[-20, 6]", -np.power(10.0, rng.uniform(-20.0, 6.0, CHUNK))
    yield "uniform [-2, 2]", rng.uniform(-2.0, 2.0, CHUNK)
    yield "uniform [-760, 720]", rng.uniform(-760.0, 720.0, CHUNK)
    yield "round-to-int boundaries", _boundary_args(CHUNK, rng)
    yield "special", np.array([0.0, -0.0, 5e-324, -5e-324, 2.0 ** -54, -(2.0 ** -54), 511.9, 512.0,
                               -512.0, 709.78, 709.79, -708.4, -745.13, -745.14, -1023.9, -1024.0,
                               -1e300, 1e300, np.inf, -np.inf, np.nan])


def test_device_exp_matches_host_exp():
    rng = np.random.default_rng(20261203)
    total = 0
    for name, x in _chunks(rng):
        y_dev = lib.eval_exp(x)
        y_host = host_exp(x)
        bad = np.flatnonzero(y_dev.view(np.uint64) != y_host.view(np.uint64))
        if bad.size:
            i = bad[0]
            pytest.fail("%s: %d of %d differ, first x=%r device %r host %r"
                        % (name, bad.size, x.size, x[i], y_dev[i], y_host[i]))
        total += x.size
        print(name, x.size, "equal", flush=True)
    assert total >= 100_000_000
