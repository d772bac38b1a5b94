"""The beam quadrature's exp (mumatscale_amd/csrc/mmca_exp.h) against the host
C library's exp, on the CPU.

The reference's Gaussian-beam temperature sums 400 glibc exp values per cell
(temperature.c:104-112).  mmca_exp.h restates glibc's table-driven double exp
so that the device reproduces those bits.  Here the same header is compiled
for the host with g++ and compared with libm's exp in the same process:
random arguments over the whole finite range, the subnormal-result range, and
arguments next to the round-to-integer boundaries of x*128/ln2, where a
wrongly fused or unfused reduction step shows.  A variant that rounds the
product before adding the shift (no fusion) must fail on those boundaries,
which shows that they discriminate.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mumatscale_amd", "csrc")

DRIVER = r"""
#include "mmca_exp.h"
#include <stdio.h>
#include <stdlib.h>

using mmca::exp_as_u64;
using mmca::exp_as_f64;

static uint64_t st = 0x9e3779b97f4a7c15ull;
static inline uint64_t xs(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static inline double unit(void) { return (double) (xs() >> 11) * 0x1p-53; }

/* the reduction without the fused multiply-add (a plausible mistake) */
static double unfused(double x)
{
    const double kd0 = 0x1.71547652b82fep0 * 128 * x;
    double kd = kd0 + 0x1.8p52;
    const uint64_t ki = exp_as_u64(kd);
    kd -= 0x1.8p52;
    const double r = fma(kd, -0x1.cf79abc9e3b3ap-47, fma(kd, -0x1.62e42fefa0000p-8, x));
    const uint32_t idx = 2u * (uint32_t) (ki % 128);
    const double tail = exp_as_f64(mmca::exp_tab[idx]);
    const uint64_t sbits = mmca::exp_tab[idx + 1] + (ki << 45);
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, 0x1.1111167a4d017p-7, 0x1.55555cf172b91p-5),
                           fma(r2, fma(r, 0x1.555555555543cp-3, 0x1.ffffffffffdbdp-2), tail + r));
    const double scale = exp_as_f64(sbits);
    return fma(scale, tmp, scale);
}

int main(int argc, char **argv)
{
    long n = atol(argv[1]);
    long bad_rand = 0, bad_sub = 0, bad_bnd = 0, bad_unfused = 0, n_bnd = 0, bad_spec = 0;
    for (long i = 0; i < n; ++i) {
        double x = -760.0 + 1480.0 * unit();
        bad_rand += exp_as_u64(exp(x)) != exp_as_u64(mmca::exp_glibc(x, mmca::exp_tab));
        double y = -745.2 + 37.5 * unit();                 /* subnormal and near-subnormal results */
        bad_sub += exp_as_u64(exp(y)) != exp_as_u64(mmca::exp_glibc(y, mmca::exp_tab));
    }
    /* x next to (k + 1/2) ln2 / 128: the rounding of x*128/ln2 decides k */
    for (long k = -60000; k < 60000; ++k) {
        const double x0 = (k + 0.5) / (0x1.71547652b82fep0 * 128);
        for (int d = -6; d <= 6; ++d) {
            const double x = exp_as_f64(exp_as_u64(x0) + d);
            if (x == 0.0 || x < -500.0 || x > 500.0)
                continue;
            ++n_bnd;
            const uint64_t g = exp_as_u64(exp(x));
            bad_bnd += g != exp_as_u64(mmca::exp_glibc(x, mmca::exp_tab));
            bad_unfused += g != exp_as_u64(unfused(x));
        }
    }
    const double spec[] = {0.0, -0.0, 1e-300, -1e-300, 0x1p-54, -0x1p-54, 0x1p-55, 511.9, 512.0, -512.0,
                           709.78, 709.79, -708.4, -745.13, -745.14, -1023.9, -1024.0, -1e6, 1e6,
                           __builtin_inf(), -__builtin_inf(), __builtin_nan("")};
    for (double x : spec)
        bad_spec += exp_as_u64(exp(x)) != exp_as_u64(mmca::exp_glibc(x, mmca::exp_tab));
    printf("%ld %ld %ld %ld %ld %ld\n", bad_rand, bad_sub, bad_bnd, bad_unfused, n_bnd, bad_spec);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("exp")
    src = d / "exp_driver.cpp"
    src.write_text(DRIVER)
    exe = d / "exp_driver"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                           "-I" + CSRC, str(src), "-o", str(exe), "-lm"])
    return str(exe)


def test_exp_matches_host_library(driver):
    out = subprocess.check_output([driver, "3000000"], text=True).split()
    bad_rand, bad_sub, bad_bnd, bad_unfused, n_bnd, bad_spec = map(int, out)
    assert n_bnd > 1_000_000
    assert bad_rand == 0
    assert bad_sub == 0
    assert bad_bnd == 0
    assert bad_spec == 0
    # the boundary arguments tell the fused reduction from the unfused one
    assert bad_unfused > 0


def test_exp_table_regenerates():
    """The committed table is what gen_exp_table.py computes from 2^(k/128)."""
    out = subprocess.check_output(["python3", os.path.join(CSRC, "gen_exp_table.py")], text=True)
    with open(os.path.join(CSRC, "mmca_exp_table.inc")) as f:
        assert f.read() == out
    rows = [ln for ln in out.splitlines() if ln.startswith("0x")]
    assert len(rows) == 128
    # k = 0 is exactly 1; k = 64 is sqrt(2) with the exponent bias folded in
    assert rows[0] == "0x0000000000000000ull, 0x3ff0000000000000ull,"
    import struct
    h64 = int(rows[64].split(",")[1].strip().rstrip("ull"), 16) + (64 << 45)
    assert struct.unpack("<d", struct.pack("<Q", h64))[0] == 2 ** 0.5
