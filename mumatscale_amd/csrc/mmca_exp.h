// exp(x) in double precision with the same bits as the host C library's exp
// on x86-64: glibc 2.35, whose exp (sysdeps/ieee754/dbl-64/e_exp.c) the
// dynamic loader resolves to its FMA build on CPUs with FMA and AVX2.  The
// reference sums 400 of these per cell for the Gaussian-beam temperature
// (temperature.c:104-112), so the device must reproduce them bit for bit.
//
// Method (the published table-driven one of that file): with N = 128,
//   x = (k/N) ln2 + r,  k = round(x N/ln2),  |r| <= ln2/(2N),
//   exp(x) = 2^(k/N) exp(r) = H (1 + T) exp(r) ~ s + s (T + r + r^2 p(r)),
// where s = 2^(k div N) H_(k mod N) is assembled in the exponent bits from the
// table of mmca_exp_table.inc and p is a degree-3 polynomial.  The FMA build
// fuses every product into its following sum except where the product has a
// use in another basic block (the scaled-down result of the underflow branch);
// the fma() calls below stand exactly where those fusions are, and everything
// else is a separately rounded operation.  Including files are compiled with
// floating-point contraction off.
//
// Pinned on the CPU against the host library over random and round-to-int
// boundary arguments (tests/test_exp_cpu.py) and on the GPU through
// mmca_eval_exp (tests/test_gpu_exp.py).
#pragma once

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define MMCA_HD __host__ __device__
#else
#define MMCA_HD
#endif

namespace mmca {

constexpr int EXP_TAB_BITS = 7;
constexpr int EXP_TAB_N = 1 << EXP_TAB_BITS;
constexpr int EXP_TAB_WORDS = 2 * EXP_TAB_N;

// {bits(T_k), bits(H_k) - (k << 45)} for k = 0..127
#ifdef __HIPCC__
__constant__
#endif
static const uint64_t exp_tab[EXP_TAB_WORDS] = {
#include "mmca_exp_table.inc"
};

MMCA_HD inline uint64_t exp_as_u64(double d) { return __builtin_bit_cast(uint64_t, d); }
MMCA_HD inline double exp_as_f64(uint64_t u) { return __builtin_bit_cast(double, u); }

// |x| >= 512 with a finite result (the scale's exponent may leave the normal range)
MMCA_HD inline double
exp_special(double tmp, uint64_t sbits, uint64_t ki)
{
    if ((ki & 0x80000000ull) == 0)
    {
        // k > 0: the exponent of the scale may have overflowed by <= 460
        sbits -= 1009ull << 52;
        const double scale = exp_as_f64(sbits);
        return 0x1p1009 * fma(scale, tmp, scale);
    }
    // k < 0: result may be subnormal; round to 53 bits before scaling into that range
    sbits += 1022ull << 52;
    const double scale = exp_as_f64(sbits);
    const double st = scale * tmp;
    double y = scale + st;
    if (y < 1.0)
    {
        double lo = scale - y + st;
        const double hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0)
            y = 0.0;
    }
    return 0x1p-1022 * y;
}

// tab: EXP_TAB_WORDS words of mmca_exp_table.inc (constant memory or a copy in LDS)
MMCA_HD inline double
exp_glibc(double x, const uint64_t *tab)
{
    const double inv_ln2_n = 0x1.71547652b82fep0 * EXP_TAB_N;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    const double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double shift = 0x1.8p52;
    const double c2 = 0x1.ffffffffffdbdp-2;
    const double c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5;
    const double c5 = 0x1.1111167a4d017p-7;

    // top 12 bits without the sign; 0x3c9 = top12(2^-54), 0x408 = top12(512), 0x409 = top12(1024)
    uint32_t abstop = (uint32_t) (exp_as_u64(x) >> 52) & 0x7ff;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u)
    {
        if (abstop - 0x3c9u >= 0x80000000u)
            return 1.0 + x;                       // |x| < 2^-54
        if (abstop >= 0x409u)
        {
            if (exp_as_u64(x) == 0xfff0000000000000ull)
                return 0.0;                       // -inf
            if (abstop >= 0x7ffu)
                return 1.0 + x;                   // +inf, nan
            return (exp_as_u64(x) >> 63) ? 0.0 : HUGE_VAL;
        }
        abstop = 0;                               // 512 <= |x| < 1024
    }
    double kd = fma(inv_ln2_n, x, shift);
    const uint64_t ki = exp_as_u64(kd);
    kd -= shift;
    const double r = fma(kd, neg_ln2_lo_n, fma(kd, neg_ln2_hi_n, x));
    const uint32_t idx = 2u * (uint32_t) (ki % EXP_TAB_N);
    const uint64_t top = ki << (52 - EXP_TAB_BITS);
    const double tail = exp_as_f64(tab[idx]);
    const uint64_t sbits = tab[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, c5, c4), fma(r2, fma(r, c3, c2), tail + r));
    if (abstop == 0)
        return exp_special(tmp, sbits, ki);
    const double scale = exp_as_f64(sbits);
    return fma(scale, tmp, scale);
}

}  // namespace mmca
