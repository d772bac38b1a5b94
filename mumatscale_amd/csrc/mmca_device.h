// Device functions shared by the CA step kernels and the adjoint kernels.
// Arithmetic follows the reference's operation order; everything that
// includes this file is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include "mmca_exp.h"
#include "mmca_internal.h"

namespace mmca {

// Device buffers reached through the domain table are generic pointers to
// the compiler, whose loads and stores are then FLAT instructions: 64-bit
// per-lane addresses, and a wait for LDS results also waits for them.  G(p)
// states that p points to global memory, so that global (and, for uniform
// addresses, scalar) loads and stores are used.
#define MMCA_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ MMCA_GLOBAL T *G(T *p)
{
    return (MMCA_GLOBAL T *) p;
}

// MIN(x,y) of globals.h:19
__device__ __forceinline__ double ref_min(double x, double y) { return (x < y) ? x : y; }

// Sum over a workgroup of NT threads (result valid in thread 0).
template <int NT>
__device__ __forceinline__ double
block_sum(double v, double *s_red)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1)
        v += __shfl_xor(v, d, 64);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0)
        s_red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int i = 0; i < NT / 64; ++i)
            s += s_red[i];
    return s;
}

// direct_temp_calc (temperature.c:24-126); x,y,z are padded indices.  AM
// selects the Gaussian-beam model at compile time, so that kernels of the
// other models do not carry its registers.
template <bool AM>
__device__ __forceinline__ double
temp_model(const Consts &c, double sim_time, int x, int y, int z)
{
    if (!AM && c.gradient == 0)
        return c.t_init - sim_time * c.cool_rate;
    double rx = c.origin[0] + ((c.off[0] + x) + 0.5) * c.cell_size;
    double ry = c.origin[1] + ((c.off[1] + y) + 0.5) * c.cell_size;
    double rz = c.origin[2] + ((c.off[2] + z) + 0.5) * c.cell_size;
    if (!AM)
    {
        double axis = ((c.gdimz == 1) ? ry : rz) - c.cell_size * 0.5;
        double tgrad = c.gradient + c.grad_coef * sim_time;
        double v = c.velocity + c.velo_coef * sim_time;
        double vt = 0.5 * (v + c.velocity) * sim_time;
        return c.t_init + tgrad * (axis - vt);
    }
    double lv = c.lv;
    double invlv = 1. / lv;
    double ls = c.lsd / (2. * 1.414);
    double xlen = c.gnsbx * c.nx * c.cell_size;
    xlen = xlen + c.lsd;
    double tline = xlen * invlv;
    double tpause = 1.5e-4 * invlv;
    tline = tline + tpause;
    double trackno = 0;
    if (c.am_tracks > 1)
        trackno = floor(sim_time / tline);
    double xnow = c.x0 + (sim_time - trackno * tline) * lv;
    double ynow = c.y0 + trackno * c.lhs;
    double znow = c.z0;
    double lpow = c.lpow;
    double lx = rx - xnow;
    double ly = ry - ynow;
    double lz = rz - znow;
    if (!(rz <= znow + c.cell_size))
        return 0.0;
    const int nt = 400;
    double Tinteg = 0.0;
    double pi15 = 5.568328;
    double lalpha = c.lthcon / (c.rho * c.lcp);
    double tsteady = 20 * lalpha * invlv * invlv;
    double ldt = tsteady / nt;
    if (sim_time < 1e-4)
        lpow = 0;
    for (int i = 0; i < nt; i++)
    {
        double lt = (0.5 + i) * ldt;
        double Tlocal = sqrt(lalpha * lt) * (ls * ls + 4. * lalpha * lt);
        Tinteg += ldt *
            exp_glibc(-((lx + lv * lt) * (lx + lv * lt) + ly * ly) / (ls * ls + 4. * lalpha * lt) -
                          (lz * lz / (4. * lalpha * lt)),
                      exp_tab) /
            Tlocal;
    }
    double Tfront = c.lab * lpow * lalpha / (pi15 * c.lthcon);
    return c.liniT + Tfront * Tinteg;
}

// The Gaussian-beam quadrature of temp_model<true> with the factors of its
// 400 terms that do not depend on the cell tabulated (same operations and
// order as temp_model<true>, so the same values):
//   vlt[i] = lv * lt,  A[i] = ls*ls + 4.*lalpha*lt,
//   Tl[i] = sqrt(lalpha*lt) * (ls*ls + 4.*lalpha*lt),
//   Z[i] = lz*lz / (4.*lalpha*lt) for one plane (lz depends on z only),
// and a copy of the exp table, so that its lookups are LDS reads.
struct BeamTab {
    double vlt[BEAM_NT], A[BEAM_NT], Tl[BEAM_NT], Z[BEAM_NT];
    uint64_t exp_tab[EXP_TAB_WORDS];
};

__device__ __forceinline__ double beam_lalpha(const Consts &c) { return c.lthcon / (c.rho * c.lcp); }

__device__ __forceinline__ double
beam_ldt(const Consts &c)
{
    const double invlv = 1. / c.lv;
    const double tsteady = 20 * beam_lalpha(c) * invlv * invlv;
    return tsteady / BEAM_NT;
}

// fills vlt, A, Tl; threads tid = 0..nthreads-1 of the workgroup share the work
__device__ __forceinline__ void
beam_tab_terms(const Consts &c, BeamTab &t, int tid, int nthreads)
{
    const double ls = c.lsd / (2. * 1.414);
    const double lalpha = beam_lalpha(c);
    const double ldt = beam_ldt(c);
    for (int i = tid; i < BEAM_NT; i += nthreads)
    {
        const double lt = (0.5 + i) * ldt;
        t.vlt[i] = c.lv * lt;
        t.A[i] = ls * ls + 4. * lalpha * lt;
        t.Tl[i] = sqrt(lalpha * lt) * (ls * ls + 4. * lalpha * lt);
    }
    for (int i = tid; i < EXP_TAB_WORDS; i += nthreads)
        t.exp_tab[i] = exp_tab[i];
}

// fills Z (BEAM_NT values) for padded plane z
__device__ __forceinline__ void
beam_tab_plane(const Consts &c, double *Z, int z, int tid, int nthreads)
{
    const double lalpha = beam_lalpha(c);
    const double ldt = beam_ldt(c);
    const double rz = c.origin[2] + ((c.off[2] + z) + 0.5) * c.cell_size;
    const double lz = rz - c.z0;
    for (int i = tid; i < BEAM_NT; i += nthreads)
    {
        const double lt = (0.5 + i) * ldt;
        Z[i] = lz * lz / (4. * lalpha * lt);
    }
}

__device__ __forceinline__ void
beam_tab_plane(const Consts &c, BeamTab &t, int z, int tid, int nthreads)
{
    beam_tab_plane(c, t.Z, z, tid, nthreads);
}

// temp_model<true> on the tables; Z holds the plane factors of plane z
__device__ __forceinline__ double
beam_temp_tab(const Consts &c, const BeamTab &t, const double *Z, double sim_time, int x, int y, int z)
{
    double rx = c.origin[0] + ((c.off[0] + x) + 0.5) * c.cell_size;
    double ry = c.origin[1] + ((c.off[1] + y) + 0.5) * c.cell_size;
    double rz = c.origin[2] + ((c.off[2] + z) + 0.5) * c.cell_size;
    double lv = c.lv;
    double invlv = 1. / lv;
    double xlen = c.gnsbx * c.nx * c.cell_size;
    xlen = xlen + c.lsd;
    double tline = xlen * invlv;
    double tpause = 1.5e-4 * invlv;
    tline = tline + tpause;
    double trackno = 0;
    if (c.am_tracks > 1)
        trackno = floor(sim_time / tline);
    double xnow = c.x0 + (sim_time - trackno * tline) * lv;
    double ynow = c.y0 + trackno * c.lhs;
    double znow = c.z0;
    double lpow = c.lpow;
    double lx = rx - xnow;
    double ly = ry - ynow;
    if (!(rz <= znow + c.cell_size))
        return 0.0;
    double pi15 = 5.568328;
    double lalpha = beam_lalpha(c);
    double ldt = beam_ldt(c);
    if (sim_time < 1e-4)
        lpow = 0;
    double Tfront = c.lab * lpow * lalpha / (pi15 * c.lthcon);
    if (Tfront == 0.0)
        return c.liniT + Tfront * 0.0;     // every term is finite, so Tfront * Tinteg is this
    double Tinteg = 0.0;
    for (int i = 0; i < BEAM_NT; i++)
    {
        const double u = lx + t.vlt[i];
        Tinteg += ldt * exp_glibc(-(u * u + ly * ly) / t.A[i] - Z[i], t.exp_tab) / t.Tl[i];
    }
    return c.liniT + Tfront * Tinteg;
}

// Octahedron half-diagonal of a solid cell (grow_octahedra, growth.c:392-403)
// from its centre offset, fs and grain g.
__device__ __forceinline__ double
octa_d(const Consts &c, const DevBufs &b, long long i, int g, double fs)
{
    const MMCA_GLOBAL double *R = G(b.g_rot) + 9 * (long long) g;
    double rx0 = 0.0 - G(b.dc)[3 * i];
    double ry0 = 0.0 - G(b.dc)[3 * i + 1];
    double rz0 = 0.0 - G(b.dc)[3 * i + 2];
    double gx0 = R[0] * rx0 + R[3] * ry0 + R[6] * rz0;
    double gy0 = R[1] * rx0 + R[4] * ry0 + R[7] * rz0;
    double gz0 = R[2] * rx0 + R[5] * ry0 + R[8] * rz0;
    double ds1 = fabs(gx0) + fabs(gy0) + fabs(gz0);
    ds1 = ref_min(ds1, 1.5);
    return ds1 + fs * c.rev_fsgrow;
}

// With track_d: a cell that stops being solid keeps, in the reference's d
// array, the half-diagonal of its last step as a solid cell (grow_octahedra
// writes solid cells only).  That value follows from the state the
// partitioning step starts from: the cell's dc and fs are the ones of its
// last growth step (capture writes only liquid cells).
__device__ __forceinline__ void
keep_stale_d(const Consts &c, const DevBufs &b, long long i, int g_old, int g_new, double fs_old)
{
    if (c.track_d && g_old > 0 && g_new <= 0 && g_old < c.max_total_grains)
        G(b.d)[i] = octa_d(c, b, i, g_old, fs_old);
}

// 64-bit words of one step's branch record
__host__ __device__ inline long long
rec_words(const Consts &c)
{
    return 3ll * c.w64 * (c.ny + 2) * (c.nz + 2);
}

// Branch record of one step: per padded row (j, k) and 64-cell segment of x,
// three 64-bit planes; bit (x - 1) % 64 of plane p is bit p of the
// MMCA_BRANCH_* code of cell x (3 bits per interior cell).
__device__ __forceinline__ long long
rec_word(const Consts &c, int seg, int j, int k)
{
    return (((long long) k * (c.ny + 2) + j) * c.w64 + seg) * 3;
}

__device__ __forceinline__ int
rec_code(const Consts &c, const unsigned char *rec, int x, int j, int k)
{
    const unsigned long long *w = reinterpret_cast<const unsigned long long *>(rec) + rec_word(c, (x - 1) >> 6, j, k);
    const int bit = (x - 1) & 63;
    return (int) ((w[0] >> bit) & 1ull) | (int) (((w[1] >> bit) & 1ull) << 1) | (int) (((w[2] >> bit) & 1ull) << 2);
}

__device__ inline double
temp_at(const Consts &c, double sim_time, int x, int y, int z)
{
    return c.am ? temp_model<true>(c, sim_time, x, y, z) : temp_model<false>(c, sim_time, x, y, z);
}

// ---------------------------------------------------------------------------
// Diffusion (growth.c:85-113, order +x,-x,+y,-y,+z,-z) of one cell, from the
// neighbours' cl and fs and their mold flags.
__device__ __forceinline__ double
diffuse_cell(const Consts &c, double ce_old, double conc, double lspfs, const double *cln,
             const double *fsn, const int *mo)
{
    double nbsum = 0;
#pragma unroll
    for (int n = 0; n < 6; ++n)
    {
        double fs_av = 0.5 * (lspfs + fsn[n]);
        double r = (c.rs * fs_av + c.rl * (1 - fs_av)) * (1 - mo[n]);
        nbsum += r * (cln[n] - conc);
    }
    return ce_old + nbsum;
}

}  // namespace mmca
