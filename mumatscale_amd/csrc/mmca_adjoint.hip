// Frozen-branch replay of diffusion + partitioning and its adjoint, for
// gfx950.
//
// A forward run records the MMCA_BRANCH_* code that fs_change_diffuse
// (growth.c:221-273) took at every cell and step.  With those codes fixed,
// one step is a smooth map of the FP64 state (ce, cl, fs):
//   ce' = ce + sum_n r_n (cl_n - cl),  r_n = (rs f_n + rl (1 - f_n)) (1 - mold_n),
//         f_n = (fs + fs_n) / 2                     (growth.c:85-113, +x,-x,+y,-y,+z,-z)
//   LIQUID, REMELT: cl' = ce', fs' = 0
//   SOLID:          cl' = ce' k_inv, fs' = fs
//   GROW:           cl' = cl + dcl, fs' = fs + dfs
//   FULL:           cl' = cl + dcl, fs' = 1
//   dcl = (c_int - cl) D_xt 2 + dce / A,  dfs = (dcl A - dce) km_inv / cl,
//   A = 1 - km fs,  dce = ce' - ce,  c_int = c0 + m_inv (T - T_liq)
// k_replay evaluates it with the forward kernel's expressions (bit-identical
// with the recorded run at the recorded parameters).  The adjoint of one
// step is two kernels:
//   k_adj_part  per cell: recomputes ce', back-propagates the branch to
//               lam = dL/dce' and to the direct dependence on (ce, cl, fs)
//   k_adj_diff  per cell: gathers lam of the 6 neighbours through the
//               diffusion stencil
// and a one-workgroup sum of the parameter partials (fixed order, so the
// gradients are reproducible run to run).

#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "host_init.h"
#include "mmca_device.h"
#include "mmca_internal.h"

// Workspace of mmca_adjoint_step_noalloc, for callers that cannot allocate:
// lam over the padded cells, then the per-workgroup partial sums.  Each GPU
// holds its own copy of this array until the process ends; every call on
// that GPU uses it, so calls that overlap on one GPU race.
constexpr long long ADJ_SCRATCH_DOUBLES = 1ll << 22;     // 32 MiB
static __device__ __align__(256) double scratch[ADJ_SCRATCH_DOUBLES];

namespace mmca {
namespace {

// Device address of `scratch` on the current device.  The host-side name of
// the array is not a device address: it is looked up once per device with
// that device current, and only the looked-up address is passed to kernels.
constexpr int ADJ_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[ADJ_MAX_DEVICES] = {};

int adj_scratch_ptr(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess)
        return MMCA_ERR_HIP;
    if (dev < 0 || dev >= ADJ_MAX_DEVICES)
        return MMCA_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev])
    {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return MMCA_ERR_HIP;
        g_scratch_dev[dev] = (double *) p;
    }
    *out = g_scratch_dev[dev];
    return MMCA_OK;
}

constexpr int ADJ_BX = 64;
constexpr int ADJ_BY = 4;
constexpr int ADJ_BS = ADJ_BX * ADJ_BY;
constexpr int SUM_BS = 256;

// Derivatives of the folded constants with respect to the parameters:
// rs = d_sol * dtx, rl = d_liq * dtx (growth.c:38-40), D_xt = d_liq / dx / dx * dt
// (growth.c:185-186), km = 1 - k, km_inv = 1 / km, k_inv = 1 / k.
struct Deriv {
    double dtx;       // d rl / d d_liq = d rs / d d_sol
    double dDxt;      // d D_xt / d d_liq
};

long long adj_blocks(const Consts &c)
{
    return (long long) ((c.nx + ADJ_BX - 1) / ADJ_BX) * ((c.ny + ADJ_BY - 1) / ADJ_BY) * c.nz;
}

struct Cell {
    int x, j, k;
    long long idx;
    bool active;
};

__device__ __forceinline__ Cell
cell_of(const Consts &c)
{
    const int gbx = (c.nx + ADJ_BX - 1) / ADJ_BX, gby = (c.ny + ADJ_BY - 1) / ADJ_BY;
    const long long blk = blockIdx.x;
    Cell q;
    q.x = (int) (blk % gbx) * ADJ_BX + threadIdx.x + 1;
    q.j = (int) ((blk / gbx) % gby) * ADJ_BY + threadIdx.y + 1;
    q.k = (int) (blk / ((long long) gbx * gby)) + 1;
    q.active = (q.x <= c.nx) && (q.j <= c.ny);
    q.idx = q.k * c.sxy + (long long) q.j * c.sx + q.x;
    return q;
}

// neighbours in the diffusion order +x,-x,+y,-y,+z,-z and their mold flags
// (only the padded faces of the box can be mold)
__device__ __forceinline__ void
stencil(const Consts &c, const Cell &q, long long nb[6], int mo[6])
{
    nb[0] = q.idx + 1;
    nb[1] = q.idx - 1;
    nb[2] = q.idx + c.sx;
    nb[3] = q.idx - c.sx;
    nb[4] = q.idx + c.sxy;
    nb[5] = q.idx - c.sxy;
    mo[0] = (q.x == c.nx) ? c.pad_hi[0] : 0;
    mo[1] = (q.x == 1) ? c.pad_lo[0] : 0;
    mo[2] = (q.j == c.ny) ? c.pad_hi[1] : 0;
    mo[3] = (q.j == 1) ? c.pad_lo[1] : 0;
    mo[4] = (q.k == c.nz) ? c.pad_hi[2] : 0;
    mo[5] = (q.k == 1) ? c.pad_lo[2] : 0;
}

template <bool AM>
__global__ void __launch_bounds__(ADJ_BS)
k_replay(Consts c, const unsigned char *code, double sim_time, const double *ce, const double *cl,
         const double *fs, double *ce_o, double *cl_o, double *fs_o)
{
    const Cell q = cell_of(c);
    if (!q.active)
        return;
    long long nb[6];
    int mo[6];
    stencil(c, q, nb, mo);
    double cln[6], fsn[6];
#pragma unroll
    for (int n = 0; n < 6; ++n)
    {
        cln[n] = cl[nb[n]];
        fsn[n] = fs[nb[n]];
    }
    const double ce_old = ce[q.idx], conc = cl[q.idx], lspfs = fs[q.idx];
    const double ce_new = diffuse_cell(c, ce_old, conc, lspfs, cln, fsn, mo);
    const int br = rec_code(c, code, q.x, q.j, q.k);
    double clv, fsv;
    if (br == MMCA_BRANCH_SOLID)
    {
        clv = ce_new * c.k_inv;
        fsv = lspfs;
    }
    else if (br == MMCA_BRANCH_GROW || br == MMCA_BRANCH_FULL)
    {
        const double Tcell = temp_model<AM>(c, sim_time, q.x, q.j, q.k);
        const double c_int = c.cinit + c.m_inv * (Tcell - c.Tliq);
        const double dce = ce_new - ce_old;
        const double dcl = (c_int - conc) * c.D_xt * 2 + dce / (1 - c.km * lspfs);
        const double dfs = (dcl * (1. - c.km * lspfs) - dce) * c.km_inv / conc;
        clv = conc + dcl;
        fsv = (br == MMCA_BRANCH_GROW) ? lspfs + dfs : 1.0;
    }
    else
    {
        clv = ce_new;
        fsv = 0.0;
    }
    ce_o[q.idx] = ce_new;
    cl_o[q.idx] = clv;
    fs_o[q.idx] = fsv;
}

// Per-cell adjoint of the branch.  Writes lam = dL/dce' and the direct
// derivatives with respect to the cell's own (ce, cl, fs); partials of
// (d_liq, part_coef) per workgroup into part[2 * blk].
template <bool AM>
__global__ void __launch_bounds__(ADJ_BS)
k_adj_part(Consts c, Deriv dv, const unsigned char *code, double sim_time, const double *ce,
           const double *cl, const double *fs, const double *g_ce, const double *g_cl, const double *g_fs,
           double *lam, double *gce_in, double *gcl_in, double *gfs_in, double *part)
{
    __shared__ double s_red[ADJ_BS / 64];
    const Cell q = cell_of(c);
    double p_dl = 0.0, p_k = 0.0;
    if (q.active)
    {
        long long nb[6];
        int mo[6];
        stencil(c, q, nb, mo);
        double cln[6], fsn[6];
#pragma unroll
        for (int n = 0; n < 6; ++n)
        {
            cln[n] = cl[nb[n]];
            fsn[n] = fs[nb[n]];
        }
        const double ce_old = ce[q.idx], conc = cl[q.idx], lspfs = fs[q.idx];
        const double ce_new = diffuse_cell(c, ce_old, conc, lspfs, cln, fsn, mo);
        const double ga = g_ce[q.idx], gb = g_cl[q.idx], ge = g_fs[q.idx];
        const int br = rec_code(c, code, q.x, q.j, q.k);
        double lamv, gce_d = 0.0, gcl_d = 0.0, gfs_d = 0.0;
        if (br == MMCA_BRANCH_SOLID)
        {
            // cl' = ce' k_inv, fs' = fs
            lamv = ga + gb * c.k_inv;
            gfs_d = ge;
            p_k = -gb * ce_new * c.k_inv * c.k_inv;
        }
        else if (br == MMCA_BRANCH_GROW || br == MMCA_BRANCH_FULL)
        {
            const double Tcell = temp_model<AM>(c, sim_time, q.x, q.j, q.k);
            const double c_int = c.cinit + c.m_inv * (Tcell - c.Tliq);
            const double A = 1 - c.km * lspfs;
            const double dce = ce_new - ce_old;
            const double D2 = c.D_xt * 2;
            const double dcl = (c_int - conc) * D2 + dce / A;
            const double u = dcl * A - dce;
            const double gdfs = (br == MMCA_BRANCH_GROW) ? ge : 0.0;   // FULL: fs' = 1
            // dfs = u km_inv / cl
            const double gu = gdfs * c.km_inv / conc;
            double gq = -gdfs * u * c.km_inv / (conc * conc);
            const double gkminv = gdfs * u / conc;
            // u = dcl A - dce ; cl' = cl + dcl
            const double gdcl = gb + gu * A;
            const double gA = gu * dcl - gdcl * dce / (A * A);
            const double gdce = -gu + gdcl / A;
            // dcl = (c_int - cl) D2 + dce / A
            gq += gb - gdcl * D2;
            const double gD2 = gdcl * (c_int - conc);
            // A = 1 - km fs ; fs' = fs + dfs
            const double gs = gdfs - gA * c.km;
            const double gkm = -gA * lspfs - gkminv * c.km_inv * c.km_inv;
            lamv = ga + gdce;
            gce_d = -gdce;          // dce = ce' - ce
            gcl_d = gq;
            gfs_d = gs;
            p_k = -gkm;             // km = 1 - k
            p_dl = gD2 * 2 * dv.dDxt;
        }
        else
        {
            // LIQUID, REMELT: cl' = ce', fs' = 0
            lamv = ga + gb;
        }
        lam[q.idx] = lamv;
        gce_in[q.idx] = gce_d;
        gcl_in[q.idx] = gcl_d;
        gfs_in[q.idx] = gfs_d;
    }
    const double s_dl = block_sum<ADJ_BS>(p_dl, s_red);
    __syncthreads();
    const double s_k = block_sum<ADJ_BS>(p_k, s_red);
    if (threadIdx.x == 0 && threadIdx.y == 0)
    {
        part[2 * (long long) blockIdx.x] = s_dl;
        part[2 * (long long) blockIdx.x + 1] = s_k;
    }
}

// Adjoint of the diffusion stencil: adds lam's pull-back to (ce, cl, fs) and
// the (d_liq, d_sol) partials per workgroup into part[2 * blk].  lam is zero
// on the halo (halo cells are never updated).
__global__ void __launch_bounds__(ADJ_BS)
k_adj_diff(Consts c, Deriv dv, const double *cl, const double *fs, const double *lam, double *gce_in,
           double *gcl_in, double *gfs_in, double *part)
{
    __shared__ double s_red[ADJ_BS / 64];
    const Cell q = cell_of(c);
    double p_dl = 0.0, p_ds = 0.0;
    if (q.active)
    {
        long long nb[6];
        int mo[6];
        stencil(c, q, nb, mo);
        const double lc = lam[q.idx], clc = cl[q.idx], fsc = fs[q.idx];
        const double hd = 0.5 * (c.rs - c.rl);
        double gcl = 0.0, gfs = 0.0, grs = 0.0, grl = 0.0;
#pragma unroll
        for (int n = 0; n < 6; ++n)
        {
            const double ln = lam[nb[n]];
            const double f = 0.5 * (fsc + fs[nb[n]]);
            const double w = 1 - mo[n];
            const double r = (c.rs * f + c.rl * (1 - f)) * w;
            const double d = cl[nb[n]] - clc;
            gcl += r * (ln - lc);
            gfs += hd * w * d * (lc - ln);
            grs += lc * f * w * d;
            grl += lc * (1 - f) * w * d;
        }
        gce_in[q.idx] += lc;
        gcl_in[q.idx] += gcl;
        gfs_in[q.idx] += gfs;
        p_dl = grl * dv.dtx;
        p_ds = grs * dv.dtx;
    }
    const double s_dl = block_sum<ADJ_BS>(p_dl, s_red);
    __syncthreads();
    const double s_ds = block_sum<ADJ_BS>(p_ds, s_red);
    if (threadIdx.x == 0 && threadIdx.y == 0)
    {
        part[2 * (long long) blockIdx.x] = s_dl;
        part[2 * (long long) blockIdx.x + 1] = s_ds;
    }
}

// g_param[0..2] += sums of the partials: (d_liq, part_coef) from pa,
// (d_liq, d_sol) from pb.  One workgroup; fixed summation order.
__global__ void __launch_bounds__(SUM_BS)
k_adj_sum(const double *pa, const double *pb, long long nblk, double *g_param)
{
    __shared__ double s_red[SUM_BS / 64];
    double v[4] = { 0.0, 0.0, 0.0, 0.0 };
    for (long long i = threadIdx.x; i < nblk; i += SUM_BS)
    {
        v[0] += pa[2 * i];
        v[1] += pa[2 * i + 1];
        v[2] += pb[2 * i];
        v[3] += pb[2 * i + 1];
    }
    double t[4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
    {
        t[m] = block_sum<SUM_BS>(v[m], s_red);
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        g_param[0] += t[0] + t[2];
        g_param[1] += t[3];
        g_param[2] += t[1];
    }
}

// lam = 0 on the halo cells of the padded box (the stencil reads them; they
// are never written by k_adj_part)
__global__ void __launch_bounds__(256)
k_zero_halo(Consts c, double *lam)
{
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.P)
        return;
    const int x = (int) (i % c.sx);
    const long long r = i / c.sx;
    const int j = (int) (r % (c.ny + 2));
    const int k = (int) (r / (c.ny + 2));
    if (x == 0 || x > c.nx || j == 0 || j > c.ny || k == 0 || k > c.nz)
        lam[i] = 0.0;
}

// The launches of one adjoint step; lam holds P doubles, pa and pb 2 * nblk each.
hipError_t
adjoint_launch(const Consts &c, const Deriv &dv, uint64_t timestep, const uint8_t *codes, const double *ce,
               const double *cl, const double *fs, const double *g_ce, const double *g_cl, const double *g_fs,
               double *g_ce_in, double *g_cl_in, double *g_fs_in, double *g_param, double *lam, double *pa,
               double *pb, hipStream_t s)
{
    const long long nblk = adj_blocks(c);
    const double sim_time = timestep * c.dt;
    const dim3 grid((unsigned) nblk), blk(ADJ_BX, ADJ_BY);
    if (c.am)
        hipLaunchKernelGGL(k_adj_part<true>, grid, blk, 0, s, c, dv, codes, sim_time, ce, cl, fs, g_ce, g_cl, g_fs,
                           lam, g_ce_in, g_cl_in, g_fs_in, pa);
    else
        hipLaunchKernelGGL(k_adj_part<false>, grid, blk, 0, s, c, dv, codes, sim_time, ce, cl, fs, g_ce, g_cl,
                           g_fs, lam, g_ce_in, g_cl_in, g_fs_in, pa);
    hipLaunchKernelGGL(k_adj_diff, grid, blk, 0, s, c, dv, cl, fs, lam, g_ce_in, g_cl_in, g_fs_in, pb);
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(SUM_BS), 0, s, pa, pb, nblk, g_param);
    return hipGetLastError();
}

int adj_consts(const mmca_params *p, Consts &c, Deriv &dv)
{
    if (!p)
        return MMCA_ERR_ARG;
    mmca_params q = *p;
    q.precision = MMCA_PREC_FP64;
    const int st = build_consts(q, c);
    if (st != MMCA_OK)
        return st;
    dv.dtx = p->dt / (p->cell_size * p->cell_size);
    dv.dDxt = 1.0 / p->cell_size / p->cell_size * p->dt;
    return MMCA_OK;
}

}  // namespace
}  // namespace mmca

extern "C" {

int mmca_replay_step(const mmca_params *params, uint64_t timestep, const uint8_t *codes, const double *ce,
                     const double *cl, const double *fs, double *ce_out, double *cl_out, double *fs_out,
                     void *stream)
{
    using namespace mmca;
    Consts c;
    Deriv dv;
    int st = adj_consts(params, c, dv);
    if (st != MMCA_OK)
        return st;
    if (!codes || !ce || !cl || !fs || !ce_out || !cl_out || !fs_out)
        return MMCA_ERR_ARG;
    if (ce_out == ce || cl_out == cl || fs_out == fs || cl_out == fs || fs_out == cl)
        return MMCA_ERR_ARG;     // the stencil reads neighbours of the inputs
    hipStream_t s = (hipStream_t) stream;
    const double sim_time = timestep * c.dt;
    const dim3 grid((unsigned) adj_blocks(c)), blk(ADJ_BX, ADJ_BY);
    if (c.am)
        hipLaunchKernelGGL(k_replay<true>, grid, blk, 0, s, c, codes, sim_time, ce, cl, fs, ce_out, cl_out, fs_out);
    else
        hipLaunchKernelGGL(k_replay<false>, grid, blk, 0, s, c, codes, sim_time, ce, cl, fs, ce_out, cl_out,
                           fs_out);
    return hipGetLastError() == hipSuccess ? MMCA_OK : MMCA_ERR_HIP;
}

int mmca_record_bytes(const mmca_params *params, int64_t *bytes_per_step)
{
    using namespace mmca;
    Consts c;
    Deriv dv;
    int st = adj_consts(params, c, dv);
    if (st != MMCA_OK)
        return st;
    if (!bytes_per_step)
        return MMCA_ERR_ARG;
    *bytes_per_step = (int64_t) sizeof(unsigned long long) * rec_words(c);
    return MMCA_OK;
}

int mmca_adjoint_workspace_bytes(const mmca_params *params, int64_t *bytes)
{
    using namespace mmca;
    Consts c;
    Deriv dv;
    int st = adj_consts(params, c, dv);
    if (st != MMCA_OK)
        return st;
    if (!bytes)
        return MMCA_ERR_ARG;
    *bytes = (int64_t) sizeof(double) * (c.P + 4 * adj_blocks(c));
    return MMCA_OK;
}

int mmca_adjoint_step(const mmca_params *params, uint64_t timestep, const uint8_t *codes, const double *ce,
                      const double *cl, const double *fs, const double *g_ce, const double *g_cl,
                      const double *g_fs, double *g_ce_in, double *g_cl_in, double *g_fs_in, double *g_param,
                      void *workspace, void *stream)
{
    using namespace mmca;
    Consts c;
    Deriv dv;
    int st = adj_consts(params, c, dv);
    if (st != MMCA_OK)
        return st;
    if (!codes || !ce || !cl || !fs || !g_ce || !g_cl || !g_fs || !g_ce_in || !g_cl_in || !g_fs_in ||
        !g_param || !workspace)
        return MMCA_ERR_ARG;
    const long long nblk = adj_blocks(c);
    double *lam = (double *) workspace;
    double *pa = lam + c.P;
    double *pb = pa + 2 * nblk;
    return adjoint_launch(c, dv, timestep, codes, ce, cl, fs, g_ce, g_cl, g_fs, g_ce_in, g_cl_in, g_fs_in, g_param,
                          lam, pa, pb, (hipStream_t) stream) == hipSuccess
               ? MMCA_OK
               : MMCA_ERR_HIP;
}

int mmca_adjoint_step_noalloc(const mmca_params *params, uint64_t timestep, const uint8_t *codes,
                              const double *ce, const double *cl, const double *fs, const double *g_ce,
                              const double *g_cl, const double *g_fs, double *g_ce_in, double *g_cl_in,
                              double *g_fs_in, double *g_param, void *stream)
{
    using namespace mmca;
    Consts c;
    Deriv dv;
    int st = adj_consts(params, c, dv);
    if (st != MMCA_OK)
        return st;
    if (!codes || !ce || !cl || !fs || !g_ce || !g_cl || !g_fs || !g_ce_in || !g_cl_in || !g_fs_in || !g_param)
        return MMCA_ERR_ARG;
    const long long nblk = adj_blocks(c);
    if (c.P + 4 * nblk > ADJ_SCRATCH_DOUBLES)
    {
        fprintf(stderr, "mmca_adjoint_step_noalloc: %lld padded cells need %lld bytes of workspace, above the "
                        "%lld of its fixed device array\n",
                (long long) c.P, (long long) sizeof(double) * (c.P + 4 * nblk),
                (long long) sizeof(double) * ADJ_SCRATCH_DOUBLES);
        return MMCA_ERR_ARG;
    }
    double *lam = nullptr;
    st = adj_scratch_ptr(&lam);
    if (st != MMCA_OK)
        return st;
    double *pa = lam + c.P;
    double *pb = pa + 2 * nblk;
    hipStream_t s = (hipStream_t) stream;
    // a previous call (another shape) may have left values where this shape's halo lies
    hipLaunchKernelGGL(k_zero_halo, dim3((unsigned) ((c.P + 255) / 256)), dim3(256), 0, s, c, lam);
    return adjoint_launch(c, dv, timestep, codes, ce, cl, fs, g_ce, g_cl, g_fs, g_ce_in, g_cl_in, g_fs_in, g_param,
                          lam, pa, pb, s) == hipSuccess
               ? MMCA_OK
               : MMCA_ERR_HIP;
}

}  // extern "C"
