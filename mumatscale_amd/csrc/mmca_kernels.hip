// HIP kernels of the CA step for gfx950 (CDNA4, wave64).
//
// One step (doiteration, calculate.c:193-352) on one subblock is seven
// launches on the context's stream:
//   k_nuc_detect    cell_nucleation (grain.c:1222-1318) over the sorted site list
//   k_nuc_activate  activateNewGrains (grain.c:841-916): ids in site order, one workgroup
//   k_solute        diffusion + partitioning (growth.c:30-170, 177-276),
//                   fused, and a 1-bit "solid" mask of gr > 0
//   k_list_count    interface-list membership counts from the mask
//                   (grow_cell_reduction, growth.c:461-538)
//   k_list_scan     exclusive scan of the counts -> raster-order slots
//   k_list_write    the list itself (diff_id) in raster order
//   k_capture       capture_octahedra_diffuse (growth.c:574-807) over the
//                   list slots, dealt evenly over the whole grid, with the
//                   octahedron half-diagonal of each candidate neighbour
//                   computed on the fly (grow_octahedra, growth.c:339-407)
//
// Arithmetic follows the reference's operation order and is compiled with
// -ffp-contract=off: FP64 results are bit-identical with the reference CPU
// build.  The Gaussian-beam temperature's exp is a restatement of the host C
// library's (mmca_exp.h).

#include <hip/hip_runtime.h>

#include <algorithm>

#include "mmca_device.h"
#include "mmca_internal.h"

namespace mmca {

// Ping-pong buffer `buf` (0/1) of a field, chosen with a select rather than an
// indexed load, so that a private copy of DevBufs stays in registers.
__device__ __forceinline__ MMCA_GLOBAL double *fs_buf(const DevBufs &b, int buf) { return G(buf ? b.fs[1] : b.fs[0]); }
__device__ __forceinline__ MMCA_GLOBAL double *cl_buf(const DevBufs &b, int buf) { return G(buf ? b.cl[1] : b.cl[0]); }
__device__ __forceinline__ MMCA_GLOBAL float *cl32_buf(const DevBufs &b, int buf) { return G(buf ? b.cl32[1] : b.cl32[0]); }

// Solute storage: FP64 values, or FP32 deviations from c0 (F32).  Arithmetic
// is FP64 in both cases.
template <bool F32>
__device__ __forceinline__ double
ld_ce(const DevBufs &b, const Consts &c, long long i)
{
    return F32 ? c.cinit + (double) G(b.ce32)[i] : G(b.ce)[i];
}

template <bool F32>
__device__ __forceinline__ double
ld_cl(const DevBufs &b, const Consts &c, int buf, long long i)
{
    return F32 ? c.cinit + (double) cl32_buf(b, buf)[i] : cl_buf(b, buf)[i];
}

// The solute kernel writes each cell of ce, cl and fs once per step and the
// later kernels of the step read few of them back (capture), so these stores
// are non-temporal: streamed past the caches, coherent at the kernel's end.
template <typename T>
__device__ __forceinline__ void
st_stream(MMCA_GLOBAL T *p, long long i, T v)
{
    __builtin_nontemporal_store(v, &p[i]);
}

// store ce; returns the FP64 solute lost to the FP32 rounding (0 for FP64)
template <bool F32>
__device__ __forceinline__ double
st_ce(const DevBufs &b, const Consts &c, long long i, double v)
{
    if (F32)
    {
        const double dev = v - c.cinit;
        const float f = (float) dev;
        st_stream(G(b.ce32), i, f);
        return dev - (double) f;
    }
    st_stream(G(b.ce), i, v);
    return 0.0;
}

template <bool F32>
__device__ __forceinline__ void
st_cl(const DevBufs &b, const Consts &c, int buf, long long i, double v)
{
    if (F32)
        st_stream(cl32_buf(b, buf), i, (float) (v - c.cinit));
    else
        st_stream(cl_buf(b, buf), i, v);
}

// Compensated (Kahan) addition of a workgroup's residual into its ledger slot.
__device__ __forceinline__ void
ledger_add(const DevBufs &b, long long slot, double v)
{
    double s = G(b.ledger)[2 * slot], comp = G(b.ledger)[2 * slot + 1];
    double y = v - comp;
    double t = s + y;
    comp = (t - s) - y;
    G(b.ledger)[2 * slot] = t;
    G(b.ledger)[2 * slot + 1] = comp;
}

// ---------------------------------------------------------------------------
// Block-wide exclusive scan of 0/1 flags with wave ballots.  NW = waves/block.
template <int NW>
__device__ __forceinline__ int
block_exscan_flag(int flag, int *total, int *s_w)
{
    unsigned long long bal = __ballot(flag);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int r = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0)
        s_w[w] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i)
    {
        int v = s_w[i];
        off += (i < w) ? v : 0;
        tot += v;
    }
    __syncthreads();
    *total = tot;
    return off + r;
}

// Block-wide exclusive scan of small non-negative ints.
template <int NW>
__device__ __forceinline__ int
block_exscan_int(int v, int *total, int *s_w)
{
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        int t = __shfl_up(incl, d, 64);
        if (lane >= d)
            incl += t;
    }
    if (lane == 63)
        s_w[w] = incl;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i)
    {
        int x = s_w[i];
        off += (i < w) ? x : 0;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// Domain of workgroup w: the last d with off[d] <= w (domains without work
// share their offset with the next domain).
__device__ __forceinline__ int
find_dom(const long long *off, int ndom, long long w)
{
    int lo = 0, hi = ndom - 1;
    while (lo < hi)
    {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// cell_nucleation (grain.c:1222-1318): one thread per site, sites sorted by
// padded index; each workgroup compacts its events in site order.
constexpr int NUC_BS = 256;

__global__ void __launch_bounds__(NUC_BS)
k_nuc_detect(View V, int advance, int cur)
{
    __shared__ int s_w[NUC_BS / 64];
    const int d = find_dom(V.off_nuc, V.ndom, blockIdx.x);
    const long long blk = blockIdx.x - V.off_nuc[d];
    const DevBufs b = V.dom[d].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[d].c;
    const MMCA_GLOBAL Counters *cnt = G(b.cnt);
    if (cnt->status != MMCA_OK)
        return;
    const unsigned long long ts = cnt->timestep + (advance ? 1ull : 0ull);
    const double sim_time = ts * c.dt;
    const int s = (int) blk * NUC_BS + threadIdx.x;
    int flag = 0;
    if (s < b.nsites)
    {
        int idx = G(b.site_idx)[s];
        int z = (int) (idx / c.sxy);
        int rem = (int) (idx - z * c.sxy);
        int y = rem / c.sx;
        int x = rem - y * c.sx;
        double T = c.beam_sep ? G(b.temp)[idx] : temp_at(c, sim_time, x, y, z);
        int layerno = 1, initnuc = 1;
        if (c.am)
        {
            // tempUpdateSB_int sets lsindex before nucleation (temperature.c:256-259)
            if (T > c.Tliq)
                G(b.site_ls)[s] = 1;
            if (c.am_layers == 1)
                layerno = 1;
            else if (c.am_layers > 1 && G(b.site_ls)[s] == 1)
                layerno = 1;
            else
                layerno = 0;
            if (sim_time <= 1e-4)
                initnuc = 1;
            else if (T >= c.Tsol)
                initnuc = 1;
            else
                initnuc = 0;
        }
        if (G(b.gr)[idx] <= 0 && layerno && initnuc)
        {
            double Tunder = c.Tliq - T;
            double m_solute0 = c.ma * T * T + c.mb * T + c.mc;
            Tunder += m_solute0 * ((c.f32 ? ld_cl<true>(b, c, cur, idx) : ld_cl<false>(b, c, cur, idx)) - c.cinit);
            if (Tunder > (double) G(b.site_thr)[s])
                flag = 1;
        }
    }
    int tot;
    int r = block_exscan_flag<NUC_BS / 64>(flag, &tot, s_w);
    if (flag)
        G(b.ev_blk)[blk * NUC_BS + r] = s;
    if (threadIdx.x == 0)
        G(b.ev_cnt)[blk] = tot;
}

// Event ordering and activateNewGrains/createGrains on one rank
// (grain.c:841-916, 348-425), one workgroup.
// mode: 1 = order the detected events, 2 = activate, 4 = advance the timestep.
__global__ void __launch_bounds__(NUC_BS)
k_nuc_activate(View V, int mode)
{
    __shared__ int s_w[NUC_BS / 64];
    __shared__ int s_nev;
    const DevBufs b = V.dom[blockIdx.x].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[blockIdx.x].c;
    const int nblk = (b.nsites + NUC_BS - 1) / NUC_BS;
    Counters *cnt = b.cnt;
    if (cnt->status != MMCA_OK)
        return;
    const int tid = threadIdx.x;
    const unsigned long long ts = cnt->timestep + ((mode & 4) ? 1ull : 0ull);
    if (mode & 1)
    {
        // exclusive scan of the per-workgroup counts, then concatenate
        int carry = 0;
        for (int base = 0; base < nblk; base += NUC_BS)
        {
            int blk = base + tid;
            int v = (blk < nblk) ? G(b.ev_cnt)[blk] : 0;
            int tot;
            int off = carry + block_exscan_int<NUC_BS / 64>(v, &tot, s_w);
            for (int i = 0; i < v; ++i)
                G(b.events)[off + i] = G(b.ev_blk)[blk * NUC_BS + i];
            carry += tot;
        }
        if (tid == 0)
        {
            s_nev = carry;
            cnt->n_events = carry;
        }
    }
    else if (tid == 0)
        s_nev = cnt->n_events;
    __syncthreads();
    const int nev = s_nev;
    if ((mode & 2) && nev > 0)
    {
        const int base_id = cnt->num_grains;
        const int rnd0 = cnt->random_used;
        const long long maxg = c.max_total_grains;
        int rused = 0;
        for (int base = 0; base < nev; base += NUC_BS)
        {
            int e = base + tid;
            int isev = e < nev;
            int s = isev ? G(b.events)[e] : 0;
            int fixed = isev ? G(b.site_fixed)[s] : -1;
            int israndom = isev && fixed < 0;
            int rtot;
            int rrank = block_exscan_flag<NUC_BS / 64>(israndom, &rtot, s_w);
            if (isev)
            {
                long long id = (long long) base_id + e;
                if (id <= maxg)
                {
                    int idx = G(b.site_idx)[s];
                    G(b.gr)[idx] = (int) id;
                    const double *rot;
                    int src;
                    if (fixed >= 0)
                    {
                        rot = b.fixed_rot + 9 * (long long) fixed;
                        src = fixed;
                    }
                    else
                    {
                        int k = rnd0 + rused + rrank;
                        rot = b.rand_rot + 9 * (long long) k;
                        src = -(k + 1);
                    }
                    for (int q = 0; q < 9; ++q)
                        b.g_rot[9 * id + q] = rot[q];
                    G(b.g_cell)[id] = idx;
                    G(b.g_step)[id] = ts;
                    G(b.g_src)[id] = src;
                }
            }
            rused += rtot;
        }
        __syncthreads();
        if (tid == 0)
        {
            long long total = (long long) base_id + nev;
            if (total > maxg + 1)
            {
                // grain.c:868-888: with every slot taken nothing is activated;
                // otherwise createGrains runs past MaxTotGrains and aborts
                // (grain.c:361-366) after creating the grains that fit.
                if ((long long) base_id != maxg + 1)
                {
                    cnt->status = MMCA_ERR_GRAIN_CAPACITY;
                    cnt->failed_step = ts;
                }
            }
            else
            {
                cnt->num_grains = (int) total;
                cnt->random_used = rnd0 + rused;
                cnt->new_grains += nev;
            }
        }
    }
    if ((mode & 4) && tid == 0)
        cnt->timestep = ts;
}

// activateNewGrains over the ranks of a coupled batch (grain.c:841-916), one
// workgroup: each domain's ordered events (k_nuc_activate mode 1) get ids
// in rank order after the grains existing so far (the displacements of
// grain.c:893-898), orientations from the domain's own stream (its fixed
// entries or its next random draws), in the one grain table the domains
// share.  Every domain advances its timestep; running past MaxTotGrains
// stops all of them, as createGrains' error() stops every rank.
__global__ void __launch_bounds__(NUC_BS)
k_nuc_activate_coupled(View V)
{
    __shared__ int s_w[NUC_BS / 64];
    const int tid = threadIdx.x;
    for (int d = 0; d < V.ndom; ++d)
        if (V.dom[d].b.cnt->status != MMCA_OK)
            return;                               // uniform: all domains stop together
    const Counters *c0 = V.dom[0].b.cnt;
    const unsigned long long ts = c0->timestep + 1ull;
    const int base_id = c0->num_grains;
    const long long maxg = V.dom[0].c.max_total_grains;
    long long total = base_id;
    for (int d = 0; d < V.ndom; ++d)
        total += V.dom[d].b.cnt->n_events;
    const bool over = total > maxg + 1;
    int next = base_id;
    for (int d = 0; d < V.ndom && !over; ++d)
    {
        const DevBufs b = V.dom[d].b;    // private copy: stores cannot alias it
        const int nev = b.cnt->n_events;
        const int rnd0 = b.cnt->random_used;
        int rused = 0;
        for (int base = 0; base < nev; base += NUC_BS)
        {
            int e = base + tid;
            int isev = e < nev;
            int s = isev ? G(b.events)[e] : 0;
            int fixed = isev ? G(b.site_fixed)[s] : -1;
            int israndom = isev && fixed < 0;
            int rtot;
            int rrank = block_exscan_flag<NUC_BS / 64>(israndom, &rtot, s_w);
            if (isev)
            {
                const long long id = (long long) next + e;
                int idx = G(b.site_idx)[s];
                G(b.gr)[idx] = (int) id;
                const double *rot;
                int src;
                if (fixed >= 0)
                {
                    rot = b.fixed_rot + 9 * (long long) fixed;
                    src = fixed;
                }
                else
                {
                    int k = rnd0 + rused + rrank;
                    rot = b.rand_rot + 9 * (long long) k;
                    src = -(k + 1);
                }
                for (int q = 0; q < 9; ++q)
                    b.g_rot[9 * id + q] = rot[q];
                G(b.g_cell)[id] = idx;
                G(b.g_step)[id] = ts;
                G(b.g_src)[id] = src;
                G(b.g_dom)[id] = d;
            }
            rused += rtot;
        }
        __syncthreads();
        if (tid == 0)
        {
            b.cnt->random_used = rnd0 + rused;
            b.cnt->new_grains += nev;
        }
        next += nev;
    }
    __syncthreads();
    if (tid == 0)
        for (int d = 0; d < V.ndom; ++d)
        {
            Counters *cnt = V.dom[d].b.cnt;
            if (over)
            {
                // grain.c:868-888, 361-366: with every slot taken nothing is
                // activated; otherwise createGrains stops the run
                if ((long long) base_id != maxg + 1)
                {
                    cnt->status = MMCA_ERR_GRAIN_CAPACITY;
                    cnt->failed_step = ts;
                }
            }
            else
                cnt->num_grains = (int) total;
            cnt->timestep = ts;
        }
}

// Halo exchange between the subblocks of a coupled batch (ExchangeFacesForVar
// / FinishExchangeForVar, calculate.c:227-330): the interior layer of each
// neighbouring subblock is copied into this subblock's halo layer, face by
// face.  Phase 0 (start of the step): cl and fs of buffer `cb` and dc.  Phase
// 1 (after partitioning): gr, fs of buffer `cb`, and the solid bits of the
// halo cells (mask words of the y/z halo rows, x-halo bytes).  Only face
// cells are exchanged: the step reads no edge or corner halo cell.
// grid.y = domain * 6 + face (face 2a: low side of axis a, 2a+1: high side).
__global__ void __launch_bounds__(256)
k_halo_exchange(View V, int phase, int cb)
{
    const int d = blockIdx.y / 6, f = blockIdx.y % 6;
    const int ax = f / 2, hi = f % 2;
    const DevBufs b = V.dom[d].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[d].c;
    if (c.nsub[0] == 0 || G(b.cnt)->status != MMCA_OK)
        return;                                   // a stopped batch keeps its halos
    // Everything indexed by the face's axis is picked with selects, not array
    // subscripts: a runtime subscript would put the descriptor in private
    // (scratch) memory for every thread.
    const int ns0 = c.nsub[0], ns1 = c.nsub[1], ns2 = c.nsub[2];
    int cx = c.off[0] / c.nx, cy = c.off[1] / c.ny, cz = c.off[2] / c.nz;
    const int na = ax == 0 ? c.nx : (ax == 1 ? c.ny : c.nz);     // cells along the axis
    const int nsa = ax == 0 ? ns0 : (ax == 1 ? ns1 : ns2);
    int ca = (ax == 0 ? cx : (ax == 1 ? cy : cz)) + (hi ? 1 : -1);
    if (ca < 0 || ca >= nsa)
    {
        // a face of the global domain: a mold wall, or a wrapped face whose
        // neighbour is the subblock at the other end of the axis
        // (sbID_from_sbcoords, functions.c:272-311); with one subblock along
        // the axis that is the subblock itself, which exchanges nothing
        // (face_util.c:128, 217) and keeps its halo
        const int wall = hi ? (ax == 0 ? c.pad_hi[0] : (ax == 1 ? c.pad_hi[1] : c.pad_hi[2]))
                            : (ax == 0 ? c.pad_lo[0] : (ax == 1 ? c.pad_lo[1] : c.pad_lo[2]));
        if (wall || nsa < 2)
            return;
        ca = (ca + nsa) % nsa;
    }
    if (ax == 0)
        cx = ca;
    else if (ax == 1)
        cy = ca;
    else
        cz = ca;
    const int r = (cz * ns1 + cy) * ns0 + cx;     // MPI Cart rank (z, y, x)
    const DevBufs nb = V.dom[r].b;
    // the two axes along the face, u (faster) and v, with their strides
    const int nu = ax == 0 ? c.ny : c.nx, nv = ax == 2 ? c.ny : c.nz;
    const long long su = ax == 0 ? c.sx : 1, sv = ax == 2 ? c.sx : c.sxy;
    const long long sa = ax == 0 ? 1 : (ax == 1 ? c.sx : c.sxy);
    const long long nface = (long long) nu * nv;
    const long long t = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    const int hal = hi ? na + 1 : 0;              // halo layer of this subblock
    const int src = hi ? 1 : na;                  // the neighbour's interior layer
    if (t < nface)
    {
        const int qu = (int) (t % nu) + 1, qv = (int) (t / nu) + 1;
        const long long base = qu * su + qv * sv;
        const long long dst = base + hal * sa;
        const long long from = base + src * sa;
        if (phase == 0)
        {
            if (c.f32)
                cl32_buf(b, cb)[dst] = cl32_buf(nb, cb)[from];
            else
                cl_buf(b, cb)[dst] = cl_buf(nb, cb)[from];
            fs_buf(b, cb)[dst] = fs_buf(nb, cb)[from];
            for (int e = 0; e < 3; ++e)
                G(b.dc)[3 * dst + e] = G(nb.dc)[3 * from + e];
        }
        else
        {
            const int g = G(nb.gr)[from];
            G(b.gr)[dst] = g;
            fs_buf(b, cb)[dst] = fs_buf(nb, cb)[from];
            if (ax == 0)
            {
                const long long row = (long long) qv * (c.ny + 2) + qu;   // x face: u = y, v = z
                G(b.xhalo)[2 * row + hi] = (unsigned char) (g > 0);
            }
        }
    }
    if (phase == 1 && ax > 0)
    {
        // mask words of the halo rows: y faces, rows (hal, k) for k = 1..nz;
        // z faces, rows (j, hal) for j = 1..ny
        const int nrow = (ax == 1) ? c.nz : c.ny;
        if (t < (long long) nrow * c.w64)
        {
            const int w = (int) (t % c.w64), i = (int) (t / c.w64) + 1;
            const long long rd = (ax == 1) ? ((long long) i * (c.ny + 2) + hal) : ((long long) hal * (c.ny + 2) + i);
            const long long rs = (ax == 1) ? ((long long) i * (c.ny + 2) + src) : ((long long) src * (c.ny + 2) + i);
            G(b.mask)[rd * c.w64 + w] = G(nb.mask)[rs * c.w64 + w];
        }
    }
}

// Partitioning of one cell (fs_change_diffuse, growth.c:221-273).  The
// temperature, temp(), is only evaluated for cells with gr > 0.  *branch
// receives the MMCA_BRANCH_* code of the branch taken.
template <typename TempFn>
__device__ __forceinline__ void
partition_cell(const Consts &c, TempFn temp, double ce_new, double ce_prev, double conc, double lspfs,
               int g_old, double *cl_out, double *fs_out, int *g_out, int *branch)
{
    int g = g_old;
    *branch = MMCA_BRANCH_LIQUID;
    if (g <= 0)
    {
        *cl_out = ce_new;
        *g_out = 0;
        *fs_out = 0;
        return;
    }
    const double Tcell = temp();
    double c_int = c.cinit + c.m_inv * (Tcell - c.Tliq);
    double Tunder = c.Tliq - Tcell;
    Tunder += c.liq_slope * (conc - c.cinit);
    if (Tcell < 1.0 || Tcell > c.Tliq)
    {
        *cl_out = ce_new;
        *g_out = 0;
        *fs_out = 0;
        return;
    }
    if (lspfs == 1.0 && (Tunder >= 0.0))
    {
        *cl_out = ce_new * c.k_inv;
        *fs_out = lspfs;
        *g_out = g;
        *branch = MMCA_BRANCH_SOLID;
        return;
    }
    double dce = ce_new - ce_prev;
    double dcl = (c_int - conc) * c.D_xt * 2 + dce / (1 - c.km * lspfs);
    double dfs = (dcl * (1. - c.km * lspfs) - dce) * c.km_inv / conc;
    double cl = conc + dcl;
    double fs = lspfs + dfs;
    *branch = (1.0 < fs) ? MMCA_BRANCH_FULL : MMCA_BRANCH_GROW;
    fs = ref_min(1.0, fs);
    if (fs < 0.)
    {
        fs = 0.;
        g = 0;
        cl = ce_new;
        *branch = MMCA_BRANCH_REMELT;
    }
    *cl_out = cl;
    *fs_out = fs;
    *g_out = g;
}

// The temperature the partitioning of cell idx reads (tempUpdateSB_int,
// temperature.c:218-262): the model itself, or with beam_sep the separable
// beam field computed into b.temp before the step (the kernels then run
// their non-AM form).
template <bool AM>
__device__ __forceinline__ double
temp_step(const DevBufs &b, const Consts &c, double sim_time, long long idx, int x, int y, int z)
{
    if constexpr (AM)
        return temp_model<true>(c, sim_time, x, y, z);
    if (c.beam_sep)
        return G(b.temp)[idx];
    return temp_model<false>(c, sim_time, x, y, z);
}

// growth.c:163-165: error() ends every rank.  In a coupled batch the status
// goes to every domain of the view, so that each rank's later kernels (and
// the halo exchange) stop at the same point; the ranks share one timestep.
__device__ __forceinline__ void
flag_negative_ce(const View &V, const Consts &c, Counters *cnt)
{
    if (c.nsub[0] == 0)
    {
        if (atomicCAS(&cnt->status, MMCA_OK, MMCA_ERR_NEGATIVE_CE) == MMCA_OK)
            cnt->failed_step = cnt->timestep;
        return;
    }
    for (int d = 0; d < V.ndom; ++d)
    {
        Counters *o = V.dom[d].b.cnt;
        if (atomicCAS(&o->status, MMCA_OK, MMCA_ERR_NEGATIVE_CE) == MMCA_OK)
            o->failed_step = o->timestep;
    }
}

// Cell-parallel kernel.  DIFF && FSC: fused step reading buffers `cur`,
// writing `nxt`.  DIFF only: ce += nbsum (the caller copies ce to oce
// beforehand).  FSC only: in place on `cur` with ce_old = oce.  FSC writes
// the solid mask (gr > 0).  REC also writes each cell's partitioning branch
// into the domain's branch record.
constexpr int SOL_BX = 64;
constexpr int SOL_BY = 4;
constexpr int SOL_BZ = 8;    // planes per workgroup: spreads the per-workgroup set-up over 8 x 256 cells

template <bool DIFF, bool FSC, bool AM, bool F32, bool REC = false>
__global__ void __launch_bounds__(SOL_BX *SOL_BY)
k_solute(View V, int cur, int nxt)
{
    __shared__ double s_red[SOL_BX * SOL_BY / 64];
    __shared__ int s_stop;
    const int dm = find_dom(V.off_cell, V.ndom, blockIdx.x);
    const DevBufs b = V.dom[dm].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[dm].c;
    // workgroup -> 64 x 4 tile and SOL_BZ planes (32-bit index arithmetic:
    // a domain has fewer than 2^31 cells)
    const unsigned blk = (unsigned) (blockIdx.x - V.off_cell[dm]);
    const unsigned gbx = (unsigned) (c.nx + SOL_BX - 1) / SOL_BX, gby = (unsigned) (c.ny + SOL_BY - 1) / SOL_BY;
    const unsigned rest = blk / gbx;
    const int bx = (int) (blk - rest * gbx);
    const int by = (int) (rest % gby);
    const int bz = (int) (rest / gby);
    Counters *cnt = b.cnt;
    // this kernel itself may raise the status: read it once per workgroup so
    // that every wave takes the same branch before the barrier of block_sum
    if (threadIdx.x == 0 && threadIdx.y == 0)
        s_stop = (G(b.cnt)->status != MMCA_OK);
    __syncthreads();
    if (s_stop)
        return;
    const int x = bx * SOL_BX + threadIdx.x + 1;   // interior x
    const int j = by * SOL_BY + threadIdx.y + 1;
    const int k0 = bz * SOL_BZ + 1;
    const int k1 = min(c.nz, k0 + SOL_BZ - 1);
    const bool rowok = (j <= c.ny);
    const bool active = rowok && (x <= c.nx);
    const double sim_time = G(b.cnt)->timestep * c.dt;
    double lost = 0.0;                                   // FP32 storage rounding of ce
    for (int k = k0; k <= k1; ++k)
    {
        int g_new = 0;
        int br_rec = 0;
        if (active)
        {
            const long long idx = k * c.sxy + (long long) j * c.sx + x;
            const MMCA_GLOBAL double *fsc = fs_buf(b, cur);
            const double ce_old = ld_ce<F32>(b, c, idx);
            const double conc = ld_cl<F32>(b, c, cur, idx);
            const double lspfs = fsc[idx];
            double ce_new = ce_old;
            if (DIFF)
            {
                const long long nb[6] = { idx + 1, idx - 1, idx + c.sx, idx - c.sx, idx + c.sxy, idx - c.sxy };
                // mold of a neighbour: a padded face of the box (interior cells are never mold)
                const int mo[6] = { (x == c.nx) ? c.pad_hi[0] : 0, (x == 1) ? c.pad_lo[0] : 0,
                                    (j == c.ny) ? c.pad_hi[1] : 0, (j == 1) ? c.pad_lo[1] : 0,
                                    (k == c.nz) ? c.pad_hi[2] : 0, (k == 1) ? c.pad_lo[2] : 0 };
                double cln[6], fsn[6];
#pragma unroll
                for (int n = 0; n < 6; ++n)
                {
                    cln[n] = ld_cl<F32>(b, c, cur, nb[n]);
                    fsn[n] = fsc[nb[n]];
                }
                ce_new = diffuse_cell(c, ce_old, conc, lspfs, cln, fsn, mo);
                if (c.check_neg && ce_new < 0)
                    flag_negative_ce(V, c, cnt);
                lost += st_ce<F32>(b, c, idx, ce_new);
            }
            if (FSC)
            {
                const double ce_prev = DIFF ? ce_old : G(b.oce)[idx];
                const int g_old = G(b.gr)[idx];
                double cl_out, fs_out;
                int g, branch;
                partition_cell(c, [&]() { return temp_step<AM>(b, c, sim_time, idx, x, j, k); }, ce_new, ce_prev,
                               conc, lspfs, g_old, &cl_out, &fs_out, &g, &branch);
                br_rec = branch;
                st_cl<F32>(b, c, nxt, idx, cl_out);
                st_stream(fs_buf(b, nxt), idx, fs_out);
                if (g != g_old)
                    G(b.gr)[idx] = g;
                keep_stale_d(c, b, idx, g_old, g, lspfs);
                g_new = g;
            }
        }
        if (FSC)
        {
            unsigned long long bits = __ballot(active && g_new > 0);
            if (threadIdx.x == 0 && rowok)
                G(b.mask)[((long long) k * (c.ny + 2) + j) * c.w64 + bx] = bits;
        }
        if (REC)
        {
            // branch record of this step, three bit planes per 64-cell segment
            // (bounds: set by the host, checked here too)
            const unsigned long long p0 = __ballot(active && (br_rec & 1));
            const unsigned long long p1 = __ballot(active && (br_rec & 2));
            const unsigned long long p2 = __ballot(active && (br_rec & 4));
            const long long rs = (long long) (G(b.cnt)->timestep - b.rec_t0) - 1;
            if (threadIdx.x == 0 && rowok && rs >= 0 && rs < b.rec_steps)
            {
                MMCA_GLOBAL unsigned long long *w = (MMCA_GLOBAL unsigned long long *) b.rec +
                                                    rs * rec_words(c) + rec_word(c, bx, j, k);
                w[0] = p0;
                w[1] = p1;
                w[2] = p2;
            }
        }
    }
    if (F32)
    {
        const double tot = block_sum<SOL_BX * SOL_BY>(lost, s_red);
        if (threadIdx.x == 0 && threadIdx.y == 0 && tot != 0.0)
            ledger_add(b, blk, tot);
    }
}

// Raw storage of ce and cl (FP64 values, or FP32 deviations from c0), kept
// unconverted while a load is in flight, so that a prefetched value is only
// waited for where it is used.
template <bool F32> struct Store { using T = double; };
template <> struct Store<true> { using T = float; };

template <bool F32>
__device__ __forceinline__ typename Store<F32>::T
raw_ce(const DevBufs &b, long long i)
{
    if constexpr (F32)
        return G(b.ce32)[i];
    else
        return G(b.ce)[i];
}

template <bool F32>
__device__ __forceinline__ typename Store<F32>::T
raw_cl(const DevBufs &b, int buf, long long i)
{
    if constexpr (F32)
        return cl32_buf(b, buf)[i];
    else
        return cl_buf(b, buf)[i];
}

template <bool F32>
__device__ __forceinline__ double
from_raw(const Consts &c, typename Store<F32>::T r)
{
    if constexpr (F32)
        return c.cinit + (double) r;
    else
        return r;
}

// What the z-marching kernel loads for plane q one plane ahead: ce and gr of
// q, the tile-edge neighbours of q, and cl and fs of q+1 (the +z neighbour at
// plane q, the cell's own value at q+1).
template <bool F32>
struct ZmAhead
{
    typename Store<F32>::T ce, cl2, clx, cly;
    double fs2, fsx, fsy;
    int g;
};

// z-marching fused kernel: a workgroup of 64 x BY threads owns a 64 x BY
// column tile and walks `kz` planes upward.  cl/fs of the planes k-1, k, k+1
// stay in registers, x neighbours come from the neighbouring lanes (wave
// shuffles), y neighbours from the rows of the other waves in LDS.
//
// Every load of plane k+1 (ce, gr, the tile-edge neighbours, and cl, fs of
// plane k+2) is issued before plane k is computed, into one of two register
// sets used in turn, so that it is in flight for a whole plane.  Lanes one
// past the last column or row of the domain load the padding cell as their
// own value (and compute nothing), so the cell next to the boundary reads it
// by shuffle or LDS like any other neighbour; then only lane 0 (x-1) and lane
// 63 (x+1), and only row 0 (y-1) and row BY-1 (y+1), read neighbours from
// memory, one of each kind per lane.
template <int BY, bool AM, bool F32>
__global__ void __launch_bounds__(64 * BY)
k_solute_zm(View V, int cur, int nxt)
{
    static_assert(BY >= 2, "row 0 and row BY-1 must be different waves");
    __shared__ double s_cl[BY][64];
    __shared__ double s_fs[BY][64];
    __shared__ double s_red[BY];
    __shared__ int s_stop;
    // Gaussian-beam tables (AM only): the cell-independent factors of the
    // 400 quadrature terms, and the plane-dependent one for the current plane
    __shared__ double s_beam[AM ? sizeof(BeamTab) / sizeof(double) : 1];
    BeamTab *beam = reinterpret_cast<BeamTab *>(s_beam);
    // Tile of this workgroup.  With xcd_remap the workgroups that the
    // hardware hands to one XCD (linear id mod 8) take a contiguous run of
    // tiles (x fastest, then y, then z, then domain), so that tiles sharing
    // halo rows and planes run on the same L2.
    long long gtile = blockIdx.x;
    if (V.xcd_remap)
    {
        const long long per = (V.tot_zm + 7) / 8;
        gtile = (blockIdx.x % 8) * per + blockIdx.x / 8;
    }
    if (gtile >= V.tot_zm)
        return;                                   // uniform over the workgroup
    const int dm = find_dom(V.off_zm, V.ndom, gtile);
    const long long tile = gtile - V.off_zm[dm];
    const DevBufs b = V.dom[dm].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[dm].c;
    const int kz = c.zm_planes;
    Counters *cnt = b.cnt;
    // this kernel itself may raise the status: read it once per workgroup so
    // that every wave takes the same branch before the barriers below
    if (threadIdx.x == 0 && threadIdx.y == 0)
        s_stop = (G(b.cnt)->status != MMCA_OK);
    __syncthreads();
    if (s_stop)
        return;
    const int lane = threadIdx.x;
    const int ty = threadIdx.y;
    if constexpr (AM)
        beam_tab_terms(c, *beam, ty * 64 + lane, 64 * BY);
    const int ntx = (c.nx + 63) / 64, nty = (c.ny + BY - 1) / BY;
    const int bx = (int) (tile % ntx);
    const int by = (int) ((tile / ntx) % nty);
    const int bz = (int) (tile / ((long long) ntx * nty));
    const int x = bx * 64 + lane + 1;
    const int j = by * BY + ty + 1;
    const int k0 = bz * kz + 1;
    const int k1 = min(c.nz, k0 + kz - 1);
    const bool rowok = (j <= c.ny);
    const bool active = rowok && (x <= c.nx);
    // Own cl/fs: every computing cell, and the padding cells one past the
    // last column and row (the neighbours of the boundary cells).  Loads are
    // issued by every lane, at addresses clamped into the padded box, so that
    // no branch skips one and each wait covers exactly the plane it needs.
    const int xc = min(x, c.nx + 1), jc = min(j, c.ny + 1);
    const long long col = (long long) jc * c.sx + xc;
    const long long row0 = (long long) jc * c.sx + bx * 64;   // x-1 of lane 0
    const MMCA_GLOBAL double *fsc = fs_buf(b, cur);
    MMCA_GLOBAL double *fsn_out = fs_buf(b, nxt);
    const double sim_time = G(b.cnt)->timestep * c.dt;
    const bool xm_mem = (lane == 0);
    const bool xp_mem = (lane == 63);
    const bool ym_mem = (ty == 0);
    const bool yp_mem = (ty == BY - 1);
    // tile-edge neighbours: lane 63 reads x+1, every other lane the x-1 of
    // lane 0 (one shared line); row 0 reads y-1, row BY-1 y+1, the other rows
    // the first cell of their row
    const long long xe = xp_mem ? col + 1 : row0;
    const long long ye = ym_mem ? col - c.sx : yp_mem ? col + c.sx : row0 + 1;
    const int mo_x1 = (x == c.nx) ? c.pad_hi[0] : 0, mo_x0 = (x == 1) ? c.pad_lo[0] : 0;
    const int mo_y1 = (j == c.ny) ? c.pad_hi[1] : 0, mo_y0 = (j == 1) ? c.pad_lo[1] : 0;
    auto fetch = [&](ZmAhead<F32> &a, int q) {
        // plane q's ce, gr and edge neighbours; plane q+1's cl and fs
        const long long pq = q * c.sxy;
        a.cl2 = raw_cl<F32>(b, cur, pq + c.sxy + col);
        a.fs2 = fsc[pq + c.sxy + col];
        a.ce = raw_ce<F32>(b, pq + col);
        a.g = G(b.gr)[pq + col];
        a.clx = raw_cl<F32>(b, cur, pq + xe);
        a.fsx = fsc[pq + xe];
        a.cly = raw_cl<F32>(b, cur, pq + ye);
        a.fsy = fsc[pq + ye];
    };
    double clm = from_raw<F32>(c, raw_cl<F32>(b, cur, (k0 - 1) * c.sxy + col));
    double fsm = fsc[(k0 - 1) * c.sxy + col];
    double clv = from_raw<F32>(c, raw_cl<F32>(b, cur, k0 * c.sxy + col));
    double fsv = fsc[k0 * c.sxy + col];
    // Loads run one plane ahead, in two register sets used in turn.  (Two
    // planes ahead, with a third set, measured slower.)
    ZmAhead<F32> P0, P1;
    fetch(P0, k0);
    // the loop starts with nothing in flight, so that its waits are the
    // steady-state ones
    __builtin_amdgcn_s_waitcnt(0);
    double lost = 0.0;
    bool neg = false;                             // a negative ce seen by this lane
    // one plane: `ld` (plane k-1's set) receives plane k+1, `use` holds plane k
    auto plane = [&](int k, ZmAhead<F32> &ld, ZmAhead<F32> &use) {
        const long long idx = k * c.sxy + col;
        if (k > k0)
        {
            clm = clv;
            fsm = fsv;
            clv = from_raw<F32>(c, ld.cl2);       // landed: it was plane k-1's +z neighbour
            fsv = ld.fs2;
        }
        // near the top this reloads plane k1 (unused): every plane issues
        // the same loads, so every wait is the same
        fetch(ld, min(k + 1, k1));
        if constexpr (AM)
            beam_tab_plane(c, *beam, k, ty * 64 + lane, 64 * BY);   // read after the barrier below
        s_cl[ty][lane] = clv;
        s_fs[ty][lane] = fsv;
        const double shxm = __shfl_up(clv, 1, 64), sfxm = __shfl_up(fsv, 1, 64);
        const double shxp = __shfl_down(clv, 1, 64), sfxp = __shfl_down(fsv, 1, 64);
        __syncthreads();
        int g_new = 0;
        if (active)
        {
            const double ex_cl = from_raw<F32>(c, use.clx), ey_cl = from_raw<F32>(c, use.cly);
            const double clxm = xm_mem ? ex_cl : shxm, fsxm = xm_mem ? use.fsx : sfxm;
            const double clxp = xp_mem ? ex_cl : shxp, fsxp = xp_mem ? use.fsx : sfxp;
            const double clym = ym_mem ? ey_cl : s_cl[ym_mem ? 0 : ty - 1][lane];
            const double fsym = ym_mem ? use.fsy : s_fs[ym_mem ? 0 : ty - 1][lane];
            const double clyp = yp_mem ? ey_cl : s_cl[yp_mem ? 0 : ty + 1][lane];
            const double fsyp = yp_mem ? use.fsy : s_fs[yp_mem ? 0 : ty + 1][lane];
            const double clp = from_raw<F32>(c, use.cl2), fsp = use.fs2;
            const double ce_old = from_raw<F32>(c, use.ce);
            const int g_old = use.g;
            const double cln[6] = { clxp, clxm, clyp, clym, clp, clm };
            const double fsn[6] = { fsxp, fsxm, fsyp, fsym, fsp, fsm };
            const int mo[6] = { mo_x1, mo_x0, mo_y1, mo_y0, (k == c.nz) ? c.pad_hi[2] : 0,
                                (k == 1) ? c.pad_lo[2] : 0 };
            const double ce_new = diffuse_cell(c, ce_old, clv, fsv, cln, fsn, mo);
            neg |= (ce_new < 0);
            lost += st_ce<F32>(b, c, idx, ce_new);
            double cl_out, fs_out;
            int branch;
            partition_cell(
                c,
                [&]() {
                    if constexpr (AM)
                        return beam_temp_tab(c, *beam, beam->Z, sim_time, x, j, k);
                    else
                        return temp_step<false>(b, c, sim_time, idx, x, j, k);
                },
                ce_new, ce_old, clv, fsv, g_old, &cl_out, &fs_out, &g_new, &branch);
            st_cl<F32>(b, c, nxt, idx, cl_out);
            st_stream(fsn_out, idx, fs_out);
            if (g_new != g_old)
                G(b.gr)[idx] = g_new;
            keep_stale_d(c, b, idx, g_old, g_new, fsv);
        }
        unsigned long long bits = __ballot(active && g_new > 0);
        if (rowok && lane == 0)
            G(b.mask)[((long long) k * (c.ny + 2) + j) * c.w64 + bx] = bits;
        __syncthreads();
    };
    // Whole rounds of two planes, so that the loop is only re-entered after
    // a whole round and its waits see the same loads on every path; then the
    // plane left over.
    int k = k0;
    for (; k + 1 <= k1; k += 2)
    {
        plane(k, P1, P0);
        plane(k + 1, P0, P1);
    }
    if (k <= k1)
        plane(k, P1, P0);
    // status and step of a negative ce (any plane of the tile: the step is
    // the same for all)
    if (c.check_neg && neg)
        flag_negative_ce(V, c, cnt);
    if (F32)
    {
        const double tot = block_sum<64 * BY>(lost, s_red);
        if (lane == 0 && ty == 0 && tot != 0.0)
            ledger_add(b, tile, tot);
    }
}

// ---------------------------------------------------------------------------
// z-marching fused kernel with its planes staged in LDS by direct global->LDS
// loads (gfx950 global_load_lds_dword): k_solute_zl.
//
// A workgroup of 64 x BY threads owns a 64 x BY column tile and walks its
// planes upward, as k_solute_zm does, with the same arithmetic per cell.  The
// planes come through a ring of ZL_RING slots in LDS instead of registers:
// each slot holds one plane's cl and fs over the tile with a one-cell halo in
// x and y ((BY + 2) rows x 66), and its ce and gr over the tile.  The loads
// write LDS directly, so a plane in flight costs no VGPRs, and two planes are
// in flight while one is computed.  Every wave issues an equal share of a
// slot's loads (64 lanes x 4 bytes each, per-lane offsets computed once: they
// are the same for every plane), then waits for its own share of the plane it
// needs next; the one barrier per plane makes every wave's share visible and
// frees the slot of the plane before the last.
//
// The loads are issued from inline asm, which the compiler's wait-count pass
// does not see: it would otherwise make every LDS read wait for all loads in
// flight.  Their completion is waited for explicitly with vmcnt(N), where N
// counts only loads of this ring issued after the awaited ones; any other
// memory operation issued in between makes that wait stricter, never looser.
// Every load address is clamped into the padded box of its field.
constexpr int ZL_RING = 4;

template <int BY, bool F32>
struct ZlSlot {
    static constexpr int E = F32 ? 4 : 8;          // bytes per stored ce, cl value
    static constexpr int HR = BY + 2, HW = 66;     // halo'd rows and row width
    static constexpr int pad64(int n) { return (n + 63) / 64 * 64; }
    static constexpr int CL_DW = pad64(HR * HW * E / 4);
    static constexpr int FS_DW = pad64(HR * HW * 2);
    static constexpr int CE_DW = pad64(BY * 64 * E / 4);
    static constexpr int GR_DW = pad64(BY * 64);
    static constexpr int CL0 = 0, FS0 = CL_DW, CE0 = FS0 + FS_DW, GR0 = CE0 + CE_DW;
    static constexpr int DW = GR0 + GR_DW;         // dwords per slot
    static constexpr int NI = DW / 64;             // load instructions per slot
    static constexpr int NPW = (NI + BY - 1) / BY; // per wave: loads NPW w .. NPW w + NPW - 1 of the slot
    // field region of load i (every region is whole loads: cl, fs, ce, gr)
    static constexpr int region_of(int i)
    {
        return i * 64 < FS0 ? 0 : i * 64 < CE0 ? 1 : i * 64 < GR0 ? 2 : 3;
    }
    // a wave's loads span at most two regions, so it keeps two base pointers
    static constexpr bool two_regions()
    {
        for (int w = 0; w < BY; ++w)
        {
            const int a = w * NPW < NI - 1 ? w * NPW : NI - 1;
            const int z = w * NPW + NPW - 1 < NI - 1 ? w * NPW + NPW - 1 : NI - 1;
            if (region_of(z) - region_of(a) > 1)
                return false;
        }
        return true;
    }
    static_assert(two_regions(), "a wave's share of a slot must lie in at most two field regions");
};

// One dword per lane from global (sbase + voff) into LDS (lds + lane * 4).
__device__ __forceinline__ void
zl_load(const void *sbase, unsigned voff, unsigned lds)
{
    unsigned save;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %3\n\t"
                 "global_load_lds_dword %1, %2\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(save)
                 : "v"(voff), "s"(sbase), "s"(lds)
                 : "memory");
}

template <int N>
__device__ __forceinline__ void
zl_wait()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

template <int BY, bool AM, bool F32>
__global__ void __launch_bounds__(64 * BY)
k_solute_zl(View V, int cur, int nxt)
{
    using L = ZlSlot<BY, F32>;
    using T = typename Store<F32>::T;
    static_assert(L::NPW * 3 + 16 < 64, "loads in flight must stay within the vmcnt range");
    __shared__ __attribute__((aligned(16))) unsigned s_ring[ZL_RING * L::DW];
    __shared__ double s_red[BY];
    __shared__ int s_stop;
    __shared__ double s_beam[AM ? sizeof(BeamTab) / sizeof(double) : 1];
    __shared__ double s_z[AM ? 2 : 1][AM ? BEAM_NT : 1];      // plane factors of the beam, double-buffered
    // the other temperature models depend on z only (on y in 2-D runs): one
    // value per plane of the tile, computed before the walk, so that their
    // constants need no registers during it; dynamic LDS of V.zm_planes_max
    // doubles (none for the beam), so that short tiles leave room for more
    // workgroups per CU
    extern __shared__ double s_T[];
    BeamTab *beam = reinterpret_cast<BeamTab *>(s_beam);
    long long gtile = blockIdx.x;
    if (V.xcd_remap)
    {
        const long long per = (V.tot_zm + 7) / 8;
        gtile = (blockIdx.x % 8) * per + blockIdx.x / 8;
    }
    if (gtile >= V.tot_zm)
        return;                                   // uniform over the workgroup
    const int dm = find_dom(V.off_zm, V.ndom, gtile);
    const long long tile = gtile - V.off_zm[dm];
    const DevBufs b = V.dom[dm].b;
    const Consts c = V.dom[dm].c;
    const int kz = c.zm_planes;
    Counters *cnt = b.cnt;
    if (threadIdx.x == 0 && threadIdx.y == 0)
        s_stop = (G(b.cnt)->status != MMCA_OK);
    __syncthreads();
    if (s_stop)
        return;
    const int lane = threadIdx.x;
    const int ty = threadIdx.y;
    const int wv = __builtin_amdgcn_readfirstlane(ty);
    const int ntx = (c.nx + 63) / 64, nty = (c.ny + BY - 1) / BY;
    const int bx = (int) (tile % ntx);
    const int by = (int) ((tile / ntx) % nty);
    const int bz = (int) (tile / ((long long) ntx * nty));
    const int x0 = bx * 64 + 1, j0 = by * BY + 1;
    const int x = x0 + lane;
    const int j = j0 + ty;
    const int k0 = bz * kz + 1;
    const int k1 = min(c.nz, k0 + kz - 1);
    const int n = k1 - k0 + 1;
    const bool rowok = (j <= c.ny);
    const bool active = rowok && (x <= c.nx);

    // this wave's share of a slot: loads i = NPW wv + m (the last load of the
    // slot repeated past its end), which lie in at most two field regions rA
    // and rB (rB from load mB on); each load reads at per-lane byte offset
    // off[m] within a plane of its field.  Few scalar registers hold this
    // state (two bases, two strides), so the kernel's constants stay in
    // scalar registers too.
    const char *fbase[4] = { (const char *) (F32 ? (const void *) cl32_buf(b, cur) : (const void *) cl_buf(b, cur)),
                             (const char *) fs_buf(b, cur),
                             (const char *) (F32 ? (const void *) b.ce32 : (const void *) b.ce),
                             (const char *) b.gr };
    const long long fplane[4] = { c.sxy * L::E, c.sxy * 8, c.sxy * L::E, c.sxy * 4 };
    const int ibase = wv * L::NPW;
    const int rA = L::region_of(min(ibase, L::NI - 1)), rB = L::region_of(min(ibase + L::NPW - 1, L::NI - 1));
    const int rstart = rB == 0 ? 0 : rB == 1 ? L::FS0 / 64 : rB == 2 ? L::CE0 / 64 : L::GR0 / 64;
    const int mB = rA == rB ? 0 : rstart - ibase;              // loads m >= mB are in rB
    const char *baseA = (rA == 0) ? fbase[0] : (rA == 1) ? fbase[1] : (rA == 2) ? fbase[2] : fbase[3];
    const char *baseB = (rB == 0) ? fbase[0] : (rB == 1) ? fbase[1] : (rB == 2) ? fbase[2] : fbase[3];
    const long long strA = (rA == 0) ? fplane[0] : (rA == 1) ? fplane[1] : (rA == 2) ? fplane[2] : fplane[3];
    const long long strB = (rB == 0) ? fplane[0] : (rB == 1) ? fplane[1] : (rB == 2) ? fplane[2] : fplane[3];
    unsigned off[L::NPW];
#pragma unroll
    for (int m = 0; m < L::NPW; ++m)
    {
        const int i = min(ibase + m, L::NI - 1);
        const int d = i * 64 + lane;
        int dw, esz, rows_halo;
        if (d < L::FS0)
            dw = d - L::CL0, esz = L::E, rows_halo = 1;
        else if (d < L::CE0)
            dw = d - L::FS0, esz = 8, rows_halo = 1;
        else if (d < L::GR0)
            dw = d - L::CE0, esz = L::E, rows_halo = 0;
        else
            dw = d - L::GR0, esz = 4, rows_halo = 0;
        const int per = esz / 4;                          // dwords per value
        const int width = rows_halo ? L::HW : 64;
        const int nval = rows_halo ? L::HR * L::HW : BY * 64;
        const int e = min(dw / per, nval - 1);            // padding dwords repeat the last value
        const int half = dw - (dw / per) * per;
        const int row = e / width, col = e - row * width;
        const int gj = min(j0 - rows_halo + row, c.ny + 1);
        const int gx = min(x0 - rows_halo + col, c.nx + 1);
        off[m] = (unsigned) ((gj * c.sx + gx) * esz + half * 4);
    }
    const unsigned ring_lds = (unsigned) (unsigned long) (__attribute__((address_space(3))) unsigned *) s_ring;
    // loads of plane sequence number s (plane k0 - 1 + s, clamped to k1 + 1) into slot s % ZL_RING
    auto issue = [&](int s) {
        const long long p = min(k0 - 1 + s, k1 + 1);
        const unsigned slot = ring_lds + (unsigned) ((s % ZL_RING) * L::DW) * 4u;
        const char *pa = baseA + p * strA, *pb = baseB + p * strB;
#pragma unroll
        for (int m = 0; m < L::NPW; ++m)
            zl_load(m >= mB ? pb : pa, off[m], slot + (unsigned) min(ibase + m, L::NI - 1) * 256u);
    };
    auto cl_at = [&](int q, int row, int col) -> double {
        const unsigned *sl = s_ring + q * L::DW + L::CL0;
        return from_raw<F32>(c, reinterpret_cast<const T *>(sl)[row * L::HW + col]);
    };
    auto fs_at = [&](int q, int row, int col) -> double {
        const unsigned *sl = s_ring + q * L::DW + L::FS0;
        return reinterpret_cast<const double *>(sl)[row * L::HW + col];
    };

    const MMCA_GLOBAL double *fsc = fs_buf(b, cur);
    (void) fsc;
    MMCA_GLOBAL double *fsn_out = fs_buf(b, nxt);
    const double sim_time = G(b.cnt)->timestep * c.dt;
    const int mo_x1 = (x == c.nx) ? c.pad_hi[0] : 0, mo_x0 = (x == 1) ? c.pad_lo[0] : 0;
    const int mo_y1 = (j == c.ny) ? c.pad_hi[1] : 0, mo_y0 = (j == 1) ? c.pad_lo[1] : 0;

    issue(0);
    issue(1);
    issue(2);
    issue(3);
    if constexpr (AM)
    {
        beam_tab_terms(c, *beam, ty * 64 + lane, 64 * BY);
        beam_tab_plane(c, s_z[1], k0, ty * 64 + lane, 64 * BY);
    }
    double t_row = 0.0;                          // 2-D runs: the row's temperature (all planes)
    if constexpr (!AM)
    {
        if (c.gdimz == 1)
            t_row = temp_model<false>(c, sim_time, x, j, k0);
        else
            for (int e = ty * 64 + lane; e < n; e += 64 * BY)
                s_T[e] = temp_model<false>(c, sim_time, x0, j0, k0 + e);   // read after the loop's first barrier
    }
    zl_wait<2 * L::NPW>();                       // planes k0 - 1 and k0 have landed
    __syncthreads();
    double clm = cl_at(0, ty + 1, lane + 1), fsm = fs_at(0, ty + 1, lane + 1);
    double clv = cl_at(1, ty + 1, lane + 1), fsv = fs_at(1, ty + 1, lane + 1);
    double lost = 0.0;
    bool neg = false;
    for (int s = 1; s <= n; ++s)
    {
        const int k = k0 - 1 + s;
        const int q = s % ZL_RING, q1 = (s + 1) % ZL_RING;
        zl_wait<L::NPW>();                       // this wave's share of plane k + 1 has landed
        __syncthreads();                         // every wave's share; the slot of plane k - 1 is free
        issue(s + 3);
        if constexpr (AM)
            if (s < n)
                beam_tab_plane(c, s_z[(s + 1) & 1], k + 1, ty * 64 + lane, 64 * BY);
        const double clp = cl_at(q1, ty + 1, lane + 1), fsp = fs_at(q1, ty + 1, lane + 1);
        int g_new = 0;
        if (active)
        {
            const long long idx = k * c.sxy + (long long) j * c.sx + x;
            const double clxm = cl_at(q, ty + 1, lane), fsxm = fs_at(q, ty + 1, lane);
            const double clxp = cl_at(q, ty + 1, lane + 2), fsxp = fs_at(q, ty + 1, lane + 2);
            const double clym = cl_at(q, ty, lane + 1), fsym = fs_at(q, ty, lane + 1);
            const double clyp = cl_at(q, ty + 2, lane + 1), fsyp = fs_at(q, ty + 2, lane + 1);
            const unsigned *sl = s_ring + q * L::DW;
            const double ce_old = from_raw<F32>(c, reinterpret_cast<const T *>(sl + L::CE0)[ty * 64 + lane]);
            const int g_old = reinterpret_cast<const int *>(sl + L::GR0)[ty * 64 + lane];
            const double cln[6] = { clxp, clxm, clyp, clym, clp, clm };
            const double fsn[6] = { fsxp, fsxm, fsyp, fsym, fsp, fsm };
            const int mo[6] = { mo_x1, mo_x0, mo_y1, mo_y0, (k == c.nz) ? c.pad_hi[2] : 0,
                                (k == 1) ? c.pad_lo[2] : 0 };
            const double ce_new = diffuse_cell(c, ce_old, clv, fsv, cln, fsn, mo);
            neg |= (ce_new < 0);
            lost += st_ce<F32>(b, c, idx, ce_new);
            double cl_out, fs_out;
            int branch;
            partition_cell(
                c,
                [&]() {
                    if constexpr (AM)
                        return beam_temp_tab(c, *beam, s_z[s & 1], sim_time, x, j, k);
                    else if (c.beam_sep)
                        return G(b.temp)[idx];
                    else
                        return c.gdimz == 1 ? t_row : s_T[s - 1];
                },
                ce_new, ce_old, clv, fsv, g_old, &cl_out, &fs_out, &g_new, &branch);
            st_cl<F32>(b, c, nxt, idx, cl_out);
            st_stream(fsn_out, idx, fs_out);
            if (g_new != g_old)
                G(b.gr)[idx] = g_new;
            keep_stale_d(c, b, idx, g_old, g_new, fsv);
        }
        unsigned long long bits = __ballot(active && g_new > 0);
        if (rowok && lane == 0)
            G(b.mask)[((long long) k * (c.ny + 2) + j) * c.w64 + bx] = bits;
        clm = clv;
        fsm = fsv;
        clv = clp;
        fsv = fsp;
    }
    zl_wait<0>();                                // no load may land in LDS after the workgroup ends
    if (c.check_neg && neg)
        flag_negative_ce(V, c, cnt);
    if (F32)
    {
        const double tot = block_sum<64 * BY>(lost, s_red);
        if (lane == 0 && ty == 0 && tot != 0.0)
            ledger_add(b, tile, tot);
    }
}

// Solid mask from gr (used when the list is built outside the fused step).
// Mask layout: per padded row (j,k), w64 words; bit b of word w is interior
// x = 64 w + b + 1.
__global__ void
k_mask_from_gr(DevBufs b, Consts c)
{
    const int x = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    bool solid = false;
    if (x <= c.nx)
        solid = G(b.gr)[k * c.sxy + (long long) j * c.sx + x] > 0;
    unsigned long long bits = __ballot(solid);
    if (threadIdx.x == 0)
        G(b.mask)[((long long) k * (c.ny + 2) + j) * c.w64 + blockIdx.x] = bits;
}

// ---------------------------------------------------------------------------
// Interface list: interior cells with gr <= 0 and a neighbour with gr > 0
// (growth.c:496-513), in raster order.  One thread per 64-bit mask word.
constexpr int LIST_BS = 256;

__host__ __device__ __forceinline__ long long
list_words_d(const Consts &c)
{
    return (long long) c.nz * c.ny * c.w64;
}

__device__ __forceinline__ unsigned long long
mask_word(const DevBufs &b, const Consts &c, int k, int j, int w)
{
    return G(b.mask)[((long long) k * (c.ny + 2) + j) * c.w64 + w];
}

// mask word wi (< nz * ny * w64 < 2^31) -> padded plane k, row j, word w;
// 32-bit arithmetic (64-bit division is a long instruction sequence)
__device__ __forceinline__ void
word_pos(const Consts &c, long long wi, int *k, int *j, int *w)
{
    const unsigned u = (unsigned) wi, w64 = (unsigned) c.w64, ny = (unsigned) c.ny;
    const unsigned nrow = u / w64;
    *w = (int) (u - nrow * w64);
    const unsigned kk = nrow / ny;
    *j = (int) (nrow - kk * ny) + 1;
    *k = (int) kk + 1;
}

// Solid bits of the x-1 (left) and x+1 (right) neighbours of the cells of
// mask word w; the halo cells x = 0 and x = nx+1 come from the row's x-halo
// bytes (solid only where another subblock of a coupled batch lies there).
__device__ __forceinline__ void
x_neighbour_bits(const DevBufs &b, const Consts &c, int k, int j, int w, unsigned long long own,
                 unsigned long long *left, unsigned long long *right)
{
    *left = own << 1;
    *right = own >> 1;
    const long long row = (long long) k * (c.ny + 2) + j;
    if (w > 0)
        *left |= mask_word(b, c, k, j, w - 1) >> 63;
    else
        *left |= (unsigned long long) (G(b.xhalo)[2 * row] != 0);
    if (w + 1 < c.w64)
        *right |= mask_word(b, c, k, j, w + 1) << 63;
    else
        *right |= (unsigned long long) (G(b.xhalo)[2 * row + 1] != 0) << (c.nx - 1 - 64 * w);
}

__device__ __forceinline__ unsigned long long
member_bits(const DevBufs &b, const Consts &c, long long wi, int *pk, int *pj, int *pw)
{
    int k, j, w;
    word_pos(c, wi, &k, &j, &w);
    *pk = k;
    *pj = j;
    *pw = w;
    unsigned long long own = mask_word(b, c, k, j, w);
    unsigned long long left, right;
    x_neighbour_bits(b, c, k, j, w, own, &left, &right);
    unsigned long long yz = mask_word(b, c, k, j - 1, w) | mask_word(b, c, k, j + 1, w) |
                            mask_word(b, c, k - 1, j, w) | mask_word(b, c, k + 1, j, w);
    // bit b of word w is interior x = 64 w + b + 1 <= nx
    const int hi = c.nx - w * 64;
    const unsigned long long inter = (hi >= 64) ? ~0ull : ((1ull << hi) - 1ull);
    return ~own & (left | right | yz) & inter;
}

__global__ void __launch_bounds__(LIST_BS)
k_list_count(View V)
{
    __shared__ int s_w[LIST_BS / 64];
    const int dm = find_dom(V.off_list, V.ndom, blockIdx.x);
    const long long blk = blockIdx.x - V.off_list[dm];
    const DevBufs b = V.dom[dm].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[dm].c;
    const long long nwords = list_words_d(c);
    if (G(b.cnt)->status != MMCA_OK)
        return;
    const long long wi = blk * LIST_BS + threadIdx.x;
    int v = 0;
    if (wi < nwords)
    {
        int k, j, w;
        v = __popcll(member_bits(b, c, wi, &k, &j, &w));
    }
    int tot;
    block_exscan_int<LIST_BS / 64>(v, &tot, s_w);
    if (threadIdx.x == 0)
        G(b.blk_count)[blk] = tot;
}

// Exclusive scan of the block counts (one workgroup).  When decide_realloc
// is set, also applies the dc_tmp growth rule of growth.c:644-660.
constexpr int SCAN_BS = 1024;

__global__ void __launch_bounds__(SCAN_BS)
k_list_scan(View V, int decide_realloc)
{
    __shared__ int s_w[SCAN_BS / 64];
    const DevBufs b = V.dom[blockIdx.x].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[blockIdx.x].c;
    const int nblk = (int) ((list_words_d(c) + LIST_BS - 1) / LIST_BS);
    Counters *cnt = b.cnt;
    if (cnt->status != MMCA_OK)
        return;
    int carry = 0;
    for (int base = 0; base < nblk; base += SCAN_BS)
    {
        int i = base + threadIdx.x;
        int v = (i < nblk) ? G(b.blk_count)[i] : 0;
        int tot;
        int ex = block_exscan_int<SCAN_BS / 64>(v, &tot, s_w);
        if (i < nblk)
            G(b.blk_count)[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0)
    {
        cnt->gindex = carry;
        if (decide_realloc)
        {
            if (carry > cnt->dc_tmp_size)
            {
                cnt->realloc_old = cnt->dc_tmp_size;
                cnt->dc_tmp_size = carry + 100;
            }
            else
                cnt->realloc_old = -1;
        }
    }
}

__global__ void
k_realloc_decide(DevBufs b)
{
    Counters *cnt = b.cnt;
    if (cnt->status != MMCA_OK)
        return;
    if (cnt->gindex > cnt->dc_tmp_size)
    {
        cnt->realloc_old = cnt->dc_tmp_size;
        cnt->dc_tmp_size = cnt->gindex + 100;
    }
    else
        cnt->realloc_old = -1;
}

// Interface list writer (grow_cell_reduction, growth.c:496-513): slot
// base+i+1 of diff_id gets the padded index of the workgroup's member i, in
// raster order.  The members of a workgroup's 256 mask words are spread
// evenly over its threads: member i is found by a binary search over the
// per-word prefix counts in LDS.
__device__ __forceinline__ int
nth_set_bit(unsigned long long m, int r)
{
    int pos = 0;
#pragma unroll
    for (int half = 32; half > 0; half >>= 1)
    {
        const unsigned long long lo = m & ((1ull << half) - 1ull);
        const int c = __popcll(lo);
        if (r >= c)
        {
            r -= c;
            m >>= half;
            pos += half;
        }
        else
            m = lo;
    }
    return pos;
}

__global__ void __launch_bounds__(LIST_BS)
k_list_write(View V)
{
    __shared__ int s_w[LIST_BS / 64];
    __shared__ unsigned long long s_mem[LIST_BS];
    __shared__ unsigned long long s_nb[6][LIST_BS];   // solid bits of the 6 neighbours, per word
    __shared__ int s_ex[LIST_BS];
    const int dm = find_dom(V.off_list, V.ndom, blockIdx.x);
    const long long blk = blockIdx.x - V.off_list[dm];
    const DevBufs b = V.dom[dm].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[dm].c;
    const long long nwords = list_words_d(c);
    if (G(b.cnt)->status != MMCA_OK)
        return;
    const long long wi0 = blk * LIST_BS;
    const long long wi = wi0 + threadIdx.x;
    unsigned long long mem = 0;
    unsigned long long nb[6] = { 0, 0, 0, 0, 0, 0 };
    if (wi < nwords)
    {
        int k, j, w;
        word_pos(c, wi, &k, &j, &w);
        const unsigned long long own = mask_word(b, c, k, j, w);
        // bit b of each word: the neighbour of interior x = 64 w + b + 1 in the
        // capture order +x, +y, -x, -y, +z, -z (growth.c:604-610)
        x_neighbour_bits(b, c, k, j, w, own, &nb[2], &nb[0]);
        nb[1] = mask_word(b, c, k, j + 1, w);
        nb[3] = mask_word(b, c, k, j - 1, w);
        nb[4] = mask_word(b, c, k + 1, j, w);
        nb[5] = mask_word(b, c, k - 1, j, w);
        const int hi = c.nx - w * 64;
        const unsigned long long inter = (hi >= 64) ? ~0ull : ((1ull << hi) - 1ull);
        mem = ~own & (nb[0] | nb[1] | nb[2] | nb[3] | nb[4] | nb[5]) & inter;
    }
    int tot;
    int ex = block_exscan_int<LIST_BS / 64>(__popcll(mem), &tot, s_w);
    s_mem[threadIdx.x] = mem;
#pragma unroll
    for (int n = 0; n < 6; ++n)
        s_nb[n][threadIdx.x] = nb[n];
    s_ex[threadIdx.x] = ex;
    __syncthreads();
    const int base = G(b.blk_count)[blk];            // members before this workgroup
    if (wi < nwords)
    {
        // slot of the first member of each row (of the next member after the
        // row's start when the row has none), for capture's row bands
        int k, j, w;
        word_pos(c, wi, &k, &j, &w);
        if (w == 0)
            G(b.row_first)[(long long) (k - 1) * c.ny + (j - 1)] = base + ex + 1;
    }
    for (int i = threadIdx.x; i < tot; i += LIST_BS)
    {
        // word of member i: the last word whose prefix count is <= i
        int lo = 0, hi = LIST_BS - 1;
        while (lo < hi)
        {
            int mid = (lo + hi + 1) >> 1;
            if (s_ex[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int bit = nth_set_bit(s_mem[lo], i - s_ex[lo]);
        int k, j, w;
        word_pos(c, wi0 + lo, &k, &j, &w);
        const int x = w * 64 + bit + 1;
        unsigned int pattern = 0;
#pragma unroll
        for (int n = 0; n < 6; ++n)
            pattern |= (unsigned int) ((s_nb[n][lo] >> bit) & 1ull) << n;
        G(b.diff_id)[base + i + 1] = (int) (k * c.sxy + (long long) j * c.sx + x);   // slots start at 1
        G(b.nbr)[base + i + 1] = (unsigned char) pattern;
    }
}

// capture_octahedra_diffuse, INDEX_SEP branch (growth.c:640-807), over the
// slots 1..gindex of the list.  Every member is independent: it reads only
// solid neighbours (through the pre-capture mask) and writes only itself and
// its slot, so the slots may be dealt to the threads in any order.  `cb` is
// the buffer holding fs/cl after fs_change; `rows` is 1 when the list writer
// of the same list has filled row_first (the row bands need it).
__global__ void __launch_bounds__(LIST_BS)
k_capture(View V, int cb, int rows)
{
    const int dm = find_dom(V.off_list, V.ndom, blockIdx.x);
    const long long blk = blockIdx.x - V.off_list[dm];
    const long long nwg = V.off_list[dm + 1] - V.off_list[dm];
    const long long nthr = nwg * LIST_BS;
    const DevBufs b = V.dom[dm].b;    // private copy: stores cannot alias it
    const Consts c = V.dom[dm].c;
    const MMCA_GLOBAL Counters *cnt = G(b.cnt);
    if (cnt->status != MMCA_OK)
        return;
    const int nmem = cnt->gindex;
    const bool realloc = cnt->realloc_old >= 0;
    const long long maxg = c.max_total_grains;
    MMCA_GLOBAL double *fsb = fs_buf(b, cb);
    MMCA_GLOBAL double *clb = cl_buf(b, cb);
    auto member = [&](long long t) {
        const int slot = (int) t + 1;
        const unsigned idx = (unsigned) G(b.diff_id)[slot];
        double tx, ty, tz;
        if (realloc)
            tx = ty = tz = 0.0;
        else
        {
            tx = G(b.dtx)[slot];
            ty = G(b.dty)[slot];
            tz = G(b.dtz)[slot];
        }
        // neighbour order +x, +y, -x, -y, +z, -z (growth.c:686-700)
        const int nnx[6] = { 1, 0, -1, 0, 0, 0 };
        const int nny[6] = { 0, 1, 0, -1, 0, 0 };
        const int nnz[6] = { 0, 0, 0, 0, 1, -1 };
        const long long ndxs[6] = { (long long) idx + 1, (long long) idx + c.sx, (long long) idx - 1,
                                    (long long) idx - c.sx, (long long) idx + c.sxy, (long long) idx - c.sxy };
        // solid neighbours before capture (recorded by the list writer); their
        // grain ids, centres and fs are independent loads issued together
        const unsigned int sol = G(b.nbr)[slot];
        int gids[6];
        double pdc[6][3], pfs[6];
#pragma unroll
        for (int n = 0; n < 6; ++n)
        {
            gids[n] = 0;
            pdc[n][0] = pdc[n][1] = pdc[n][2] = pfs[n] = 0.0;
            if ((sol >> n) & 1u)
            {
                gids[n] = G(b.gr)[ndxs[n]];
                pdc[n][0] = G(b.dc)[3 * ndxs[n]];
                pdc[n][1] = G(b.dc)[3 * ndxs[n] + 1];
                pdc[n][2] = G(b.dc)[3 * ndxs[n] + 2];
                pfs[n] = fsb[ndxs[n]];
            }
        }
        int found = -1;
        int gfound = 0;
        double g00 = 0, g10 = 0, g20 = 0, g01 = 0, g11 = 0, g21 = 0, g02 = 0, g12 = 0, g22 = 0;
        double dx = 0, dy = 0, dz = 0, dn = 0, ddx = 0, ddy = 0, ddz = 0;
#pragma unroll
        for (int n = 0; n < 6; ++n)
        {
            const int gid = gids[n];
            if (found >= 0 || !(gid > 0 && gid < maxg))
                continue;
            const MMCA_GLOBAL double *R = G(b.g_rot) + 9 * (long long) gid;
            g00 = R[0]; g01 = R[1]; g02 = R[2];
            g10 = R[3]; g11 = R[4]; g12 = R[5];
            g20 = R[6]; g21 = R[7]; g22 = R[8];
            ddx = pdc[n][0];
            ddy = pdc[n][1];
            ddz = pdc[n][2];
            // grow_octahedra (growth.c:392-403)
            double rx0 = 0.0 - ddx;
            double ry0 = 0.0 - ddy;
            double rz0 = 0.0 - ddz;
            double gx0 = g00 * rx0 + g10 * ry0 + g20 * rz0;
            double gy0 = g01 * rx0 + g11 * ry0 + g21 * rz0;
            double gz0 = g02 * rx0 + g12 * ry0 + g22 * rz0;
            double ds1 = fabs(gx0) + fabs(gy0) + fabs(gz0);
            ds1 = ref_min(ds1, 1.5);
            dn = ds1 + pfs[n] * c.rev_fsgrow;
            // growth.c:717-735
            double rx1 = -1.0 * nnx[n] - ddx;
            double ry1 = -1.0 * nny[n] - ddy;
            double rz1 = -1.0 * nnz[n] - ddz;
            dx = g00 * rx1 + g10 * ry1 + g20 * rz1;
            dy = g01 * rx1 + g11 * ry1 + g21 * rz1;
            dz = g02 * rx1 + g12 * ry1 + g22 * rz1;
            double dfs = dn - (fabs(dx) + fabs(dy) + fabs(dz));
            if (dfs > 0)
            {
                found = n;
                gfound = gid;
            }
        }
        if (found > -1)
        {
            double ncx = 0., ncz = 0., ncy = 0.;
            double ncd = ref_min(dn, 1.0);
            double dd = dn - ncd;
            if (dx + dy >= 0 && dx - dy > 0 && dx + dz > 0 && dx - dz >= 0)
                ncx = dd;
            else if (dx + dy <= 0 && dx - dy < 0 && dx + dz < 0 && dx - dz <= 0)
                ncx = -1. * dd;
            else if (dy + dx > 0 && dy - dx >= 0 && dy + dz >= 0 && dy - dz > 0)
                ncy = dd;
            else if (dy + dx < 0 && dy - dx <= 0 && dy + dz <= 0 && dy - dz < 0)
                ncy = -1. * dd;
            else if (dz + dx >= 0 && dz - dx > 0 && dz + dy > 0 && dz - dy >= 0)
                ncz = dd;
            else
                ncz = -1. * dd;
            tx = ddx + nnx[found] + g00 * ncx + g01 * ncy + g02 * ncz;
            ty = ddy + nny[found] + g10 * ncx + g11 * ncy + g12 * ncz;
            tz = ddz + nnz[found] + g20 * ncx + g21 * ncy + g22 * ncz;
            G(b.gr)[idx] = gfound;
            fsb[idx] = 0.;
            if (c.f32)
                cl32_buf(b, cb)[idx] = G(b.ce32)[idx];
            else
                clb[idx] = G(b.ce)[idx];
        }
        if (found > -1 || realloc)
        {
            G(b.dtx)[slot] = tx;
            G(b.dty)[slot] = ty;
            G(b.dtz)[slot] = tz;
        }
        G(b.dc)[3 * (long long) idx] = tx;
        G(b.dc)[3 * (long long) idx + 1] = ty;
        G(b.dc)[3 * (long long) idx + 2] = tz;
    };
    const int band = c.capture_band;
    const int nband = band > 0 ? (c.ny + band - 1) / band : 0;
    const int chunk = c.capture_chunk;
    if (rows && nband >= 8 && nwg >= 8)
    {
        // Units of `band` rows of one plane.  Band q of every plane belongs
        // to the XCD q % 8 (the hardware places workgroup w on XCD w % 8), and
        // the workgroups of this domain on one XCD take its units plane by
        // plane, one unit per workgroup at a time.  A member's solid
        // neighbours lie in its own row, in the rows next to it (mostly in
        // the same unit: one workgroup, one L1) and in the planes next to it
        // (units of the same XCD in flight together: one L2).
        const long long o = V.off_list[dm];
        const int xcd = (int) (blockIdx.x % 8);
        const long long first = o + ((xcd - (int) (o % 8)) + 8) % 8;
        const long long rank = (blockIdx.x - first) / 8;
        const long long nxcd = (o + nwg - 1 - first) / 8 + 1;
        const int mine = (nband - 1 - xcd) / 8 + 1;            // bands of this XCD in a plane
        const long long nunit = (long long) c.nz * mine;
        const long long nrow = (long long) c.nz * c.ny;
        for (long long u = rank; u < nunit; u += nxcd)
        {
            const int k = (int) (u / mine);                    // plane k + 1
            const int q = xcd + 8 * (int) (u - (long long) k * mine);
            const long long r0 = (long long) k * c.ny + (long long) q * band;
            const long long r1 = (long long) k * c.ny + min((q + 1) * band, c.ny);
            const int s0 = max(1, min(G(b.row_first)[r0], nmem + 1));
            const int s1 = r1 < nrow ? max(s0, min(G(b.row_first)[r1], nmem + 1)) : nmem + 1;
            for (int t = s0 - 1 + (int) threadIdx.x; t < s1 - 1; t += LIST_BS)
                member(t);
        }
    }
    else if (chunk > 0 && nwg >= 8)
    {
        // Contiguous chunks of `chunk` slots, chunk q on the XCD q % 8: the
        // members of one chunk (about chunk / (nx * ny) planes of the
        // interface) share most of their solid neighbours, whose lines then
        // stay in that XCD's L2.  The hardware places workgroup w on XCD
        // w % 8; the workgroups of this domain on one XCD stride over the
        // 256-slot groups of that XCD's chunks.
        const int S = chunk / LIST_BS;
        const long long o = V.off_list[dm];
        const int xcd = (int) (blockIdx.x % 8);
        const long long first = o + ((xcd - (int) (o % 8)) + 8) % 8;
        const long long rank = (blockIdx.x - first) / 8;
        const long long nxcd = (o + nwg - 1 - first) / 8 + 1;
        const long long nchunk = (nmem + chunk - 1) / chunk;
        const long long mine = (nchunk > xcd) ? (nchunk - 1 - xcd) / 8 + 1 : 0;
        for (long long u = rank; u < mine * S; u += nxcd)
        {
            const long long t = (xcd + 8 * (u / S)) * (long long) chunk + (u % S) * LIST_BS + threadIdx.x;
            if (t < nmem)
                member(t);
        }
    }
    else
        for (long long t = blk * LIST_BS + threadIdx.x; t < nmem; t += nthr)
            member(t);
}

// ---------------------------------------------------------------------------
// Sum of min(1,fs) (reduce_fs, growth.c:541-566) and of ce over the interior.
constexpr int RED_BS = 256;

__global__ void __launch_bounds__(RED_BS)
k_reduce_partial(DevBufs b, Consts c, int cb)
{
    __shared__ double s_a[RED_BS], s_b[RED_BS];
    double a = 0.0, s = 0.0;
    const long long N = (long long) c.nx * c.ny * c.nz;
    const MMCA_GLOBAL double *fs = fs_buf(b, cb);
    for (long long i = (long long) blockIdx.x * RED_BS + threadIdx.x; i < N;
         i += (long long) gridDim.x * RED_BS)
    {
        const unsigned ui = (unsigned) i, unx = (unsigned) c.nx, uny = (unsigned) c.ny;   // N < 2^31
        const unsigned row = ui / unx;
        int x = (int) (ui - row * unx) + 1;
        const unsigned kk = row / uny;
        int j = (int) (row - kk * uny) + 1;
        int k = (int) kk + 1;
        long long idx = k * c.sxy + (long long) j * c.sx + x;
        a += ref_min(1.0, fs[idx]);
        s += c.f32 ? (double) G(b.ce32)[idx] : G(b.ce)[idx];     // FP32: deviation from c0
    }
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = s;
    __syncthreads();
    for (int d = RED_BS / 2; d > 0; d >>= 1)
    {
        if (threadIdx.x < d)
        {
            s_a[threadIdx.x] += s_a[threadIdx.x + d];
            s_b[threadIdx.x] += s_b[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        G(b.red_part)[2 * blockIdx.x] = s_a[0];
        G(b.red_part)[2 * blockIdx.x + 1] = s_b[0];
    }
}

__global__ void __launch_bounds__(RED_BS)
k_reduce_final(DevBufs b, int nblk, double *out)
{
    __shared__ double s_a[RED_BS], s_b[RED_BS];
    double a = 0.0, s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += RED_BS)
    {
        a += G(b.red_part)[2 * i];
        s += G(b.red_part)[2 * i + 1];
    }
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = s;
    __syncthreads();
    for (int d = RED_BS / 2; d > 0; d >>= 1)
    {
        if (threadIdx.x < d)
        {
            s_a[threadIdx.x] += s_a[threadIdx.x + d];
            s_b[threadIdx.x] += s_b[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        out[0] = s_a[0];
        out[1] = s_b[0];
    }
}

// ---------------------------------------------------------------------------
// Separable Gaussian-beam temperature (Consts::beam_sep, FP32-deviation AM).
// Term i of the quadrature of temp_model<true> (temperature.c:104-112),
//   ldt * exp(-((lx + lv t_i)^2 + ly^2) / A_i - lz^2 / (4 alpha t_i)) / Tl_i,
// is the product X_i(x) Y_i(y) Zf_i(z) of
//   X_i = exp(-(lx + lv t_i)^2 / A_i),   Y_i = ldt / Tl_i * exp(-ly^2 / A_i),
//   Zf_i = exp(-lz^2 / (4 alpha t_i)),
// so that over the interior of a block T = liniT + Tfront * sum_i X_i(x)
// W_i(y, z) with W_i(y, z) = Y_i(y) Zf_i(z): a product of a (BEAM_NT x nx)
// and a ((nz ny) x BEAM_NT) matrix, 400 multiply-adds per cell and
// 400 (nx + ny + nz) exps per step in place of 400 exps per cell.  The three
// factors round separately and the sum runs in another order, so T differs
// from the quadrature's in its last bits.  Every factor is at most 1 (the
// exponents are not positive), so a factor that underflows stands for a term
// below the smallest double.

// beam position and amplitude at sim_time (temperature.c:62-103)
struct BeamNow {
    double xnow, ynow, znow, Tfront;
};

__device__ __forceinline__ BeamNow
beam_now(const Consts &c, double sim_time)
{
    const double invlv = 1. / c.lv;
    double xlen = c.gnsbx * c.nx * c.cell_size;
    xlen = xlen + c.lsd;
    double tline = xlen * invlv;
    const double tpause = 1.5e-4 * invlv;
    tline = tline + tpause;
    double trackno = 0;
    if (c.am_tracks > 1)
        trackno = floor(sim_time / tline);
    BeamNow n;
    n.xnow = c.x0 + (sim_time - trackno * tline) * c.lv;
    n.ynow = c.y0 + trackno * c.lhs;
    n.znow = c.z0;
    const double lpow = (sim_time < 1e-4) ? 0.0 : c.lpow;
    n.Tfront = c.lab * lpow * beam_lalpha(c) / (5.568328 * c.lthcon);
    return n;
}

// Table shapes: X rows (one per node) of beam_x_pitch columns, zero past
// nx; W rows (one per (y, z) row of the block, q = (z-1) ny + y-1) of
// BEAM_NT nodes, beam_w_rows of them, zero past ny nz.
constexpr int BF_COLS = 128;                      // columns per workgroup of k_beam_field
constexpr int BF_ROWS = 64;                       // W rows per workgroup
__host__ __device__ __forceinline__ int beam_x_pitch(const Consts &c) { return (c.nx + BF_COLS - 1) / BF_COLS * BF_COLS; }
__host__ __device__ __forceinline__ long long
beam_w_rows(const Consts &c)
{
    return ((long long) c.ny * c.nz + BF_ROWS - 1) / BF_ROWS * BF_ROWS;
}

// The factor tables at the time of timestep cnt->timestep + adv:
// X[i][x-1], Y[y-1][i], Zf[z-1][i].  W depends on the time only through
// the track's y position; beam_key[0] holds the position W was built for,
// and beam_key[1] tells k_beam_w whether to build it again.
__global__ void __launch_bounds__(256)
k_beam_tabs(DevBufs b, Consts c, int adv)
{
    const MMCA_GLOBAL Counters *cnt = G(b.cnt);
    if (cnt->status != MMCA_OK)
        return;
    const double sim_time = (cnt->timestep + (unsigned long long) adv) * c.dt;
    const BeamNow bn = beam_now(c, sim_time);
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        const double held = G(b.beam_key)[0];
        const bool same = __double_as_longlong(held) == __double_as_longlong(bn.ynow);
        G(b.beam_key)[1] = same ? 0.0 : 1.0;
        G(b.beam_key)[0] = bn.ynow;
    }
    const double ls = c.lsd / (2. * 1.414);
    const double lalpha = beam_lalpha(c);
    const double ldt = beam_ldt(c);
    const int xp = beam_x_pitch(c);
    const long long nX = (long long) BEAM_NT * xp, nY = (long long) BEAM_NT * c.ny;
    const long long n = nX + nY + (long long) BEAM_NT * c.nz;
    for (long long e = (long long) blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long) gridDim.x * blockDim.x)
    {
        if (e < nX)
        {
            const int i = (int) (e / xp), q = (int) (e - (long long) i * xp);
            const double lt = (0.5 + i) * ldt;
            const double A = ls * ls + 4. * lalpha * lt;
            double v = 0.0;                       // columns past nx pad the row
            if (q < c.nx)
            {
                const double rx = c.origin[0] + ((c.off[0] + q + 1) + 0.5) * c.cell_size;
                const double u = (rx - bn.xnow) + c.lv * lt;
                v = exp_glibc(-(u * u) / A, exp_tab);
            }
            G(b.beam_x)[e] = v;
        }
        else if (e < nX + nY)
        {
            const long long f = e - nX;
            const int y = (int) (f / BEAM_NT), i = (int) (f - (long long) y * BEAM_NT);
            const double lt = (0.5 + i) * ldt;
            const double A = ls * ls + 4. * lalpha * lt;
            const double ry = c.origin[1] + ((c.off[1] + y + 1) + 0.5) * c.cell_size;
            const double ly = ry - bn.ynow;
            const double Tl = sqrt(lalpha * lt) * (ls * ls + 4. * lalpha * lt);
            G(b.beam_y)[f] = ldt / Tl * exp_glibc(-(ly * ly) / A, exp_tab);
        }
        else
        {
            const long long f = e - nX - nY;
            const int z = (int) (f / BEAM_NT), i = (int) (f - (long long) z * BEAM_NT);
            const double lt = (0.5 + i) * ldt;
            const double rz = c.origin[2] + ((c.off[2] + z + 1) + 0.5) * c.cell_size;
            const double lz = rz - bn.znow;
            G(b.beam_z)[f] = exp_glibc(-(lz * lz) / (4. * lalpha * lt), exp_tab);
        }
    }
}

// W[q][i] = Y[y-1][i] * Zf[z-1][i] for q = (z-1) ny + y-1, when the track's
// y position changed (beam_key[1]); one row per wave at a time
__global__ void __launch_bounds__(256)
k_beam_w(DevBufs b, Consts c)
{
    if (G(b.cnt)->status != MMCA_OK || G(b.beam_key)[1] == 0.0)
        return;
    const long long rows = (long long) c.nz * c.ny;
    const long long waves = (long long) gridDim.x * (blockDim.x / 64);
    for (long long q = (long long) blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; q < rows; q += waves)
    {
        const int z = (int) (q / c.ny), y = (int) (q - (long long) z * c.ny);
        const MMCA_GLOBAL double *Y = G(b.beam_y) + (long long) y * BEAM_NT;
        const MMCA_GLOBAL double *Z = G(b.beam_z) + (long long) z * BEAM_NT;
        for (int i = threadIdx.x % 64; i < BEAM_NT; i += 64)
            G(b.beam_w)[q * BEAM_NT + i] = Y[i] * Z[i];
    }
}

// T over the interior into out (padded indexing): the product of W (rows q)
// and X on the matrix cores.  v_mfma_f64_16x16x4f64 multiplies a 16 x 4
// tile of W (lane l: row l % 16, node slot l / 16) by a 4 x 16 tile of X
// (lane l: node slot l / 16, column l % 16) into a 16 x 16 tile held as four
// doubles per lane (lane l, register r: row l / 16 + 4 r, column l % 16).
// The sum over the nodes may take them in any order, so the node slot k of
// step s in a block of 16 nodes 16 t is node 16 t + 4 k + s % 4: each lane
// then loads four consecutive W values of its row at once.  A wave owns 64
// rows x 32 columns (4 x 2 tiles, 32 accumulators per lane); the four waves
// of a workgroup share its 64 rows and take 32 columns each.
typedef double dvec4 __attribute__((ext_vector_type(4)));
static_assert(BEAM_NT % 16 == 0, "nodes come in blocks of 16");

__global__ void __launch_bounds__(256)
k_beam_field(DevBufs b, Consts c, int adv, double *out)
{
    const MMCA_GLOBAL Counters *cnt = G(b.cnt);
    if (cnt->status != MMCA_OK)
        return;                                   // uniform over the workgroup
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long q0 = (long long) blockIdx.y * BF_ROWS;
    const int x0 = blockIdx.x * BF_COLS + wave * 32;      // 0-based interior column
    const int kk = lane >> 4, lr = lane & 15;
    const double sim_time = (cnt->timestep + (unsigned long long) adv) * c.dt;
    const BeamNow bn = beam_now(c, sim_time);
    const long long nq = (long long) c.ny * c.nz;
    // stores tile (a, bb) register r: row q0 + 16 a + kk + 4 r, column x0 + 16 bb + lr
    auto store = [&](int a, int bb, int r, double sum, bool sep) {
        const long long q = q0 + 16 * a + kk + 4 * r;
        const int x = x0 + 16 * bb + lr;
        if (q >= nq || x >= c.nx)
            return;
        const int z = (int) (q / c.ny) + 1, y = (int) (q - (long long) (z - 1) * c.ny) + 1;
        const double rz = c.origin[2] + ((c.off[2] + z) + 0.5) * c.cell_size;
        double t;
        if (!(rz <= bn.znow + c.cell_size))
            t = 0.0;
        else if (!sep)
            t = c.liniT + bn.Tfront * 0.0;        // every term is finite: liniT + 0
        else
            t = c.liniT + bn.Tfront * sum;
        out[(long long) z * c.sxy + (long long) y * c.sx + x + 1] = t;
    };
    if (bn.Tfront == 0.0)
    {
        for (int a = 0; a < 4; ++a)
            for (int bb = 0; bb < 2; ++bb)
                for (int r = 0; r < 4; ++r)
                    store(a, bb, r, 0.0, false);
        return;
    }
    const int xp = beam_x_pitch(c);
    // this lane's W rows (one per row tile) and X columns (one per column tile)
    const MMCA_GLOBAL double *Wr[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        Wr[a] = G(b.beam_w) + (q0 + 16 * a + lr) * BEAM_NT + 4 * kk;
    const MMCA_GLOBAL double *Xc = G(b.beam_x) + x0 + lr + (long long) (4 * kk) * xp;
    dvec4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        for (int bb = 0; bb < 2; ++bb)
            acc[a][bb] = dvec4{0.0, 0.0, 0.0, 0.0};
    // one block of 16 nodes: loads, then 4 steps of 8 MFMAs
    dvec4 wa[4];
    double xb[2][4];
    auto load = [&](int t) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
            wa[a] = *reinterpret_cast<const MMCA_GLOBAL dvec4 *>(Wr[a] + 16 * t);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
                xb[bb][m] = Xc[(long long) (16 * t + m) * xp + 16 * bb];
    };
    load(0);
    for (int t = 0; t < BEAM_NT / 16; ++t)
    {
        const dvec4 ca[4] = { wa[0], wa[1], wa[2], wa[3] };
        const double cb[2][4] = { { xb[0][0], xb[0][1], xb[0][2], xb[0][3] },
                                  { xb[1][0], xb[1][1], xb[1][2], xb[1][3] } };
        if (t + 1 < BEAM_NT / 16)
            load(t + 1);                          // the next block's loads overlap this block's MFMAs
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
                    acc[a][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[a][m], cb[bb][m], acc[a][bb], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                store(a, bb, r, acc[a][bb][r], true);
}

// The separable temperature of timestep cnt->timestep + adv into b.temp
hipError_t launch_beam_field(const DevBufs &b, const Consts &c, int adv, hipStream_t s)
{
    const long long tabs = (long long) BEAM_NT * (beam_x_pitch(c) + c.ny + c.nz);
    hipLaunchKernelGGL(k_beam_tabs, dim3((unsigned) std::min<long long>((tabs + 255) / 256, 4096)), dim3(256), 0, s,
                       b, c, adv);
    const long long rows = (long long) c.ny * c.nz;
    hipLaunchKernelGGL(k_beam_w, dim3((unsigned) std::min<long long>((rows + 3) / 4, 2048)), dim3(256), 0, s, b, c);
    const dim3 grid((unsigned) (beam_x_pitch(c) / BF_COLS), (unsigned) (beam_w_rows(c) / BF_ROWS));
    hipLaunchKernelGGL(k_beam_field, grid, dim3(256), 0, s, b, c, adv, b.temp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Derived fields for download.  With beam_sep the interior temperature is
// the separable one, refreshed into b.temp by the caller; the halo cells
// (read by no step) keep the quadrature.
__device__ __forceinline__ double
field_temp(const DevBufs &b, const Consts &c, long long i, int x, int y, int z)
{
    if (c.beam_sep && x >= 1 && x <= c.nx && y >= 1 && y <= c.ny && z >= 1 && z <= c.nz)
        return G(b.temp)[i];
    return temp_at(c, G(b.cnt)->timestep * c.dt, x, y, z);
}

__global__ void
k_fill_temp(DevBufs b, Consts c, double *out)
{
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.P)
        return;
    int z = (int) (i / c.sxy);
    long long rem = i - z * c.sxy;
    int y = (int) (rem / c.sx);
    int x = (int) (rem - (long long) y * c.sx);
    out[i] = field_temp(b, c, i, x, y, z);
}

// cache_io_data (loop.c:74-113): FP32 copies of fs, ce and the temperature
// and a copy of gr over the padded box, for the output writers.  A null
// output is skipped.
template <bool F32>
__global__ void
k_io_fields(DevBufs b, Consts c, int cur, float *fs_io, float *ce_io, float *t_io, int *gr_io)
{
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.P)
        return;
    if (fs_io)
        fs_io[i] = (float) fs_buf(b, cur)[i];
    if (ce_io)
        ce_io[i] = (float) ld_ce<F32>(b, c, i);
    if (gr_io)
        gr_io[i] = G(b.gr)[i];
    if (t_io)
    {
        int z = (int) (i / c.sxy);
        long long rem = i - z * c.sxy;
        int y = (int) (rem / c.sx);
        int x = (int) (rem - (long long) y * c.sx);
        t_io[i] = (float) field_temp(b, c, i, x, y, z);
    }
}

hipError_t launch_io_fields(const DevBufs &b, const Consts &c, int cur, float *fs_io, float *ce_io, float *t_io,
                            int *gr_io, hipStream_t s)
{
    const dim3 grid((unsigned) ((c.P + 255) / 256));
    if (c.f32)
        hipLaunchKernelGGL(k_io_fields<true>, grid, dim3(256), 0, s, b, c, cur, fs_io, ce_io, t_io, gr_io);
    else
        hipLaunchKernelGGL(k_io_fields<false>, grid, dim3(256), 0, s, b, c, cur, fs_io, ce_io, t_io, gr_io);
    return hipGetLastError();
}

// grow_octahedra (growth.c:339-407) into the d buffer
__global__ void
k_grow_octahedra(DevBufs b, Consts c, int cb)
{
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.P || b.cnt->status != MMCA_OK)
        return;
    int z = (int) (i / c.sxy);
    long long rem = i - z * c.sxy;
    int y = (int) (rem / c.sx);
    int x = (int) (rem - (long long) y * c.sx);
    if (x < 1 || x > c.nx || y < 1 || y > c.ny || z < 1 || z > c.nz)
        return;
    int grid = G(b.gr)[i];
    // with track_d the solid set is the one of the last growth step (the
    // mask the solute kernel wrote, before capture), as in the reference
    const bool solid = c.track_d
        ? ((G(b.mask)[((long long) z * (c.ny + 2) + y) * c.w64 + (x - 1) / 64] >> ((x - 1) % 64)) & 1ull) != 0
        : grid > 0;
    if (solid && grid > 0 && grid < c.max_total_grains)
        G(b.d)[i] = octa_d(c, b, i, grid, fs_buf(b, cb)[i]);
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side).
hipError_t launch_nucleate(const View &v, int mode, int cur, hipStream_t s)
{
    if ((mode & 1) && v.tot_nuc > 0)
        hipLaunchKernelGGL(k_nuc_detect, dim3((unsigned) v.tot_nuc), dim3(NUC_BS), 0, s, v,
                           (mode & 4) ? 1 : 0, cur);
    hipLaunchKernelGGL(k_nuc_activate, dim3(v.ndom), dim3(NUC_BS), 0, s, v, mode);
    return hipGetLastError();
}

// Nucleation of a coupled batch: detection and event order per domain, then
// activation across the domains in rank order.
hipError_t launch_nucleate_coupled(const View &v, int cur, hipStream_t s)
{
    if (v.tot_nuc > 0)
        hipLaunchKernelGGL(k_nuc_detect, dim3((unsigned) v.tot_nuc), dim3(NUC_BS), 0, s, v, 1, cur);
    hipLaunchKernelGGL(k_nuc_activate, dim3(v.ndom), dim3(NUC_BS), 0, s, v, 1);
    hipLaunchKernelGGL(k_nuc_activate_coupled, dim3(1), dim3(NUC_BS), 0, s, v);
    return hipGetLastError();
}

// `max_face`: the largest face (cells) or halo-row word count of any domain
hipError_t launch_halo_exchange(const View &v, int phase, int cb, long long max_face, hipStream_t s)
{
    const unsigned gx = (unsigned) ((max_face + 255) / 256);
    hipLaunchKernelGGL(k_halo_exchange, dim3(gx, (unsigned) (6 * v.ndom)), dim3(256), 0, s, v, phase, cb);
    return hipGetLastError();
}

// zm_by: rows of the z-marching kernel (all domains of a view share it); 0
// selects the cell-parallel kernel.
template <bool AM, bool F32>
static void launch_solute_t(const View &v, int zm_by, int lds, int diff, int fsc, int rec, int cur, int nxt,
                            hipStream_t s)
{
    dim3 blk(SOL_BX, SOL_BY);
    if (rec)
    {
        // branch recording: FP64 storage, fused cell-parallel kernel (the caller checks)
        if (diff && fsc && !F32)
            hipLaunchKernelGGL((k_solute<true, true, AM, false, true>), dim3((unsigned) v.tot_cell), blk, 0, s, v,
                               cur, nxt);
        return;
    }
    if (diff && fsc && zm_by > 0 && lds)
    {
        const long long ng = v.xcd_remap ? ((v.tot_zm + 7) / 8) * 8 : v.tot_zm;
        const size_t tlds = AM ? 0 : (size_t) v.zm_planes_max * sizeof(double);   // per-plane temperatures
        if (zm_by == 16)
            hipLaunchKernelGGL((k_solute_zl<16, AM, F32>), dim3((unsigned) ng), dim3(64, 16), tlds, s, v, cur, nxt);
        else if (zm_by == 8)
            hipLaunchKernelGGL((k_solute_zl<8, AM, F32>), dim3((unsigned) ng), dim3(64, 8), tlds, s, v, cur, nxt);
        else
            hipLaunchKernelGGL((k_solute_zl<4, AM, F32>), dim3((unsigned) ng), dim3(64, 4), tlds, s, v, cur, nxt);
    }
    else if (diff && fsc && zm_by > 0)
    {
        const long long ng = v.xcd_remap ? ((v.tot_zm + 7) / 8) * 8 : v.tot_zm;
        if (zm_by == 16)
            hipLaunchKernelGGL((k_solute_zm<16, AM, F32>), dim3((unsigned) ng), dim3(64, 16), 0, s, v, cur, nxt);
        else if (zm_by == 8)
            hipLaunchKernelGGL((k_solute_zm<8, AM, F32>), dim3((unsigned) ng), dim3(64, 8), 0, s, v, cur, nxt);
        else
            hipLaunchKernelGGL((k_solute_zm<4, AM, F32>), dim3((unsigned) ng), dim3(64, 4), 0, s, v, cur, nxt);
    }
    else if (diff && fsc)
        hipLaunchKernelGGL((k_solute<true, true, AM, F32>), dim3((unsigned) v.tot_cell), blk, 0, s, v, cur, nxt);
    else if (F32)
        return;     // single operators run on FP64 storage only (the caller checks)
    else if (diff)
        hipLaunchKernelGGL((k_solute<true, false, AM, false>), dim3((unsigned) v.tot_cell), blk, 0, s, v, cur,
                           nxt);
    else
        hipLaunchKernelGGL((k_solute<false, true, AM, false>), dim3((unsigned) v.tot_cell), blk, 0, s, v, cur,
                           nxt);
}

hipError_t launch_solute(const View &v, int am, int f32, int zm_by, int lds, int diff, int fsc, int rec, int cur,
                         int nxt, hipStream_t s)
{
    if (am)
    {
        if (f32)
            launch_solute_t<true, true>(v, zm_by, lds, diff, fsc, rec, cur, nxt, s);
        else
            launch_solute_t<true, false>(v, zm_by, lds, diff, fsc, rec, cur, nxt, s);
    }
    else
    {
        if (f32)
            launch_solute_t<false, true>(v, zm_by, lds, diff, fsc, rec, cur, nxt, s);
        else
            launch_solute_t<false, false>(v, zm_by, lds, diff, fsc, rec, cur, nxt, s);
    }
    return hipGetLastError();
}

hipError_t launch_mask_from_gr(const DevBufs &b, const Consts &c, hipStream_t s)
{
    hipLaunchKernelGGL(k_mask_from_gr, dim3(c.w64, c.ny + 2, c.nz + 2), dim3(64), 0, s, b, c);
    return hipGetLastError();
}

long long list_words(const Consts &c) { return list_words_d(c); }
int list_blocks(const Consts &c) { return (int) ((list_words(c) + LIST_BS - 1) / LIST_BS); }
int nuc_blocks(int nsites) { return (nsites + NUC_BS - 1) / NUC_BS; }
long long cell_blocks(const Consts &c)
{
    return (long long) ((c.nx + SOL_BX - 1) / SOL_BX) * ((c.ny + SOL_BY - 1) / SOL_BY) *
           ((c.nz + SOL_BZ - 1) / SOL_BZ);
}
long long zm_tiles(const Consts &c, int by)
{
    return (long long) ((c.nx + 63) / 64) * ((c.ny + by - 1) / by) * ((c.nz + c.zm_planes - 1) / c.zm_planes);
}

hipError_t launch_list(const View &v, int cb, int capture, int decide_realloc, hipStream_t s)
{
    if (v.tot_list > 0)
        hipLaunchKernelGGL(k_list_count, dim3((unsigned) v.tot_list), dim3(LIST_BS), 0, s, v);
    hipLaunchKernelGGL(k_list_scan, dim3(v.ndom), dim3(SCAN_BS), 0, s, v, decide_realloc);
    if (capture >= 0 && v.tot_list > 0)
        hipLaunchKernelGGL(k_list_write, dim3((unsigned) v.tot_list), dim3(LIST_BS), 0, s, v);
    if (capture > 0 && v.tot_list > 0)
        hipLaunchKernelGGL(k_capture, dim3((unsigned) v.tot_list), dim3(LIST_BS), 0, s, v, cb, 1);
    return hipGetLastError();
}

hipError_t launch_capture_only(const View &v, const DevBufs &b, int cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_realloc_decide, dim3(1), dim3(1), 0, s, b);
    if (v.tot_list > 0)   // the list may come from a state image, which holds no row_first
        hipLaunchKernelGGL(k_capture, dim3((unsigned) v.tot_list), dim3(LIST_BS), 0, s, v, cb, 0);
    return hipGetLastError();
}

int reduce_blocks(const Consts &c)
{
    long long N = (long long) c.nx * c.ny * c.nz;
    long long nb = (N + RED_BS * 8 - 1) / (RED_BS * 8);
    if (nb > 2048)
        nb = 2048;
    if (nb < 1)
        nb = 1;
    return (int) nb;
}

hipError_t launch_reduce(const DevBufs &b, const Consts &c, int cb, double *out2, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_partial, dim3(b.nblk_red), dim3(RED_BS), 0, s, b, c, cb);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_BS), 0, s, b, b.nblk_red, out2);
    return hipGetLastError();
}

hipError_t launch_fill_temp(const DevBufs &b, const Consts &c, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_temp, dim3((unsigned) ((c.P + 255) / 256)), dim3(256), 0, s, b, c, out);
    return hipGetLastError();
}

hipError_t launch_grow_octahedra(const DevBufs &b, const Consts &c, int cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_grow_octahedra, dim3((unsigned) ((c.P + 255) / 256)), dim3(256), 0, s, b, c, cb);
    return hipGetLastError();
}

}  // namespace mmca

namespace mmca {

__global__ void
k_fill_f64(double *p, long long n, double v)
{
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

hipError_t launch_fill_f64(double *p, long long n, double v, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// Streaming probes of the bandwidth a kernel with the solute kernel's field
// mix can reach: no arithmetic beyond a dependency that keeps every load
// live.  Each workgroup streams one contiguous run of PROBE_RUN cells: every
// thread issues the loads of PROBE_UNROLL cells 256 apart, then their stores.
// Short contiguous runs on many workgroups stream faster here than a grid
// stride over the whole array (about 5.3 against 4.3-5.2 TB/s for mix 1,
// depending on the grid).
//   mix 0: copy          read a f64, write b f64                         16 B/cell
//   mix 1: FP64 fields   read ce, cl, fs f64 and gr i32; write ce, cl, fs 52 B/cell
//   mix 2: FP32 fields   read ce, cl f32, fs f64, gr i32; write ce, cl, fs 36 B/cell
constexpr int PROBE_UNROLL = 4;
constexpr int PROBE_RUN = 256 * PROBE_UNROLL;

template <int MIX>
__global__ void __launch_bounds__(256)
k_probe(long long n, double *a, double *b, double *f, float *ea, float *eb, int *g)
{
    const long long i0 = (long long) blockIdx.x * PROBE_RUN + threadIdx.x;
    double v0[PROBE_UNROLL], v1[PROBE_UNROLL], v2[PROBE_UNROLL];
    int gv[PROBE_UNROLL];
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u)
    {
        const long long i = min(i0 + u * 256, n - 1);
        if constexpr (MIX == 0)
            v0[u] = G(a)[i];
        else if constexpr (MIX == 1)
        {
            v0[u] = G(a)[i];
            v1[u] = G(b)[i];
            v2[u] = G(f)[i];
            gv[u] = G(g)[i];
        }
        else
        {
            v0[u] = G(ea)[i];
            v1[u] = G(eb)[i];
            v2[u] = G(f)[i];
            gv[u] = G(g)[i];
        }
    }
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u)
    {
        const long long i = i0 + u * 256;
        if (i >= n)
            break;
        if constexpr (MIX == 0)
            G(b)[i] = v0[u];
        else if constexpr (MIX == 1)
        {
            G(a)[i] = v0[u] + gv[u];
            G(b)[i] = v1[u];
            G(f)[i] = v2[u];
        }
        else
        {
            G(ea)[i] = (float) v0[u] + gv[u];
            G(eb)[i] = (float) v1[u];
            G(f)[i] = v2[u];
        }
    }
}

hipError_t launch_probe(int mix, long long n, double *a, double *b, double *f, float *ea, float *eb, int *g,
                        hipStream_t s)
{
    const unsigned grid = (unsigned) ((n + PROBE_RUN - 1) / PROBE_RUN);
    if (mix == 0)
        hipLaunchKernelGGL(k_probe<0>, dim3(grid), dim3(256), 0, s, n, a, b, f, ea, eb, g);
    else if (mix == 1)
        hipLaunchKernelGGL(k_probe<1>, dim3(grid), dim3(256), 0, s, n, a, b, f, ea, eb, g);
    else
        hipLaunchKernelGGL(k_probe<2>, dim3(grid), dim3(256), 0, s, n, a, b, f, ea, eb, g);
    return hipGetLastError();
}

// y[i] = exp(x[i]) through the beam quadrature's exp (mmca_exp.h), with the
// table read from constant memory as temp_model<true> reads it
__global__ void
k_eval_exp(const double *x, double *y, long long n)
{
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        y[i] = exp_glibc(x[i], exp_tab);
}

hipError_t launch_eval_exp(const double *x, double *y, long long n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_eval_exp, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

}  // namespace mmca
