// Internal definitions shared by the HIP kernels and the host side of the
// CA step library.  Not part of the C ABI (include/mmca.h is).
#pragma once

#include <stdint.h>
#include <vector>

#include "../../include/mmca.h"

namespace mmca {

// Grid geometry and per-run constants, passed to every kernel by value.
// The folded constants are computed on the host with the same operations
// and order as the reference (growth.c:38-40, 181-189; read_ctrl.c:551).
struct Consts {
    int nx, ny, nz;         // interior cells
    int sx;                 // nx + 2 (row pitch in cells)
    long long sxy;          // (nx+2)*(ny+2) (plane pitch in cells)
    long long P;            // padded cells
    int w64;                // 64-bit mask words per padded row
    int pad_lo[3], pad_hi[3];   // mold wall on the low / high face of each axis: FaceCtrl pad on a
                                // face of the global domain (blank_moldface, temperature.c:133-152)
    int off[3];             // subblock offset in global cells (sb_coords * n)
    int nsub[3];            // subblocks per axis of a coupled batch (0: not coupled)
    int gdimz;              // global z extent (2-D runs use y as the axis)
    int zm_planes;          // > 0: z-marching solute kernel with this many planes per workgroup
    int zm_by;              // rows per z-marching workgroup (4, 8, 16)
    int xcd_remap;          // 1: contiguous tile runs per XCD
    int zm_lds;             // 1: z-marching kernel with its planes staged in LDS (k_solute_zl)
    int capture_chunk;      // > 0: capture deals list slots in chunks of this many, pinned to XCDs
    int capture_band;       // > 0: capture works on units of this many rows of one plane, band q on XCD q % 8
    int track_d;            // 1: keep the reference's d array (stale values of cells no longer solid)
    int f32;                // 1: ce, cl stored as FP32 deviations from cinit
    // diffusion (growth.c:38-40)
    double rs, rl;
    // partitioning (growth.c:181-189)
    double liq_slope, k_inv, km, km_inv, m_inv, D_xt, cinit, Tliq, Tsol;
    // nucleation undercooling m(T) = a T^2 + b T + c (grain.c:1265-1267)
    double ma, mb, mc;
    double rev_fsgrow;
    long long max_total_grains;
    int check_neg;
    // temperature (temperature.c:24-126)
    int am;                 // 1: Gaussian-beam model
    int beam_sep;           // 1 (FP32-deviation AM only): the beam temperature from its separable
                            // factors, computed per step into DevBufs::temp
    int am_layers, am_tracks, gnsbx;
    double dt, cell_size, origin[3];
    double t_init, cool_rate, gradient, grad_coef, velocity, velo_coef;
    double lpow, lv, lsd, lab, lhs, liniT, lthcon, rho, lcp, x0, y0, z0;
};

// quadrature nodes of the Gaussian-beam temperature (temperature.c:104)
constexpr int BEAM_NT = 400;

// Device-resident counters of one subblock.
struct Counters {
    unsigned long long timestep;     // bp->timestep
    unsigned long long failed_step;  // timestep of the first error
    int num_grains;                  // bp->num_grains (includes grain 0)
    int random_used;                 // random orientations consumed
    int gindex;                      // interface-list length
    int dc_tmp_size;                 // emulated size of the reference's dc_tmp
    int realloc_old;                 // slots zeroed this step (old size), -1 if none
    int status;                      // MMCA_OK or first error
    int n_events;                    // nucleation events of the last detection
    int new_grains;                  // grains created since the host last reset it
    int pad_[2];
};

// Device buffers of one subblock.
struct DevBufs {
    double *ce;
    double *cl[2];
    float *ce32;                     // FP32 storage: ce - cinit
    float *cl32[2];                  // FP32 storage: cl - cinit
    double *ledger;                  // 2 per workgroup slot: Kahan sum of FP32 storage rounding of ce
    long long nledger;
    double *fs[2];
    int *gr;
    double *dc;                      // decentered octahedron centres, x,y,z per cell (decentered_t)
    unsigned long long *mask;        // 1 bit per padded cell: gr > 0
    unsigned char *xhalo;            // per padded row: gr > 0 of its x = 0 and x = nx+1 halo cells (2 bytes)
    // interface list and dc_tmp emulation
    int *diff_id;                    // slots 1..gindex
    unsigned char *nbr;              // per slot: solid neighbours before capture, bit n = capture direction n
    double *dtx, *dty, *dtz;         // dc_tmp slots 0..capacity
    int *blk_count;                  // per-block member counts -> offsets
    int *row_first;                  // per interior row (k-1) ny + (j-1): slot of its first list member
    int nblk_list;
    // nucleation sites, sorted by padded index
    int *site_idx;
    float *site_thr;
    int *site_fixed;                 // index into the fixed orientation table, -1 if random
    int *site_ls;                    // AM lsindex of the site cell
    int *events;                     // site numbers that nucleate this step, in order
    int *ev_blk;                     // per-workgroup event lists (NUC_BS per workgroup)
    int *ev_cnt;                     // per-workgroup event counts
    int nsites;
    // grains
    double *g_rot;                   // 9 per grain, rotmat row-major
    int *g_cell;                     // padded index of the nucleation cell
    unsigned long long *g_step;      // nucleation timestep
    int *g_src;                      // >=0: fixed table index; <0: -(k+1) random draw k
    int *g_dom;                      // coupled batches: domain (rank) that created the grain
    double *fixed_rot;               // 9 per fixed nucleation point
    double *rand_rot;                // 9 per random orientation draw
    int rand_cap;                    // random orientations uploaded
    // separable beam temperature (Consts::beam_sep): T per padded cell, the
    // factor tables X[i][x], Y[y][i], Zf[z][i], W[(z-1) ny + y-1][i] =
    // Y[y][i] Zf[z][i], and the track position W was built for (2 doubles)
    double *temp;
    double *beam_x, *beam_y, *beam_z, *beam_w, *beam_key;
    // single-operator buffers (allocated on first use)
    double *oce;
    double *d;
    // reductions
    double *red_part;                // 2 * nblk_red partial sums
    int nblk_red;
    Counters *cnt;
    // partitioning-branch record (caller's device memory; null when off):
    // the step with timestep t goes to rec + (t - rec_t0 - 1) * rec_words(c) * 8
    // bytes (3 bits per cell, layout of rec_code)
    unsigned char *rec;
    unsigned long long rec_t0;
    long long rec_steps;
};

// One domain (subblock) as the hot kernels see it.
struct DomainDesc {
    DevBufs b;
    Consts c;
};

// The domains advanced by one launch.  Workgroup w of a kernel belongs to
// domain d when off[d] <= w < off[d+1] (prefix sums of each domain's
// workgroup count for that kernel).
struct View {
    const DomainDesc *dom;       // device array, ndom entries
    int ndom;
    int xcd_remap;               // contiguous tile runs per XCD for the solute kernel
    int zm_planes_max;           // largest zm_planes of the domains (LDS of the per-plane temperatures)
    const long long *off_nuc;    // k_nuc_detect workgroups
    const long long *off_zm;     // z-marching solute tiles
    const long long *off_cell;   // cell-parallel solute workgroups
    const long long *off_list;   // list count / capture workgroups
    long long tot_nuc, tot_zm, tot_cell, tot_list;
};

}  // namespace mmca
