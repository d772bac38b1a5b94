// Host side of the CA step library: the C ABI of include/mmca.h, the
// nucleation set-up on the host (the reference's RNG stream), and the
// ordering of kernel launches on one HIP stream per context.

#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mmca_internal.h"

namespace mmca {
hipError_t launch_nucleate(const DevBufs &b, const Consts &c, int mode, int cur, hipStream_t s);
hipError_t launch_solute(const DevBufs &b, const Consts &c, int diff, int fsc, int cur, int nxt,
                         hipStream_t s);
hipError_t launch_mask_from_gr(const DevBufs &b, const Consts &c, hipStream_t s);
hipError_t launch_list(const DevBufs &b, const Consts &c, int cb, int capture, int decide_realloc,
                       hipStream_t s);
hipError_t launch_capture_only(const DevBufs &b, const Consts &c, int cb, hipStream_t s);
hipError_t launch_reduce(const DevBufs &b, const Consts &c, int cb, double *out2, hipStream_t s);
hipError_t launch_fill_temp(const DevBufs &b, const Consts &c, double *out, hipStream_t s);
hipError_t launch_grow_octahedra(const DevBufs &b, const Consts &c, int cb, hipStream_t s);
hipError_t launch_fill_f64(double *p, long long n, double v, hipStream_t s);
long long list_words(const Consts &c);
int list_blocks(const Consts &c);
int reduce_blocks(const Consts &c);
}  // namespace mmca

using namespace mmca;

#define HIPCHK(expr)                                                               \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "mmca: %s failed: %s (%s:%d)\n", #expr,                \
                    hipGetErrorString(e_), __FILE__, __LINE__);                    \
            return MMCA_ERR_HIP;                                                   \
        }                                                                          \
    } while (0)

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif

namespace {

// glibc rand(): TYPE_3 additive feedback generator x[i] = x[i-3] + x[i-31]
// (published algorithm of glibc srandom_r/random_r), seeded as the reference
// does with srand(iproc + 1 + seed) (ca_main.c:215).
struct GlibcRand {
    int32_t r[31];
    int f = 3, b = 0;
    bool gauss_have = false;   // functions.c:380
    float gauss_next = 0.0f;   // functions.c:381

    void seed(unsigned int s)
    {
        if (s == 0)
            s = 1;
        int32_t word = (int32_t) s;
        r[0] = word;
        for (int i = 1; i < 31; i++) {
            long hi = word / 127773;
            long lo = word % 127773;
            word = (int32_t) (16807 * lo - 2836 * hi);
            if (word < 0)
                word += 2147483647;
            r[i] = word;
        }
        f = 3;
        b = 0;
        for (int i = 0; i < 310; i++)
            next();
        gauss_have = false;
        gauss_next = 0.0f;
    }
    int32_t next()
    {
        uint32_t val = (uint32_t) r[f] + (uint32_t) r[b];
        r[f] = (int32_t) val;
        int32_t out = (int32_t) (val >> 1);
        if (++f >= 31) {
            f = 0;
            ++b;
        } else if (++b >= 31)
            b = 0;
        return out;
    }
    // getRandScale (functions.c:441-445)
    float scale() { return (float) next() / (float) 2147483647; }
    // getRandomGaussian (functions.c:377-405)
    float gaussian()
    {
        if (gauss_have) {
            gauss_have = false;
            return gauss_next;
        }
        float n0, n1, s;
        do {
            n0 = ((float) next() / (float) (2147483647 / 2.0)) - 1.0;
            n1 = ((float) next() / (float) (2147483647 / 2.0)) - 1.0;
            s = n0 * n0 + n1 * n1;
        } while (s >= 1.0 || s == 0);
        float multiplier = sqrtf(-2.0 * log(s) / s);
        float val = n0 * multiplier;
        gauss_next = n1 * multiplier;
        gauss_have = true;
        return val;
    }
};

// nucleation_t (globals.h:147-160)
struct NucPt {
    double x, y, z;
    size_t cx, cy, cz;
    double gx, gy, gz;
    double threshold;
};

// rotation matrix of nucleate_grain (grain.c:301-344), row-major.  The
// reference, built with gcc, evaluates each cos(a)/sin(a) pair with one
// glibc sincos() call (gcc merges them); sincos can differ from separate
// sin/cos in the last bit, so it is called here explicitly.
void rotmat_of(double a0, double a1, double a2, double *R)
{
    double c0, s0, c1, s1, c2, s2;
    sincos(a0, &s0, &c0);
    sincos(a1, &s1, &c1);
    sincos(a2, &s2, &c2);
    R[0 * 3 + 0] = c2 * c0 - s2 * s1 * s0;
    R[1 * 3 + 0] = c2 * s0 + s2 * c1 * c0;
    R[2 * 3 + 0] = s2 * s1;
    R[0 * 3 + 1] = -1. * s2 * c0 - c2 * c1 * s0;
    R[1 * 3 + 1] = -1. * s2 * s0 + c2 * c1 * c0;
    R[2 * 3 + 1] = c2 * s1;
    R[0 * 3 + 2] = s1 * s0;
    R[1 * 3 + 2] = -1. * s1 * c0;
    R[2 * 3 + 2] = c1;
}

}  // namespace

struct mmca_ctx {
    mmca_params p;
    Consts c;
    DevBufs b;
    hipStream_t stream = nullptr;
    int device = 0;
    int cur = 0;                 // buffer index of the latest cl/fs
    bool have_oce = false;       // single-operator diffusion ran
    // host mirror of the counters after the last synchronisation
    Counters hc;
    // nucleation set-up
    GlibcRand rng;
    int oriented = 0;
    bool nucleation_ready = false;
    std::vector<NucPt> nuc_pt;
    std::vector<int> site_idx;
    std::vector<float> site_thr;
    std::vector<int> site_fixed;
    std::vector<double> rand_angles;   // 3 per random orientation draw
    int rand_filled = 0;
    int rand_capacity = 0;
    double *d_red = nullptr;           // 2 doubles
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double time_ms = 0.0;
    long long time_n = 0;
};

namespace {

long long padded_cells(const mmca_params &p)
{
    return (long long) (p.nx + 2) * (p.ny + 2) * (p.nz + 2);
}

int build_consts(const mmca_params &p, Consts &c)
{
    if (p.nx < 1 || p.ny < 1 || p.nz < 1)
        return MMCA_ERR_ARG;
    if (padded_cells(p) >= (1LL << 31) - 1)
        return MMCA_ERR_ARG;
    if (p.max_total_grains < 1 || p.max_total_grains > (1LL << 30))
        return MMCA_ERR_ARG;
    if (p.fsgrow == 0.0 || p.part_coef == 0.0 || p.part_coef == 1.0 || p.m_solute[2] == 0.0)
        return MMCA_ERR_ARG;
    memset(&c, 0, sizeof(c));
    c.nx = p.nx;
    c.ny = p.ny;
    c.nz = p.nz;
    c.sx = p.nx + 2;
    c.sxy = (long long) (p.nx + 2) * (p.ny + 2);
    c.P = padded_cells(p);
    c.w64 = (c.sx + 63) / 64;
    for (int a = 0; a < 3; a++) {
        c.pad[a] = p.pad_bndy[a] ? 1 : 0;
    }
    c.off[0] = p.sb_coords[0] * p.nx;
    c.off[1] = p.sb_coords[1] * p.ny;
    c.off[2] = p.sb_coords[2] * p.nz;
    c.gdimz = p.gdim[2];
    // growth.c:38-40
    double dtx = p.dt / (p.cell_size * p.cell_size);
    c.rs = (p.d_sol * dtx);
    c.rl = (p.d_liq * dtx);
    // growth.c:181-189
    c.liq_slope = p.m_solute[2];
    c.k_inv = 1.0 / p.part_coef;
    c.km = 1.0 - p.part_coef;
    c.km_inv = 1.0 / c.km;
    c.m_inv = 1.0 / c.liq_slope;
    double D_x = p.d_liq / p.cell_size / p.cell_size;
    c.D_xt = D_x * p.dt;
    c.cinit = p.c_init;
    c.Tliq = p.liquidus_temp;
    c.Tsol = p.solidus_temp;
    c.ma = p.m_solute[0];
    c.mb = p.m_solute[1];
    c.mc = p.m_solute[2];
    c.rev_fsgrow = 1.0 / p.fsgrow;   // read_ctrl.c:551
    c.max_total_grains = p.max_total_grains;
    c.check_neg = p.check_negative_ce ? 1 : 0;
    c.am = (p.temp_model == MMCA_TEMP_AM_BEAM) ? 1 : 0;
    c.am_layers = p.am_layers;
    c.am_tracks = p.am_tracks;
    c.gnsbx = p.gnsbx;
    c.dt = p.dt;
    c.cell_size = p.cell_size;
    for (int a = 0; a < 3; a++)
        c.origin[a] = p.origin[a];
    c.t_init = p.t_init;
    c.cool_rate = p.cool_rate;
    c.gradient = p.gradient;
    c.grad_coef = p.grad_coef;
    c.velocity = p.velocity;
    c.velo_coef = p.velo_coef;
    c.lpow = p.laser_power;
    c.lv = p.laser_speed;
    c.lsd = p.laser_spot;
    c.lab = p.laser_absorp;
    c.lhs = p.laser_hatch;
    c.liniT = p.sub_init_temp;
    c.lthcon = p.th_conductivity;
    c.rho = p.density;
    c.lcp = p.specific_heat;
    c.x0 = p.laser_start[0];
    c.y0 = p.laser_start[1];
    c.z0 = p.laser_start[2];
    return MMCA_OK;
}

// direct_temp_calc on the host, for the initial lsindex of AM sites
// (tempPrepareSB_int, temperature.c:156-192)
double host_temp(const Consts &c, double sim_time, int x, int y, int z)
{
    if (!c.am && c.gradient == 0)
        return c.t_init - sim_time * c.cool_rate;
    double rx = c.origin[0] + ((c.off[0] + x) + 0.5) * c.cell_size;
    double ry = c.origin[1] + ((c.off[1] + y) + 0.5) * c.cell_size;
    double rz = c.origin[2] + ((c.off[2] + z) + 0.5) * c.cell_size;
    if (!c.am) {
        double axis = ((c.gdimz == 1) ? ry : rz) - c.cell_size * 0.5;
        double tgrad = c.gradient + c.grad_coef * sim_time;
        double v = c.velocity + c.velo_coef * sim_time;
        double vt = 0.5 * (v + c.velocity) * sim_time;
        return c.t_init + tgrad * (axis - vt);
    }
    double lv = c.lv, invlv = 1. / lv, ls = c.lsd / (2. * 1.414);
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
    double lx = rx - xnow, ly = ry - ynow, lz = rz - znow;
    if (!(rz <= znow + c.cell_size))
        return 0.0;
    double Tinteg = 0.0, pi15 = 5.568328;
    double lalpha = c.lthcon / (c.rho * c.lcp);
    double tsteady = 20 * lalpha * invlv * invlv;
    double ldt = tsteady / 400;
    if (sim_time < 1e-4)
        lpow = 0;
    for (int i = 0; i < 400; i++) {
        double lt = (0.5 + i) * ldt;
        double Tlocal = sqrt(lalpha * lt) * (ls * ls + 4. * lalpha * lt);
        Tinteg += ldt *
            exp(-((lx + lv * lt) * (lx + lv * lt) + ly * ly) / (ls * ls + 4. * lalpha * lt) -
                (lz * lz / (4. * lalpha * lt))) / Tlocal;
    }
    double Tfront = c.lab * lpow * lalpha / (pi15 * c.lthcon);
    return c.liniT + Tfront * Tinteg;
}

template <typename T>
int dev_alloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0)
        n = 1;
    HIPCHK(hipMalloc((void **) p, n * sizeof(T)));
    return MMCA_OK;
}

#define TRY(x)                 \
    do {                       \
        int st_ = (x);         \
        if (st_ != MMCA_OK)    \
            return st_;        \
    } while (0)

int sync_counters(mmca_ctx *ctx)
{
    HIPCHK(hipMemcpyAsync(&ctx->hc, ctx->b.cnt, sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MMCA_OK;
}

int push_counters(mmca_ctx *ctx)
{
    HIPCHK(hipMemcpyAsync(ctx->b.cnt, &ctx->hc, sizeof(Counters), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MMCA_OK;
}

int courant_check(const mmca_ctx *ctx)
{
    // growth.c:43-50
    if (ctx->c.rs > ctx->p.courant_limit || ctx->c.rl > ctx->p.courant_limit)
        return MMCA_ERR_COURANT;
    return MMCA_OK;
}

// Draw random orientations up to `upto` (3 getRandScale each, grain.c:118-129)
// and upload their rotation matrices.
int ensure_random(mmca_ctx *ctx, long long upto)
{
    if (upto > ctx->rand_capacity)
        upto = ctx->rand_capacity;
    if (upto <= ctx->rand_filled)
        return MMCA_OK;
    int from = ctx->rand_filled;
    int n = (int) (upto - from);
    std::vector<double> rot((size_t) n * 9);
    ctx->rand_angles.resize((size_t) upto * 3);
    for (int k = 0; k < n; k++) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (ctx->oriented) {
            float acon = 4.0;
            gx = M_PI / acon - 2 * (M_PI / acon * ctx->rng.scale());
            gy = M_PI / acon - 2 * (M_PI / acon * ctx->rng.scale());
            gz = M_PI / acon - 2 * (M_PI / acon * ctx->rng.scale());
        }
        double *a = &ctx->rand_angles[(size_t) (from + k) * 3];
        a[0] = gx;
        a[1] = gy;
        a[2] = gz;
        rotmat_of(gx, gy, gz, &rot[(size_t) k * 9]);
    }
    HIPCHK(hipMemcpyAsync(ctx->b.rand_rot + (size_t) from * 9, rot.data(), rot.size() * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->rand_filled = (int) upto;
    return MMCA_OK;
}

int ensure_oce(mmca_ctx *ctx)
{
    if (!ctx->b.oce)
        TRY(dev_alloc(&ctx->b.oce, (size_t) ctx->c.P));
    return MMCA_OK;
}

int ensure_d(mmca_ctx *ctx)
{
    if (!ctx->b.d) {
        TRY(dev_alloc(&ctx->b.d, (size_t) ctx->c.P));
        HIPCHK(hipMemsetAsync(ctx->b.d, 0, (size_t) ctx->c.P * sizeof(double), ctx->stream));
    }
    return MMCA_OK;
}

int check_status(mmca_ctx *ctx)
{
    TRY(sync_counters(ctx));
    return ctx->hc.status;
}

// one doiteration() on the device (calculate.c:193-352)
int enqueue_iteration(mmca_ctx *ctx, bool advance)
{
    const int nxt = ctx->cur ^ 1;
    HIPCHK(launch_nucleate(ctx->b, ctx->c, advance ? 7 : 3, ctx->cur, ctx->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, ctx->stream));
    }
    HIPCHK(launch_solute(ctx->b, ctx->c, 1, 1, ctx->cur, nxt, ctx->stream));
    if (ctx->timing) {
        HIPCHK(hipEventRecord(e1, ctx->stream));
        ctx->ev.push_back(e0);
        ctx->ev.push_back(e1);
    }
    HIPCHK(launch_list(ctx->b, ctx->c, nxt, 1, 1, ctx->stream));
    ctx->cur = nxt;
    return MMCA_OK;
}

int collect_timing(mmca_ctx *ctx)
{
    double tot = 0.0;
    long long n = 0;
    for (size_t i = 0; i + 1 < ctx->ev.size(); i += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
        tot += ms;
        n++;
        (void) hipEventDestroy(ctx->ev[i]);
        (void) hipEventDestroy(ctx->ev[i + 1]);
    }
    ctx->ev.clear();
    ctx->time_ms = n ? tot / n : 0.0;
    ctx->time_n = n;
    return MMCA_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

int mmca_abi_version(void) { return MMCA_ABI_VERSION; }

const char *mmca_status_string(int status)
{
    switch (status) {
    case MMCA_OK: return "ok";
    case MMCA_ERR_ARG: return "bad argument";
    case MMCA_ERR_HIP: return "HIP runtime error";
    case MMCA_ERR_COURANT: return "Instability by Courant criterion";
    case MMCA_ERR_NEGATIVE_CE: return "negative ce (Courant instability)";
    case MMCA_ERR_GRAIN_CAPACITY: return "grain capacity exceeded";
    case MMCA_ERR_NUC_THRESHOLD: return "nucleation threshold draw does not terminate";
    case MMCA_ERR_NO_DEVICE: return "no GPU device";
    case MMCA_ERR_STATE: return "call out of order";
    default: return "unknown status";
    }
}

int mmca_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        n = 0;
    if (count)
        *count = n;
    return MMCA_OK;
}

int mmca_create(const mmca_params *params, int device, mmca_ctx **out)
{
    if (!params || !out)
        return MMCA_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return MMCA_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev)
        return MMCA_ERR_ARG;
    mmca_ctx *ctx = new mmca_ctx();
    ctx->p = *params;
    ctx->device = device;
    int st = build_consts(ctx->p, ctx->c);
    if (st != MMCA_OK) {
        delete ctx;
        return st;
    }
    memset(&ctx->b, 0, sizeof(ctx->b));
    *out = ctx;     // destroy() frees partial allocations on failure
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreate(&ctx->stream));
    const Consts &c = ctx->c;
    DevBufs &b = ctx->b;
    const size_t P = (size_t) c.P;
    const size_t N = (size_t) c.nx * c.ny * c.nz;
    TRY(dev_alloc(&b.ce, P));
    TRY(dev_alloc(&b.cl[0], P));
    TRY(dev_alloc(&b.cl[1], P));
    TRY(dev_alloc(&b.fs[0], P));
    TRY(dev_alloc(&b.fs[1], P));
    TRY(dev_alloc(&b.gr, P));
    TRY(dev_alloc(&b.dcx, P));
    TRY(dev_alloc(&b.dcy, P));
    TRY(dev_alloc(&b.dcz, P));
    const size_t mask_words = (size_t) (c.ny + 2) * (c.nz + 2) * c.w64;
    TRY(dev_alloc(&b.mask, mask_words));
    TRY(dev_alloc(&b.diff_id, N + 2));
    // dc_tmp slots 0 .. N + 101 (dc_tmp_size <= N + 100, slot dc_tmp_size used)
    TRY(dev_alloc(&b.dtx, N + 102));
    TRY(dev_alloc(&b.dty, N + 102));
    TRY(dev_alloc(&b.dtz, N + 102));
    b.nblk_list = list_blocks(c);
    TRY(dev_alloc(&b.blk_count, (size_t) b.nblk_list));
    b.nblk_red = reduce_blocks(c);
    TRY(dev_alloc(&b.red_part, (size_t) 2 * b.nblk_red));
    TRY(dev_alloc(&ctx->d_red, 2));
    const size_t G = (size_t) ctx->p.max_total_grains + 1;
    TRY(dev_alloc(&b.g_rot, G * 9));
    TRY(dev_alloc(&b.g_cell, G));
    TRY(dev_alloc(&b.g_step, G));
    TRY(dev_alloc(&b.g_src, G));
    ctx->rand_capacity = (int) ctx->p.max_total_grains;
    TRY(dev_alloc(&b.rand_rot, (size_t) ctx->rand_capacity * 9));
    TRY(dev_alloc(&b.cnt, 1));
    b.nsites = 0;
    // SetupInitialConditions (setup.c:58-65): ce = cl = Cinit, all else 0
    HIPCHK(launch_fill_f64(b.ce, (long long) P, c.cinit, ctx->stream));
    HIPCHK(launch_fill_f64(b.cl[0], (long long) P, c.cinit, ctx->stream));
    HIPCHK(launch_fill_f64(b.cl[1], (long long) P, c.cinit, ctx->stream));
    HIPCHK(hipMemsetAsync(b.fs[0], 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.fs[1], 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.gr, 0, P * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dcx, 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dcy, 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dcz, 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.mask, 0, mask_words * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemsetAsync(b.diff_id, 0, (N + 2) * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dtx, 0, (N + 102) * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dty, 0, (N + 102) * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dtz, 0, (N + 102) * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_rot, 0, G * 9 * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_cell, 0, G * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_step, 0, G * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_src, 0, G * sizeof(int), ctx->stream));
    memset(&ctx->hc, 0, sizeof(ctx->hc));
    ctx->hc.timestep = ctx->p.timestep;
    ctx->hc.num_grains = 1;      // read_ctrl.c:52: the "zero" grain
    ctx->hc.realloc_old = -1;
    TRY(push_counters(ctx));
    return MMCA_OK;
}

int mmca_destroy(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_OK;
    (void) hipSetDevice(ctx->device);
    if (ctx->stream)
        (void) hipStreamSynchronize(ctx->stream);
    DevBufs &b = ctx->b;
    void *ptrs[] = { b.ce, b.cl[0], b.cl[1], b.fs[0], b.fs[1], b.gr, b.dcx, b.dcy, b.dcz, b.mask,
                     b.diff_id, b.dtx, b.dty, b.dtz, b.blk_count, b.site_idx, b.site_thr, b.site_fixed,
                     b.site_ls, b.events, b.g_rot, b.g_cell, b.g_step, b.g_src, b.fixed_rot, b.rand_rot,
                     b.oce, b.d, b.red_part, b.cnt, ctx->d_red };
    for (void *p : ptrs)
        if (p)
            (void) hipFree(p);
    for (auto e : ctx->ev)
        (void) hipEventDestroy(e);
    if (ctx->stream)
        (void) hipStreamDestroy(ctx->stream);
    delete ctx;
    return MMCA_OK;
}

// init_nucleation_sites (grain.c:603-713) + init_sb_nucleation (grain.c:931-1010)
int mmca_init_nucleation(mmca_ctx *ctx, const mmca_nucleation_cfg *cfg)
{
    if (!ctx || !cfg)
        return MMCA_ERR_ARG;
    if (ctx->nucleation_ready)
        return MMCA_ERR_STATE;
    if ((cfg->n_fixed > 0 && !cfg->fixed) || (cfg->n_volumes > 0 && !cfg->volumes))
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const mmca_params &p = ctx->p;
    const Consts &c = ctx->c;
    GlibcRand &rng = ctx->rng;
    rng.seed((unsigned int) (cfg->rank + 1 + cfg->random_seed));
    ctx->oriented = cfg->oriented;
    auto real2cell = [&](double x, double y, double z, size_t *i, size_t *j, size_t *k) {
        // realcoord2scoord (functions.c:92-106)
        x = x - p.origin[0] + p.cell_size * 0.1;
        y = y - p.origin[1] + p.cell_size * 0.1;
        z = z - p.origin[2] + p.cell_size * 0.1;
        *i = (size_t) (x / p.cell_size);
        *j = (size_t) (y / p.cell_size);
        *k = (size_t) (z / p.cell_size);
    };
    auto cell2real = [&](int axis, long i) { return p.origin[axis] + (i + 0.5) * p.cell_size; };
    auto set_orientation = [&](NucPt &n) {   // setNucOrientation (grain.c:118-129)
        n.gx = n.gy = n.gz = 0.0;
        if (ctx->oriented) {
            float acon = 4.0;
            n.gx = M_PI / acon - 2 * (M_PI / acon * rng.scale());
            n.gy = M_PI / acon - 2 * (M_PI / acon * rng.scale());
            n.gz = M_PI / acon - 2 * (M_PI / acon * rng.scale());
        }
    };
    // Knuth Algorithm S (distribute_nucleations, grain.c:251-296)
    auto distribute = [&](size_t nc, size_t num_to_select, auto &&select) -> int {
        size_t n = std::min(num_to_select, nc);
        size_t N = nc, t = 0, m = 0, idx = 0;
        if (n == 0)
            return MMCA_OK;
        do {
            double U = rng.scale();
            if ((N - t) * U >= (n - m)) {
                t++;
            } else {
                int st = select(idx);
                if (st != MMCA_OK)
                    return st;
                t++;
                m++;
            }
            idx++;
        } while (m < n);
        return MMCA_OK;
    };

    // g_nuc_pt: fixed nucleation points in linked-list order (ll_insert puts
    // each new entry at the head, ll.c:59-84), then nucleation volumes
    std::vector<NucPt> g;
    for (int i = cfg->n_fixed - 1; i >= 0; i--) {
        const mmca_fixed_nuc &f = cfg->fixed[i];
        NucPt n;
        n.x = f.x;
        n.y = f.y;
        n.z = f.z;
        n.gx = (M_PI / 180.0) * f.ax;    // grain.c:532-535
        n.gy = (M_PI / 180.0) * f.ay;
        n.gz = (M_PI / 180.0) * f.az;
        n.threshold = f.threshold;
        real2cell(n.x, n.y, n.z, &n.cx, &n.cy, &n.cz);
        g.push_back(n);
    }
    for (int vi = cfg->n_volumes - 1; vi >= 0; vi--) {
        mmca_nuc_volume v = cfg->volumes[vi];
        // grain.c:651-693
        v.minx = std::max(v.minx, p.origin[0]);
        v.maxx = std::min(v.maxx, p.origin[0] + (p.gdim[0] * p.cell_size));
        v.miny = std::max(v.miny, p.origin[1]);
        v.maxy = std::min(v.maxy, p.origin[1] + (p.gdim[1] * p.cell_size));
        v.minz = std::max(v.minz, p.origin[2]);
        v.maxz = std::min(v.maxz, p.origin[2] + (p.gdim[2] * p.cell_size));
        size_t mnx, mny, mnz, mxx, mxy, mxz;
        real2cell(v.minx, v.miny, v.minz, &mnx, &mny, &mnz);
        real2cell(v.maxx, v.maxy, v.maxz, &mxx, &mxy, &mxz);
        size_t volume_count = (mxx - mnx) * (mxy - mny) * (mxz - mnz);
        size_t max_nucs = volume_count * v.density;
        // select_nuc_volume (grain.c:168-201)
        int st = distribute(volume_count, max_nucs, [&](size_t index) {
            size_t nx = mxx - mnx, ny = mxy - mny;
            size_t vx = index % nx;
            size_t vy = (index % (nx * ny)) / (nx);
            size_t vz = index / (nx * ny);
            NucPt n;
            n.cx = mnx + vx;
            n.cy = mny + vy;
            n.cz = mnz + vz;
            n.threshold = v.threshold;
            set_orientation(n);
            n.x = cell2real(0, (long) n.cx);
            n.y = cell2real(1, (long) n.cy);
            n.z = cell2real(2, (long) n.cz);
            g.push_back(n);
            return MMCA_OK;
        });
        if (st != MMCA_OK)
            return st;
    }
    // points of this subblock (grain.c:695-711)
    ctx->nuc_pt.clear();
    for (const NucPt &n : g) {
        size_t ix, iy, iz;
        real2cell(n.x, n.y, n.z, &ix, &iy, &iz);
        if ((int) (ix / p.nx) == p.sb_coords[0] && (int) (iy / p.ny) == p.sb_coords[1] &&
            (int) (iz / p.nz) == p.sb_coords[2])
            ctx->nuc_pt.push_back(n);
    }

    // init_sb_nucleation: dense thresholds while sampling
    const long long P = c.P;
    std::vector<float> thr((size_t) P, INFINITY);
    auto is_mold = [&](long x, long y, long z) {
        return (c.pad[0] && (x == 0 || x == c.nx + 1)) || (c.pad[1] && (y == 0 || y == c.ny + 1)) ||
               (c.pad[2] && (z == 0 || z == c.nz + 1));
    };
    for (const NucPt &n : ctx->nuc_pt) {
        size_t x = 1 + n.cx % p.nx, y = 1 + n.cy % p.ny, z = 1 + n.cz % p.nz;
        size_t index = z * (p.ny + 2) * (p.nx + 2) + y * (p.nx + 2) + x;
        if (!is_mold((long) x, (long) y, (long) z))
            thr[index] = (float) n.threshold;
    }
    std::vector<long long> vol_cells, suf_cells;
    vol_cells.reserve((size_t) c.nx * c.ny * c.nz);
    for (int z = 1; z <= c.nz; z++)
        for (int y = 1; y <= c.ny; y++)
            for (int x = 1; x <= c.nx; x++) {
                long long i = x + (long long) y * c.sx + z * c.sxy;
                if (is_mold(x, y, z))
                    continue;
                vol_cells.push_back(i);
                int surf = is_mold(x - 1, y, z) + is_mold(x + 1, y, z) + is_mold(x, y - 1, z) +
                           is_mold(x, y + 1, z) + is_mold(x, y, z - 1) + is_mold(x, y, z + 1);
                if (surf)
                    suf_cells.push_back(i);
            }
    int status = MMCA_OK;
    auto gen_threshold = [&](double center, double sigma) -> float {
        // genRandomThreshold (grain.c:141-154), bounded
        float t = 0.0;
        long tries = 0;
        do {
            if (++tries > 100000000L) {
                status = MMCA_ERR_NUC_THRESHOLD;
                return INFINITY;
            }
            t = (float) (sigma * rng.gaussian() + center);
        } while ((t < 1.0e-5) || (fabs(t - center) > (6.0 * sigma)));
        return t;
    };
    auto sample = [&](const std::vector<long long> &cells, double density, double exponent,
                      double center, double sigma) -> int {
        size_t gr_count = density * (cells.size() * pow(p.cell_size, exponent));
        return distribute(cells.size(), gr_count, [&](size_t index) {
            // select_if_not_fixed (grain.c:226-237)
            long long myindex = cells[index];
            if (thr[(size_t) myindex] == INFINITY)
                thr[(size_t) myindex] = gen_threshold(center, sigma);
            return status;
        });
    };
    TRY(sample(vol_cells, cfg->bulk_density, 3.0, cfg->bulk_mean, cfg->bulk_sigma));
    if (cfg->do_surface && (cfg->surf_density > 0.0))
        TRY(sample(suf_cells, cfg->surf_density, 2.0, cfg->surf_mean, cfg->surf_sigma));

    // sites in raster order, with the FixedNuc/NucVolume entry each matches
    // (createGrains takes the first match, grain.c:392-401)
    ctx->site_idx.clear();
    ctx->site_thr.clear();
    ctx->site_fixed.clear();
    std::vector<int> site_ls;
    std::vector<double> fixed_rot(ctx->nuc_pt.size() * 9 + 9);
    for (size_t j = 0; j < ctx->nuc_pt.size(); j++)
        rotmat_of(ctx->nuc_pt[j].gx, ctx->nuc_pt[j].gy, ctx->nuc_pt[j].gz, &fixed_rot[j * 9]);
    const double t0 = ctx->hc.timestep * c.dt;
    for (long long i = 0; i < P; i++) {
        if (thr[(size_t) i] == INFINITY)
            continue;
        int z = (int) (i / c.sxy);
        long long rem = i - z * c.sxy;
        int y = (int) (rem / c.sx);
        int x = (int) (rem - (long long) y * c.sx);
        size_t gx = (size_t) (x - 1 + c.off[0]), gy = (size_t) (y - 1 + c.off[1]),
               gz = (size_t) (z - 1 + c.off[2]);
        int fixed = -1;
        for (size_t j = 0; j < ctx->nuc_pt.size(); j++)
            if (ctx->nuc_pt[j].cx == gx && ctx->nuc_pt[j].cy == gy && ctx->nuc_pt[j].cz == gz) {
                fixed = (int) j;
                break;
            }
        ctx->site_idx.push_back((int) i);
        ctx->site_thr.push_back(thr[(size_t) i]);
        ctx->site_fixed.push_back(fixed);
        int ls = 0;
        if (c.am)   // tempPrepareSB_int (temperature.c:183-189)
            ls = host_temp(c, t0, x, y, z) >= c.Tliq ? 1 : 0;
        site_ls.push_back(ls);
    }
    DevBufs &b = ctx->b;
    const size_t ns = ctx->site_idx.size();
    b.nsites = (int) ns;
    TRY(dev_alloc(&b.site_idx, ns));
    TRY(dev_alloc(&b.site_thr, ns));
    TRY(dev_alloc(&b.site_fixed, ns));
    TRY(dev_alloc(&b.site_ls, ns));
    TRY(dev_alloc(&b.events, ns));
    TRY(dev_alloc(&b.fixed_rot, fixed_rot.size()));
    if (ns) {
        HIPCHK(hipMemcpy(b.site_idx, ctx->site_idx.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.site_thr, ctx->site_thr.data(), ns * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.site_fixed, ctx->site_fixed.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.site_ls, site_ls.data(), ns * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(b.fixed_rot, fixed_rot.data(), fixed_rot.size() * sizeof(double),
                     hipMemcpyHostToDevice));
    ctx->nucleation_ready = true;
    return MMCA_OK;
}

int mmca_step(mmca_ctx *ctx, int nsteps, int want_sums, mmca_report *report)
{
    if (!ctx || nsteps < 0)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    if (ctx->hc.status != MMCA_OK)
        return ctx->hc.status;
    const int64_t start_new = ctx->hc.new_grains;
    (void) start_new;
    ctx->hc.new_grains = 0;
    HIPCHK(hipMemcpyAsync(&ctx->b.cnt->new_grains, &ctx->hc.new_grains, sizeof(int),
                          hipMemcpyHostToDevice, ctx->stream));
    int st = courant_check(ctx);
    if (st != MMCA_OK && nsteps > 0) {
        ctx->hc.status = st;
        ctx->hc.failed_step = ctx->hc.timestep + 1;
        TRY(push_counters(ctx));
    } else {
        long long need = (long long) ctx->hc.random_used + (long long) nsteps * ctx->b.nsites;
        TRY(ensure_random(ctx, need));
        for (int n = 0; n < nsteps; n++)
            TRY(enqueue_iteration(ctx, true));
        if (want_sums)
            HIPCHK(launch_reduce(ctx->b, ctx->c, ctx->cur, ctx->d_red, ctx->stream));
        TRY(sync_counters(ctx));
        if (ctx->timing)
            TRY(collect_timing(ctx));
    }
    if (report) {
        memset(report, 0, sizeof(*report));
        report->timestep = ctx->hc.timestep;
        report->num_grains = ctx->hc.num_grains;
        report->new_grains = ctx->hc.new_grains;
        report->gindex = ctx->hc.gindex;
        report->status = ctx->hc.status;
        report->failed_step = ctx->hc.failed_step;
        if (want_sums && ctx->hc.status == MMCA_OK) {
            double s2[2];
            HIPCHK(hipMemcpy(s2, ctx->d_red, sizeof(s2), hipMemcpyDeviceToHost));
            report->solid_sum = s2[0];
            report->solute_sum = s2[1];
        }
    }
    return ctx->hc.status;
}

int mmca_set_timestep(mmca_ctx *ctx, uint64_t timestep)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    ctx->hc.timestep = timestep;
    return push_counters(ctx);
}

int mmca_get_timestep(mmca_ctx *ctx, uint64_t *timestep)
{
    if (!ctx || !timestep)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    *timestep = ctx->hc.timestep;
    return MMCA_OK;
}

int mmca_cell_nucleation(mmca_ctx *ctx, int32_t *n_new)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_nucleate(ctx->b, ctx->c, 1, ctx->cur, ctx->stream));
    TRY(sync_counters(ctx));
    if (n_new)
        *n_new = ctx->hc.n_events;
    return ctx->hc.status;
}

int mmca_activate_new_grains(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    TRY(ensure_random(ctx, (long long) ctx->hc.random_used + ctx->hc.n_events));
    HIPCHK(launch_nucleate(ctx->b, ctx->c, 2, ctx->cur, ctx->stream));
    return check_status(ctx);
}

int mmca_sb_diffuse_alloy_decentered(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    int st = courant_check(ctx);
    if (st != MMCA_OK)
        return st;
    TRY(ensure_oce(ctx));
    HIPCHK(launch_solute(ctx->b, ctx->c, 1, 0, ctx->cur, ctx->cur, ctx->stream));
    ctx->have_oce = true;
    return check_status(ctx);
}

int mmca_fs_change_diffuse(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    if (!ctx->have_oce)
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_solute(ctx->b, ctx->c, 0, 1, ctx->cur, ctx->cur, ctx->stream));
    return check_status(ctx);
}

int mmca_grow_octahedra(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(ensure_d(ctx));
    HIPCHK(launch_grow_octahedra(ctx->b, ctx->c, ctx->cur, ctx->stream));
    return check_status(ctx);
}

int mmca_grow_cell_reduction(mmca_ctx *ctx, int32_t *gindex)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_mask_from_gr(ctx->b, ctx->c, ctx->stream));
    HIPCHK(launch_list(ctx->b, ctx->c, ctx->cur, 0, 0, ctx->stream));
    TRY(sync_counters(ctx));
    if (gindex)
        *gindex = ctx->hc.gindex;
    return ctx->hc.status;
}

int mmca_capture_octahedra_diffuse(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_capture_only(ctx->b, ctx->c, ctx->cur, ctx->stream));
    return check_status(ctx);
}

int mmca_doiteration(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    if (ctx->hc.status != MMCA_OK)
        return ctx->hc.status;
    int st = courant_check(ctx);
    if (st != MMCA_OK)
        return st;
    TRY(ensure_random(ctx, (long long) ctx->hc.random_used + ctx->b.nsites));
    TRY(enqueue_iteration(ctx, false));
    TRY(check_status(ctx));
    if (ctx->timing)
        TRY(collect_timing(ctx));
    return MMCA_OK;
}

int mmca_solid_volume(mmca_ctx *ctx, double *volume)
{
    if (!ctx || !volume)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_reduce(ctx->b, ctx->c, ctx->cur, ctx->d_red, ctx->stream));
    double s2[2];
    HIPCHK(hipMemcpyAsync(s2, ctx->d_red, sizeof(s2), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *volume = s2[0] * pow(ctx->p.cell_size, 3.0);   // growth.c:977
    return MMCA_OK;
}

int mmca_upload_field(mmca_ctx *ctx, int field, const void *host)
{
    if (!ctx || !host)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t P = (size_t) ctx->c.P;
    DevBufs &b = ctx->b;
    switch (field) {
    case MMCA_FIELD_CE:
        HIPCHK(hipMemcpy(b.ce, host, P * sizeof(double), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_CL:
        HIPCHK(hipMemcpy(b.cl[0], host, P * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.cl[1], host, P * sizeof(double), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_FS:
        HIPCHK(hipMemcpy(b.fs[0], host, P * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.fs[1], host, P * sizeof(double), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_GR:
        HIPCHK(hipMemcpy(b.gr, host, P * sizeof(int), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_DC: {
        const double *h = (const double *) host;
        std::vector<double> tmp(P);
        double *dst[3] = { b.dcx, b.dcy, b.dcz };
        for (int a = 0; a < 3; a++) {
            for (size_t i = 0; i < P; i++)
                tmp[i] = h[3 * i + a];
            HIPCHK(hipMemcpy(dst[a], tmp.data(), P * sizeof(double), hipMemcpyHostToDevice));
        }
        break;
    }
    case MMCA_FIELD_OCE:
        TRY(ensure_oce(ctx));
        HIPCHK(hipMemcpy(b.oce, host, P * sizeof(double), hipMemcpyHostToDevice));
        ctx->have_oce = true;
        break;
    default:
        return MMCA_ERR_ARG;
    }
    return MMCA_OK;
}

int mmca_download_field(mmca_ctx *ctx, int field, void *host)
{
    if (!ctx || !host)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const Consts &c = ctx->c;
    const size_t P = (size_t) c.P;
    DevBufs &b = ctx->b;
    switch (field) {
    case MMCA_FIELD_CE:
        HIPCHK(hipMemcpy(host, b.ce, P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_CL:
        HIPCHK(hipMemcpy(host, b.cl[ctx->cur], P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_FS:
        HIPCHK(hipMemcpy(host, b.fs[ctx->cur], P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_GR:
        HIPCHK(hipMemcpy(host, b.gr, P * sizeof(int), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_DC: {
        std::vector<double> tmp(P);
        double *h = (double *) host;
        double *src[3] = { b.dcx, b.dcy, b.dcz };
        for (int a = 0; a < 3; a++) {
            HIPCHK(hipMemcpy(tmp.data(), src[a], P * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < P; i++)
                h[3 * i + a] = tmp[i];
        }
        break;
    }
    case MMCA_FIELD_D:
        TRY(ensure_d(ctx));
        HIPCHK(launch_grow_octahedra(ctx->b, c, ctx->cur, ctx->stream));
        HIPCHK(hipMemcpyAsync(host, b.d, P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        break;
    case MMCA_FIELD_TEMP: {
        double *tmp = nullptr;
        TRY(dev_alloc(&tmp, P));
        HIPCHK(launch_fill_temp(b, c, tmp, ctx->stream));
        HIPCHK(hipMemcpyAsync(host, tmp, P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(tmp));
        break;
    }
    case MMCA_FIELD_THRESHOLD: {
        float *h = (float *) host;
        for (size_t i = 0; i < P; i++)
            h[i] = INFINITY;
        for (size_t s = 0; s < ctx->site_idx.size(); s++)
            h[ctx->site_idx[s]] = ctx->site_thr[s];
        break;
    }
    case MMCA_FIELD_MOLD: {
        int8_t *h = (int8_t *) host;
        for (int z = 0; z < c.nz + 2; z++)
            for (int y = 0; y < c.ny + 2; y++)
                for (int x = 0; x < c.sx; x++) {
                    bool m = (c.pad[0] && (x == 0 || x == c.nx + 1)) ||
                             (c.pad[1] && (y == 0 || y == c.ny + 1)) ||
                             (c.pad[2] && (z == 0 || z == c.nz + 1));
                    h[x + (size_t) y * c.sx + (size_t) z * c.sxy] = m ? 1 : 0;
                }
        break;
    }
    case MMCA_FIELD_OCE:
        if (!ctx->have_oce)
            return MMCA_ERR_STATE;
        HIPCHK(hipMemcpy(host, b.oce, P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    default:
        return MMCA_ERR_ARG;
    }
    return MMCA_OK;
}

int mmca_get_grains(mmca_ctx *ctx, mmca_grain *out, int32_t max, int32_t *num_grains)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    const int ng = ctx->hc.num_grains;
    if (num_grains)
        *num_grains = ng;
    if (!out)
        return MMCA_OK;
    const int n = std::min(ng, (int) max);
    if (n <= 0)
        return MMCA_OK;
    std::vector<double> rot((size_t) n * 9);
    std::vector<int> cell(n), src(n);
    std::vector<unsigned long long> step(n);
    HIPCHK(hipMemcpy(rot.data(), ctx->b.g_rot, rot.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cell.data(), ctx->b.g_cell, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(src.data(), ctx->b.g_src, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(step.data(), ctx->b.g_step, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    const Consts &c = ctx->c;
    const mmca_params &p = ctx->p;
    memset(out, 0, sizeof(mmca_grain) * (size_t) n);
    for (int id = 1; id < n; id++) {
        mmca_grain &g = out[id];
        for (int q = 0; q < 9; q++)
            g.rotmat[q / 3][q % 3] = rot[(size_t) id * 9 + q];
        g.nuc_timestep = step[id];
        if (src[id] >= 0) {      // nucleate_grain with a FixedNuc/NucVolume entry
            const NucPt &np = ctx->nuc_pt[src[id]];
            g.nuc_x = np.x;
            g.nuc_y = np.y;
            g.nuc_z = np.z;
            g.nuc_threshold = np.threshold;
            g.rotang[0] = np.gx;
            g.rotang[1] = np.gy;
            g.rotang[2] = np.gz;
        } else {                 // createGrains' random branch (grain.c:402-414)
            int k = -src[id] - 1;
            int idx = cell[id];
            int z = (int) (idx / c.sxy);
            long long rem = idx - z * c.sxy;
            int y = (int) (rem / c.sx);
            int x = (int) (rem - (long long) y * c.sx);
            int gx = x - 1 + c.off[0], gy = y - 1 + c.off[1], gz = z - 1 + c.off[2];
            g.nuc_x = p.origin[0] + (gx + 0.5) * p.cell_size;
            g.nuc_y = p.origin[1] + (gy + 0.5) * p.cell_size;
            g.nuc_z = p.origin[2] + (gz + 0.5) * p.cell_size;
            auto it = std::lower_bound(ctx->site_idx.begin(), ctx->site_idx.end(), idx);
            g.nuc_threshold = (it != ctx->site_idx.end() && *it == idx)
                                  ? (double) ctx->site_thr[it - ctx->site_idx.begin()]
                                  : INFINITY;
            if (k < (int) (ctx->rand_angles.size() / 3)) {
                g.rotang[0] = ctx->rand_angles[(size_t) k * 3 + 0];
                g.rotang[1] = ctx->rand_angles[(size_t) k * 3 + 1];
                g.rotang[2] = ctx->rand_angles[(size_t) k * 3 + 2];
            }
        }
    }
    return MMCA_OK;
}

int mmca_get_interface_list(mmca_ctx *ctx, int32_t *out, int32_t max, int32_t *gindex)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    const int g = ctx->hc.gindex;
    if (gindex)
        *gindex = g;
    const int n = std::min(g, (int) max);
    if (out && n > 0)
        HIPCHK(hipMemcpy(out, ctx->b.diff_id + 1, (size_t) n * sizeof(int), hipMemcpyDeviceToHost));
    return MMCA_OK;
}

int mmca_padded_cells(mmca_ctx *ctx, int64_t *cells)
{
    if (!ctx || !cells)
        return MMCA_ERR_ARG;
    *cells = ctx->c.P;
    return MMCA_OK;
}

int mmca_set_kernel_timing(mmca_ctx *ctx, int enable)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    ctx->timing = enable != 0;
    return MMCA_OK;
}

int mmca_kernel_time_ms(mmca_ctx *ctx, double *avg_ms, int64_t *launches)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    if (avg_ms)
        *avg_ms = ctx->time_ms;
    if (launches)
        *launches = ctx->time_n;
    return MMCA_OK;
}

}  // extern "C"
