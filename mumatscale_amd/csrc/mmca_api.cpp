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

#include "host_init.h"
#include "mmca_internal.h"

namespace mmca {
hipError_t launch_nucleate(const View &v, int mode, int cur, hipStream_t s);
hipError_t launch_beam_field(const DevBufs &b, const Consts &c, int adv, hipStream_t s);
hipError_t launch_solute(const View &v, int am, int f32, int zm_by, int lds, int diff, int fsc, int rec, int cur, int nxt,
                         hipStream_t s);
hipError_t launch_mask_from_gr(const DevBufs &b, const Consts &c, hipStream_t s);
hipError_t launch_list(const View &v, int cb, int capture, int decide_realloc, hipStream_t s);
hipError_t launch_nucleate_coupled(const View &v, int cur, hipStream_t s);
hipError_t launch_halo_exchange(const View &v, int phase, int cb, long long max_face, hipStream_t s);
hipError_t launch_io_fields(const DevBufs &b, const Consts &c, int cur, float *fs_io, float *ce_io, float *t_io,
                            int *gr_io, hipStream_t s);
hipError_t launch_capture_only(const View &v, const DevBufs &b, int cb, hipStream_t s);
int nuc_blocks(int nsites);
long long cell_blocks(const Consts &c);
long long zm_tiles(const Consts &c, int by);
hipError_t launch_reduce(const DevBufs &b, const Consts &c, int cb, double *out2, hipStream_t s);
hipError_t launch_fill_temp(const DevBufs &b, const Consts &c, double *out, hipStream_t s);
hipError_t launch_grow_octahedra(const DevBufs &b, const Consts &c, int cb, hipStream_t s);
hipError_t launch_fill_f64(double *p, long long n, double v, hipStream_t s);
hipError_t launch_eval_exp(const double *x, double *y, long long n, hipStream_t s);
hipError_t launch_probe(int mix, long long n, double *a, double *b, double *f, float *ea, float *eb, int *g,
                        hipStream_t s);
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
    // device view of this context as a one-domain launch
    DomainDesc *d_desc = nullptr;
    long long *d_off = nullptr;
    View view;
    bool view_dirty = true;
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev;       // pool, reused
    size_t ev_used = 0;
    double time_ms = 0.0;
    long long time_n = 0;
    // grain records set by mmca_set_grains (ids below ext_grains.size())
    std::vector<mmca_grain> ext_grains;
    // staging buffer of mmca_io_fields: 4 x P 32-bit values (allocated on first use)
    float *io_buf = nullptr;
    // The random stream at the start of the orientation queue (before draw
    // q0).  mmca_rand draws from rng_now, the stream right after the
    // orientations consumed so far; the queue from there on is then redrawn
    // before the next step (rand_dirty).
    GlibcRand rng_q0;
    long long q0 = 0;
    GlibcRand rng_now;
    bool rand_dirty = false;
    // a domain of a coupled batch other than rank 0: its grain table is rank
    // 0's (its own was released)
    bool shares_grains = false;
    // the domains of a coupled batch in rank order (empty otherwise): grain
    // records are read through the domain that created the grain
    std::vector<mmca_ctx *> peers;
};

namespace {

// g_src of a grain whose record came from mmca_set_grains
constexpr int GRAIN_SET = -2147483647 - 1;

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
    draw_orientations(ctx->rng, ctx->oriented, n, &ctx->rand_angles[(size_t) from * 3], rot.data());
    HIPCHK(hipMemcpyAsync(ctx->b.rand_rot + (size_t) from * 9, rot.data(), rot.size() * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->rand_filled = (int) upto;
    return MMCA_OK;
}

// After mmca_rand: the queued orientations past the consumed ones were drawn
// from the stream before those raw values; drop them so that the next ones
// are drawn after them.
int replan_random(mmca_ctx *ctx)
{
    if (!ctx->rand_dirty)
        return MMCA_OK;
    ctx->rng = ctx->rng_now;
    ctx->rng_q0 = ctx->rng_now;
    ctx->q0 = ctx->hc.random_used;
    ctx->rand_filled = ctx->hc.random_used;
    ctx->rand_angles.resize((size_t) ctx->rand_filled * 3);
    ctx->rand_dirty = false;
    return MMCA_OK;
}

int ensure_oce(mmca_ctx *ctx)
{
    if (!ctx->b.oce) {
        TRY(dev_alloc(&ctx->b.oce, (size_t) ctx->c.P));
        ctx->view_dirty = true;
    }
    return MMCA_OK;
}

int ensure_d(mmca_ctx *ctx)
{
    if (!ctx->b.d) {
        TRY(dev_alloc(&ctx->b.d, (size_t) ctx->c.P));
        ctx->view_dirty = true;
        HIPCHK(hipMemsetAsync(ctx->b.d, 0, (size_t) ctx->c.P * sizeof(double), ctx->stream));
    }
    return MMCA_OK;
}

// Grain ids of a gr field from outside (an upload, a restart image): every
// id indexes the grain table of max_total_grains + 1 slots.
int check_grain_ids(const void *gr, size_t n, long long max_total_grains)
{
    for (size_t i = 0; i < n; i++) {
        int32_t g;
        memcpy(&g, (const char *) gr + i * sizeof(int32_t), sizeof(g));
        if ((long long) g > max_total_grains)
            return MMCA_ERR_ARG;
    }
    return MMCA_OK;
}

// buffers of the separable beam temperature (Consts::beam_sep)
int ensure_beam(mmca_ctx *ctx)
{
    DevBufs &b = ctx->b;
    const Consts &c = ctx->c;
    if (!b.temp) {
        // X rows of whole 128-column tiles, W rows in whole 64-row tiles (zero past ny nz)
        const size_t wrows = ((size_t) c.ny * c.nz + 63) / 64 * 64;
        TRY(dev_alloc(&b.temp, (size_t) c.P));
        TRY(dev_alloc(&b.beam_x, (size_t) BEAM_NT * ((c.nx + 127) / 128 * 128)));
        TRY(dev_alloc(&b.beam_y, (size_t) BEAM_NT * c.ny));
        TRY(dev_alloc(&b.beam_z, (size_t) BEAM_NT * c.nz));
        TRY(dev_alloc(&b.beam_w, (size_t) BEAM_NT * wrows));
        TRY(dev_alloc(&b.beam_key, 2));
        HIPCHK(hipMemsetAsync(b.temp, 0, (size_t) c.P * sizeof(double), ctx->stream));
        HIPCHK(hipMemsetAsync(b.beam_w, 0, (size_t) BEAM_NT * wrows * sizeof(double), ctx->stream));
        const double key[2] = { NAN, 1.0 };      // no position yet: W is built on the first step
        HIPCHK(hipMemcpyAsync(b.beam_key, key, sizeof(key), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx->view_dirty = true;
    }
    return MMCA_OK;
}

// the separable beam field of timestep + adv, before a kernel that reads it
int refresh_beam(mmca_ctx *ctx, int adv, hipStream_t s)
{
    if (ctx->c.beam_sep)
        HIPCHK(launch_beam_field(ctx->b, ctx->c, adv, s));
    return MMCA_OK;
}

int check_status(mmca_ctx *ctx)
{
    TRY(sync_counters(ctx));
    return ctx->hc.status;
}

int zm_rows(const Consts &c)
{
    return c.zm_planes > 0 ? c.zm_by : 0;
}

// Upload the descriptors of n domains and the per-kernel workgroup prefix
// tables; fills `v`.  `desc`/`off` are (re)allocated to hold n domains.
int build_view(mmca_ctx *const *ctxs, int n, int xcd_remap, int zm_by, DomainDesc **desc, long long **off,
               int *cap, View *v, hipStream_t s)
{
    if (*cap < n) {
        if (*desc)
            (void) hipFree(*desc);
        if (*off)
            (void) hipFree(*off);
        *desc = nullptr;
        *off = nullptr;
        TRY(dev_alloc(desc, (size_t) n));
        TRY(dev_alloc(off, (size_t) 4 * (n + 1)));
        *cap = n;
    }
    std::vector<DomainDesc> h(n);
    std::vector<long long> t((size_t) 4 * (n + 1), 0);
    long long *on = &t[0], *oz = &t[n + 1], *oc = &t[2 * (n + 1)], *ol = &t[3 * (n + 1)];
    for (int d = 0; d < n; d++) {
        h[d].b = ctxs[d]->b;
        h[d].c = ctxs[d]->c;
        on[d + 1] = on[d] + nuc_blocks(ctxs[d]->b.nsites);
        oz[d + 1] = oz[d] + (zm_by > 0 ? zm_tiles(ctxs[d]->c, zm_by) : 0);
        oc[d + 1] = oc[d] + cell_blocks(ctxs[d]->c);
        ol[d + 1] = ol[d] + list_blocks(ctxs[d]->c);
    }
    HIPCHK(hipMemcpyAsync(*desc, h.data(), sizeof(DomainDesc) * (size_t) n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(*off, t.data(), sizeof(long long) * t.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    v->dom = *desc;
    v->ndom = n;
    v->xcd_remap = xcd_remap;
    v->zm_planes_max = 0;
    for (int d = 0; d < n; d++)
        v->zm_planes_max = std::max(v->zm_planes_max, ctxs[d]->c.zm_planes);
    v->off_nuc = *off;
    v->off_zm = *off + (n + 1);
    v->off_cell = *off + 2 * (n + 1);
    v->off_list = *off + 3 * (n + 1);
    v->tot_nuc = on[n];
    v->tot_zm = oz[n];
    v->tot_cell = oc[n];
    v->tot_list = ol[n];
    return MMCA_OK;
}

int ensure_view(mmca_ctx *ctx)
{
    if (!ctx->view_dirty)
        return MMCA_OK;
    int cap = ctx->d_desc ? 1 : 0;
    mmca_ctx *one[1] = { ctx };
    TRY(build_view(one, 1, ctx->c.xcd_remap, zm_rows(ctx->c), &ctx->d_desc, &ctx->d_off, &cap, &ctx->view,
                   ctx->stream));
    ctx->view_dirty = false;
    return MMCA_OK;
}

// solute lost to FP32 storage rounding: sum over workgroup slots of the
// Kahan pairs (sum - compensation)
int read_ledger(mmca_ctx *ctx, double *out)
{
    std::vector<double> l((size_t) (2 * ctx->b.nledger));
    HIPCHK(hipMemcpy(l.data(), ctx->b.ledger, l.size() * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0.0, comp = 0.0;
    for (long long i = 0; i < ctx->b.nledger; i++) {
        double v = l[(size_t) (2 * i)] - l[(size_t) (2 * i + 1)];
        double y = v - comp, t = s + y;
        comp = (t - s) - y;
        s = t;
    }
    *out = s;
    return MMCA_OK;
}

// one doiteration() on the device (calculate.c:193-352)
int enqueue_iteration(mmca_ctx *ctx, bool advance)
{
    const int nxt = ctx->cur ^ 1;
    TRY(ensure_view(ctx));
    TRY(refresh_beam(ctx, advance ? 1 : 0, ctx->stream));
    HIPCHK(launch_nucleate(ctx->view, advance ? 7 : 3, ctx->cur, ctx->stream));
    const size_t ev0 = ctx->ev_used;
    if (ctx->timing) {
        while (ctx->ev.size() < ev0 + 2) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ctx->ev.push_back(e);
        }
        HIPCHK(hipEventRecord(ctx->ev[ev0], ctx->stream));
    }
    const int rec = ctx->b.rec != nullptr;
    HIPCHK(launch_solute(ctx->view, ctx->c.am && !ctx->c.beam_sep, ctx->c.f32, rec ? 0 : zm_rows(ctx->c),
                         ctx->c.zm_lds, 1, 1, rec, ctx->cur, nxt, ctx->stream));
    if (ctx->timing) {
        HIPCHK(hipEventRecord(ctx->ev[ev0 + 1], ctx->stream));
        ctx->ev_used = ev0 + 2;
    }
    HIPCHK(launch_list(ctx->view, nxt, 1, 1, ctx->stream));
    ctx->cur = nxt;
    return MMCA_OK;
}

int collect_timing(mmca_ctx *ctx)
{
    double tot = 0.0;
    long long n = 0;
    for (size_t i = 0; i + 1 < ctx->ev_used; i += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
        tot += ms;
        n++;
    }
    ctx->ev_used = 0;
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

int mmca_eval_exp(int device, const double *x, double *y, int64_t n)
{
    if (n < 0 || (n > 0 && (!x || !y)))
        return MMCA_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return MMCA_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev)
        return MMCA_ERR_ARG;
    if (n == 0)
        return MMCA_OK;
    if (hipSetDevice(device) != hipSuccess)
        return MMCA_ERR_HIP;
    const int64_t chunk = std::min<int64_t>(n, (int64_t) 1 << 24);
    double *dx = nullptr, *dy = nullptr;
    int st = MMCA_OK;
    if (hipMalloc(&dx, (size_t) chunk * sizeof(double)) != hipSuccess ||
        hipMalloc(&dy, (size_t) chunk * sizeof(double)) != hipSuccess)
        st = MMCA_ERR_HIP;
    for (int64_t o = 0; st == MMCA_OK && o < n; o += chunk)
    {
        const int64_t m = std::min(chunk, n - o);
        if (hipMemcpy(dx, x + o, (size_t) m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            launch_eval_exp(dx, dy, (long long) m, nullptr) != hipSuccess ||
            hipMemcpy(y + o, dy, (size_t) m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            st = MMCA_ERR_HIP;
    }
    if (dx)
        (void) hipFree(dx);
    if (dy)
        (void) hipFree(dy);
    return st;
}

int mmca_bandwidth_probe(int device, int mix, int64_t cells, int reps, double *ms, double *bytes_per_cell)
{
    static const double bpc[3] = { 16.0, 52.0, 36.0 };
    if (mix < 0 || mix > 2 || cells <= 0 || cells > (1ll << 40) || reps <= 0 || !ms)
        return MMCA_ERR_ARG;   // one workgroup per 1024 cells: the grid stays below 2^31
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return MMCA_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess)
        return MMCA_ERR_ARG;
    const size_t n = (size_t) cells;
    double *a = nullptr, *b = nullptr, *f = nullptr;
    float *ea = nullptr, *eb = nullptr;
    int *g = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = MMCA_OK;
    bool ok = hipStreamCreate(&s) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess;
    if (ok && mix != 2)
        ok = hipMalloc(&a, n * sizeof(double)) == hipSuccess && hipMalloc(&b, n * sizeof(double)) == hipSuccess;
    if (ok && mix != 0)
        ok = hipMalloc(&f, n * sizeof(double)) == hipSuccess && hipMalloc(&g, n * sizeof(int)) == hipSuccess;
    if (ok && mix == 2)
        ok = hipMalloc(&ea, n * sizeof(float)) == hipSuccess && hipMalloc(&eb, n * sizeof(float)) == hipSuccess;
    if (ok)
        ok = (!a || hipMemsetAsync(a, 0, n * sizeof(double), s) == hipSuccess) &&
             (!b || hipMemsetAsync(b, 0, n * sizeof(double), s) == hipSuccess) &&
             (!f || hipMemsetAsync(f, 0, n * sizeof(double), s) == hipSuccess) &&
             (!g || hipMemsetAsync(g, 0, n * sizeof(int), s) == hipSuccess) &&
             (!ea || hipMemsetAsync(ea, 0, n * sizeof(float), s) == hipSuccess) &&
             (!eb || hipMemsetAsync(eb, 0, n * sizeof(float), s) == hipSuccess);
    if (ok)
        ok = launch_probe(mix, (long long) n, a, b, f, ea, eb, g, s) == hipSuccess;   // warm-up
    if (ok)
        ok = hipEventRecord(e0, s) == hipSuccess;
    for (int r = 0; ok && r < reps; ++r)
        ok = launch_probe(mix, (long long) n, a, b, f, ea, eb, g, s) == hipSuccess;
    float t = 0.0f;
    if (ok)
        ok = hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&t, e0, e1) == hipSuccess;
    if (!ok)
        st = MMCA_ERR_HIP;
    else
    {
        *ms = (double) t / reps;
        if (bytes_per_cell)
            *bytes_per_cell = bpc[mix];
    }
    for (void *p : { (void *) a, (void *) b, (void *) f, (void *) ea, (void *) eb, (void *) g })
        if (p)
            (void) hipFree(p);
    if (e0)
        (void) hipEventDestroy(e0);
    if (e1)
        (void) hipEventDestroy(e1);
    if (s)
        (void) hipStreamDestroy(s);
    return st;
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
    if (c.f32) {
        TRY(dev_alloc(&b.ce32, P));
        TRY(dev_alloc(&b.cl32[0], P));
        TRY(dev_alloc(&b.cl32[1], P));
        b.nledger = (long long) ((c.nx + 63) / 64) * ((c.ny + 3) / 4) * c.nz;
        TRY(dev_alloc(&b.ledger, (size_t) (2 * b.nledger)));
    } else {
        TRY(dev_alloc(&b.ce, P));
        TRY(dev_alloc(&b.cl[0], P));
        TRY(dev_alloc(&b.cl[1], P));
    }
    TRY(dev_alloc(&b.fs[0], P));
    TRY(dev_alloc(&b.fs[1], P));
    TRY(dev_alloc(&b.gr, P));
    TRY(dev_alloc(&b.dc, 3 * P));
    const size_t mask_words = (size_t) (c.ny + 2) * (c.nz + 2) * c.w64;
    TRY(dev_alloc(&b.mask, mask_words));
    const size_t xhalo_bytes = (size_t) 2 * (c.ny + 2) * (c.nz + 2);
    TRY(dev_alloc(&b.xhalo, xhalo_bytes));
    TRY(dev_alloc(&b.diff_id, N + 2));
    TRY(dev_alloc(&b.nbr, N + 2));
    // dc_tmp slots 0 .. N + 101 (dc_tmp_size <= N + 100, slot dc_tmp_size used)
    TRY(dev_alloc(&b.dtx, N + 102));
    TRY(dev_alloc(&b.dty, N + 102));
    TRY(dev_alloc(&b.dtz, N + 102));
    b.nblk_list = list_blocks(c);
    TRY(dev_alloc(&b.blk_count, (size_t) b.nblk_list));
    TRY(dev_alloc(&b.row_first, (size_t) c.ny * c.nz));
    b.nblk_red = reduce_blocks(c);
    TRY(dev_alloc(&b.red_part, (size_t) 2 * b.nblk_red));
    TRY(dev_alloc(&ctx->d_red, 2));
    const size_t G = (size_t) ctx->p.max_total_grains + 1;
    TRY(dev_alloc(&b.g_rot, G * 9));
    TRY(dev_alloc(&b.g_cell, G));
    TRY(dev_alloc(&b.g_step, G));
    TRY(dev_alloc(&b.g_src, G));
    TRY(dev_alloc(&b.g_dom, G));
    ctx->rand_capacity = (int) ctx->p.max_total_grains;
    TRY(dev_alloc(&b.rand_rot, (size_t) ctx->rand_capacity * 9));
    TRY(dev_alloc(&b.cnt, 1));
    if (c.beam_sep)
        TRY(ensure_beam(ctx));
    b.nsites = 0;
    // SetupInitialConditions (setup.c:58-65): ce = cl = Cinit, all else 0
    if (c.f32) {   // deviations from cinit
        HIPCHK(hipMemsetAsync(b.ce32, 0, P * sizeof(float), ctx->stream));
        HIPCHK(hipMemsetAsync(b.cl32[0], 0, P * sizeof(float), ctx->stream));
        HIPCHK(hipMemsetAsync(b.cl32[1], 0, P * sizeof(float), ctx->stream));
        HIPCHK(hipMemsetAsync(b.ledger, 0, (size_t) (2 * b.nledger) * sizeof(double), ctx->stream));
    } else {
        HIPCHK(launch_fill_f64(b.ce, (long long) P, c.cinit, ctx->stream));
        HIPCHK(launch_fill_f64(b.cl[0], (long long) P, c.cinit, ctx->stream));
        HIPCHK(launch_fill_f64(b.cl[1], (long long) P, c.cinit, ctx->stream));
    }
    HIPCHK(hipMemsetAsync(b.fs[0], 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.fs[1], 0, P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.gr, 0, P * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dc, 0, 3 * P * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.mask, 0, mask_words * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemsetAsync(b.xhalo, 0, xhalo_bytes, ctx->stream));
    HIPCHK(hipMemsetAsync(b.diff_id, 0, (N + 2) * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dtx, 0, (N + 102) * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dty, 0, (N + 102) * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.dtz, 0, (N + 102) * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_rot, 0, G * 9 * sizeof(double), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_cell, 0, G * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_step, 0, G * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_src, 0, G * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(b.g_dom, 0, G * sizeof(int), ctx->stream));
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
    if (ctx->shares_grains)
        b.g_rot = nullptr, b.g_cell = nullptr, b.g_step = nullptr, b.g_src = nullptr, b.g_dom = nullptr;
    void *ptrs[] = { b.ce, b.cl[0], b.cl[1], b.ce32, b.cl32[0], b.cl32[1], b.ledger, b.fs[0], b.fs[1], b.gr, b.dc, b.mask,
                     b.xhalo,
                     b.diff_id, b.nbr, b.dtx, b.dty, b.dtz, b.blk_count, b.row_first, b.site_idx, b.site_thr, b.site_fixed,
                     b.site_ls, b.events, b.ev_blk, b.ev_cnt, b.g_rot, b.g_cell, b.g_step, b.g_src, b.g_dom, b.fixed_rot,
                     b.rand_rot,
                     b.oce, b.d, b.temp, b.beam_x, b.beam_y, b.beam_z, b.beam_w, b.beam_key, b.red_part, b.cnt, ctx->d_red,
                     ctx->d_desc, ctx->d_off, ctx->io_buf };
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
    SitePlan plan;
    int st = plan_nucleation(ctx->p, ctx->c, cfg, ctx->hc.timestep, ctx->rng, plan);
    if (st != MMCA_OK)
        return st;
    ctx->oriented = cfg->oriented;
    ctx->nuc_pt = plan.nuc_pt;
    ctx->site_idx = plan.site_idx;
    ctx->site_thr = plan.site_thr;
    ctx->site_fixed = plan.site_fixed;
    std::vector<int> &site_ls = plan.site_ls;
    std::vector<double> &fixed_rot = plan.fixed_rot;
    DevBufs &b = ctx->b;
    const size_t ns = ctx->site_idx.size();
    b.nsites = (int) ns;
    TRY(dev_alloc(&b.site_idx, ns));
    TRY(dev_alloc(&b.site_thr, ns));
    TRY(dev_alloc(&b.site_fixed, ns));
    TRY(dev_alloc(&b.site_ls, ns));
    TRY(dev_alloc(&b.events, ns));
    const size_t nblk_nuc = (ns + 255) / 256;
    TRY(dev_alloc(&b.ev_blk, nblk_nuc * 256));
    TRY(dev_alloc(&b.ev_cnt, nblk_nuc));
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
    ctx->view_dirty = true;
    ctx->rng_q0 = ctx->rng;
    ctx->q0 = 0;
    ctx->rand_dirty = false;
    // draw the first orientations now, outside any timed stepping
    TRY(ensure_random(ctx, std::min<long long>(ctx->rand_capacity,
                                               std::max<long long>(65536, 256LL * (long long) ns))));
    return MMCA_OK;
}

int mmca_step(mmca_ctx *ctx, int nsteps, int want_sums, mmca_report *report)
{
    if (!ctx || nsteps < 0)
        return MMCA_ERR_ARG;
    if (!ctx->peers.empty())
        return MMCA_ERR_STATE;   // a rank of a coupled batch steps with the batch
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    if (ctx->hc.status != MMCA_OK)
        return ctx->hc.status;
    if (ctx->b.rec && ctx->hc.timestep + (uint64_t) nsteps > ctx->b.rec_t0 + (uint64_t) ctx->b.rec_steps)
        return MMCA_ERR_ARG;     // past the end of the branch record
    TRY(replan_random(ctx));
    ctx->hc.new_grains = 0;
    HIPCHK(hipMemcpyAsync(&ctx->b.cnt->new_grains, &ctx->hc.new_grains, sizeof(int),
                          hipMemcpyHostToDevice, ctx->stream));
    const int cur0 = ctx->cur;
    const uint64_t t0 = ctx->hc.timestep;
    int st = courant_check(ctx);
    if (st != MMCA_OK && nsteps > 0) {
        ctx->hc.status = st;
        ctx->hc.failed_step = ctx->hc.timestep + 1;
        TRY(push_counters(ctx));
    } else {
        // Steps are issued in chunks small enough that the orientations
        // already drawn cover every grain the chunk can create (at most one
        // per site and step); the counters are read back between chunks.
        int done = 0;
        while (done < nsteps) {
            long long k = nsteps - done;
            if (ctx->b.nsites > 0 && ctx->rand_filled < ctx->rand_capacity) {
                long long spare = (long long) ctx->rand_filled - ctx->hc.random_used;
                if (spare < (long long) ctx->b.nsites * 16) {
                    long long want = (long long) ctx->hc.random_used +
                                     std::max((long long) ctx->b.nsites * 256, 2 * spare + 65536);
                    TRY(ensure_random(ctx, want));
                    spare = (long long) ctx->rand_filled - ctx->hc.random_used;
                }
                if (ctx->rand_filled < ctx->rand_capacity)
                    k = std::min(k, spare / ctx->b.nsites);
            }
            if (k < 1)
                return MMCA_ERR_STATE;
            for (long long n = 0; n < k; n++)
                TRY(enqueue_iteration(ctx, true));
            done += (int) k;
            if (done < nsteps) {
                TRY(sync_counters(ctx));
                if (ctx->hc.status != MMCA_OK)
                    break;
            }
        }
        if (want_sums)
            HIPCHK(launch_reduce(ctx->b, ctx->c, ctx->cur, ctx->d_red, ctx->stream));
        TRY(sync_counters(ctx));
        // stopped on an error: the buffers of the failed step stay current
        // (the steps issued after it skipped every kernel)
        if (ctx->hc.status != MMCA_OK)
            ctx->cur = cur0 ^ (int) ((ctx->hc.timestep - t0) & 1);
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
            if (ctx->c.f32) {
                report->solute_sum += (double) ctx->c.nx * ctx->c.ny * ctx->c.nz * ctx->c.cinit;
                TRY(read_ledger(ctx, &report->solute_ledger));
            }
        }
    }
    return ctx->hc.status;
}

int mmca_set_record(mmca_ctx *ctx, uint8_t *codes, int64_t n_steps)
{
    if (!ctx || n_steps < 0)
        return MMCA_ERR_ARG;
    if (!ctx->peers.empty())
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    if (codes && n_steps > 0 && ctx->c.f32)
        return MMCA_ERR_ARG;     // the adjoint works on FP64 storage
    TRY(sync_counters(ctx));
    if (codes && n_steps > 0) {
        ctx->b.rec = codes;
        ctx->b.rec_t0 = ctx->hc.timestep;
        ctx->b.rec_steps = n_steps;
    } else {
        ctx->b.rec = nullptr;
        ctx->b.rec_t0 = 0;
        ctx->b.rec_steps = 0;
    }
    ctx->view_dirty = true;
    return MMCA_OK;
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
    TRY(ensure_view(ctx));
    TRY(refresh_beam(ctx, 0, ctx->stream));
    HIPCHK(launch_nucleate(ctx->view, 1, ctx->cur, ctx->stream));
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
    TRY(replan_random(ctx));
    TRY(ensure_random(ctx, (long long) ctx->hc.random_used + ctx->hc.n_events));
    TRY(ensure_view(ctx));
    HIPCHK(launch_nucleate(ctx->view, 2, ctx->cur, ctx->stream));
    return check_status(ctx);
}

int mmca_sb_diffuse_alloy_decentered(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    if (ctx->c.f32)
        return MMCA_ERR_STATE;   // single operators run on FP64 storage
    HIPCHK(hipSetDevice(ctx->device));
    int st = courant_check(ctx);
    if (st != MMCA_OK)
        return st;
    TRY(ensure_oce(ctx));
    // oce = ce over the padded box (calculate.c:267-280)
    HIPCHK(hipMemcpyAsync(ctx->b.oce, ctx->b.ce, (size_t) ctx->c.P * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    TRY(ensure_view(ctx));
    HIPCHK(launch_solute(ctx->view, ctx->c.am, 0, 0, 0, 1, 0, 0, ctx->cur, ctx->cur, ctx->stream));
    ctx->have_oce = true;
    return check_status(ctx);
}

int mmca_fs_change_diffuse(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    if (ctx->c.f32)
        return MMCA_ERR_STATE;   // single operators run on FP64 storage
    if (!ctx->have_oce)
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(ensure_view(ctx));
    HIPCHK(launch_solute(ctx->view, ctx->c.am, 0, 0, 0, 0, 1, 0, ctx->cur, ctx->cur, ctx->stream));
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
    TRY(ensure_view(ctx));
    HIPCHK(launch_list(ctx->view, ctx->cur, 0, 0, ctx->stream));
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
    TRY(ensure_view(ctx));
    HIPCHK(launch_capture_only(ctx->view, ctx->b, ctx->cur, ctx->stream));
    return check_status(ctx);
}

int mmca_doiteration(mmca_ctx *ctx)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    if (!ctx->peers.empty())
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    if (ctx->hc.status != MMCA_OK)
        return ctx->hc.status;
    int st = courant_check(ctx);
    if (st != MMCA_OK)
        return st;
    TRY(replan_random(ctx));
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
    case MMCA_FIELD_CL:
        if (ctx->c.f32) {
            const double *h = (const double *) host;
            std::vector<float> tmp(P);
            for (size_t i = 0; i < P; i++)
                tmp[i] = (float) (h[i] - ctx->c.cinit);
            if (field == MMCA_FIELD_CE) {
                HIPCHK(hipMemcpy(b.ce32, tmp.data(), P * sizeof(float), hipMemcpyHostToDevice));
            } else {
                HIPCHK(hipMemcpy(b.cl32[0], tmp.data(), P * sizeof(float), hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(b.cl32[1], tmp.data(), P * sizeof(float), hipMemcpyHostToDevice));
            }
        } else if (field == MMCA_FIELD_CE) {
            HIPCHK(hipMemcpy(b.ce, host, P * sizeof(double), hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemcpy(b.cl[0], host, P * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(b.cl[1], host, P * sizeof(double), hipMemcpyHostToDevice));
        }
        break;
    case MMCA_FIELD_FS:
        HIPCHK(hipMemcpy(b.fs[0], host, P * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.fs[1], host, P * sizeof(double), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_GR:
        TRY(check_grain_ids(host, P, ctx->p.max_total_grains));
        HIPCHK(hipMemcpy(b.gr, host, P * sizeof(int), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_DC:
        HIPCHK(hipMemcpy(b.dc, host, 3 * P * sizeof(double), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_D:
        // the reference's d array (kept with MMCA_TUNE_TRACK_D; a restart's input)
        TRY(ensure_d(ctx));
        HIPCHK(hipMemcpy(b.d, host, P * sizeof(double), hipMemcpyHostToDevice));
        break;
    case MMCA_FIELD_OCE:
        if (ctx->c.f32)
            return MMCA_ERR_STATE;
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
    case MMCA_FIELD_CL:
        if (c.f32) {
            std::vector<float> tmp(P);
            const float *src = (field == MMCA_FIELD_CE) ? b.ce32 : b.cl32[ctx->cur];
            HIPCHK(hipMemcpy(tmp.data(), src, P * sizeof(float), hipMemcpyDeviceToHost));
            double *h = (double *) host;
            for (size_t i = 0; i < P; i++)
                h[i] = c.cinit + (double) tmp[i];
        } else {
            const double *src = (field == MMCA_FIELD_CE) ? b.ce : b.cl[ctx->cur];
            HIPCHK(hipMemcpy(host, src, P * sizeof(double), hipMemcpyDeviceToHost));
        }
        break;
    case MMCA_FIELD_FS:
        HIPCHK(hipMemcpy(host, b.fs[ctx->cur], P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_GR:
        HIPCHK(hipMemcpy(host, b.gr, P * sizeof(int), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_DC:
        HIPCHK(hipMemcpy(host, b.dc, 3 * P * sizeof(double), hipMemcpyDeviceToHost));
        break;
    case MMCA_FIELD_D:
        TRY(ensure_d(ctx));
        HIPCHK(launch_grow_octahedra(ctx->b, c, ctx->cur, ctx->stream));
        HIPCHK(hipMemcpyAsync(host, b.d, P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        break;
    case MMCA_FIELD_TEMP: {
        double *tmp = nullptr;
        TRY(dev_alloc(&tmp, P));
        TRY(refresh_beam(ctx, 0, ctx->stream));
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
                    bool m = (c.pad_lo[0] && x == 0) || (c.pad_hi[0] && x == c.nx + 1) ||
                             (c.pad_lo[1] && y == 0) || (c.pad_hi[1] && y == c.ny + 1) ||
                             (c.pad_lo[2] && z == 0) || (c.pad_hi[2] && z == c.nz + 1);
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

int mmca_get_nucleation_points(mmca_ctx *ctx, mmca_nuc_point *out, int32_t max, int32_t *n)
{
    static_assert(sizeof(mmca_nuc_point) == sizeof(NucPt), "nucleation_t layout");
    if (!ctx || max < 0 || (max > 0 && !out))
        return MMCA_ERR_ARG;
    if (!ctx->nucleation_ready)
        return MMCA_ERR_STATE;
    const int32_t np = (int32_t) ctx->nuc_pt.size();
    for (int32_t i = 0; i < np && i < max; i++) {
        const NucPt &q = ctx->nuc_pt[(size_t) i];
        out[i] = mmca_nuc_point{ q.x, q.y, q.z, (uint64_t) q.cx, (uint64_t) q.cy, (uint64_t) q.cz,
                                 q.gx, q.gy, q.gz, q.threshold };
    }
    if (n)
        *n = np;
    return MMCA_OK;
}

int mmca_get_list_slots(mmca_ctx *ctx, int32_t *out, int64_t n)
{
    if (!ctx || !out || n < 0 || n > (int64_t) ctx->c.nx * ctx->c.ny * ctx->c.nz + 2)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (n > 0)
        HIPCHK(hipMemcpy(out, ctx->b.diff_id, (size_t) n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return MMCA_OK;
}

int mmca_set_list_slots(mmca_ctx *ctx, const int32_t *in, int64_t n)
{
    if (!ctx || !in || n < 0 || n > (int64_t) ctx->c.nx * ctx->c.ny * ctx->c.nz + 2)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    const int64_t first = (int64_t) ctx->hc.gindex + 1;
    if (n > first)
        HIPCHK(hipMemcpy(ctx->b.diff_id + first, in + first, (size_t) (n - first) * sizeof(int32_t),
                         hipMemcpyHostToDevice));
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
    std::vector<int> cell(n), src(n), dom(n, 0);
    std::vector<unsigned long long> step(n);
    HIPCHK(hipMemcpy(rot.data(), ctx->b.g_rot, rot.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cell.data(), ctx->b.g_cell, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(src.data(), ctx->b.g_src, n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(step.data(), ctx->b.g_step, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (!ctx->peers.empty())
        HIPCHK(hipMemcpy(dom.data(), ctx->b.g_dom, n * sizeof(int), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(mmca_grain) * (size_t) n);
    for (int id = 1; id < n; id++) {
        mmca_grain &g = out[id];
        // the subblock (rank) that created the grain: its set-up and stream
        const mmca_ctx *owner = ctx->peers.empty() ? ctx : ctx->peers[(size_t) dom[id]];
        const Consts &c = owner->c;
        const mmca_params &p = owner->p;
        if (src[id] == GRAIN_SET && id < (int) owner->ext_grains.size()) {
            g = owner->ext_grains[id];   // record given to mmca_set_grains
            continue;
        }
        for (int q = 0; q < 9; q++)
            g.rotmat[q / 3][q % 3] = rot[(size_t) id * 9 + q];
        g.nuc_timestep = step[id];
        if (src[id] >= 0) {      // nucleate_grain with a FixedNuc/NucVolume entry
            const NucPt &np = owner->nuc_pt[src[id]];
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
            auto it = std::lower_bound(owner->site_idx.begin(), owner->site_idx.end(), idx);
            g.nuc_threshold = (it != owner->site_idx.end() && *it == idx)
                                  ? (double) owner->site_thr[it - owner->site_idx.begin()]
                                  : INFINITY;
            if (k < (int) (owner->rand_angles.size() / 3)) {
                g.rotang[0] = owner->rand_angles[(size_t) k * 3 + 0];
                g.rotang[1] = owner->rand_angles[(size_t) k * 3 + 1];
                g.rotang[2] = owner->rand_angles[(size_t) k * 3 + 2];
            }
        }
    }
    return MMCA_OK;
}

int mmca_rand(mmca_ctx *ctx, int32_t *out, int32_t n)
{
    if (!ctx || n < 0 || (n > 0 && !out))
        return MMCA_ERR_ARG;
    if (!ctx->nucleation_ready)
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    if (!ctx->rand_dirty) {
        // the stream after the orientations consumed so far (3 draws each,
        // grain.c:118-129; none when orientations are not randomised)
        ctx->rng_now = ctx->rng_q0;
        const long long draws = (ctx->oriented ? 3LL : 0LL) * ((long long) ctx->hc.random_used - ctx->q0);
        for (long long i = 0; i < draws; i++)
            (void) ctx->rng_now.next();
        ctx->rand_dirty = true;
    }
    for (int32_t i = 0; i < n; i++)
        out[i] = ctx->rng_now.next();
    return MMCA_OK;
}

}  // extern "C"

namespace {

// Restart image of one block: a header, the counters, the random stream at
// the orientations consumed so far, then the device arrays in a fixed order.
struct StateHeader {
    char magic[8];
    int32_t version, precision;
    int64_t P;
    int32_t nx, ny, nz, nsites;
    int64_t max_total_grains;
    int32_t cur, oriented, n_ext, flags;    // flags: MMCA_STATE_NO_FIELDS (0 in images without it)
    int64_t nledger, n_angles;
};

constexpr char STATE_MAGIC[8] = { 'M', 'M', 'C', 'A', 'S', 'T', 'A', 'T' };

// Restart image of a batch: this header, then the image of every domain in
// order; a coupled batch stores its shared grain table once, in rank 0's.
struct BatchStateHeader {
    char magic[8];
    int32_t version, ndom, coupled, cur;
    int32_t nsub[3], pad_;
};

constexpr char BATCH_MAGIC[8] = { 'M', 'M', 'C', 'A', 'B', 'T', 'C', 'H' };

// the stream right after the orientations consumed so far
GlibcRand stream_at_consumption(const mmca_ctx *ctx)
{
    if (ctx->rand_dirty)
        return ctx->rng_now;
    GlibcRand g = ctx->rng_q0;
    const long long draws = (ctx->oriented ? 3LL : 0LL) * ((long long) ctx->hc.random_used - ctx->q0);
    for (long long i = 0; i < draws; i++)
        (void) g.next();
    return g;
}

// Appends (save) or extracts (load) each piece in the same order.
struct StateIO {
    char *buf;           // null: count only
    const char *in;      // load source
    int64_t pos = 0, cap = 0;
    bool load;
    mmca_ctx *ctx;
    hipStream_t s;
    int host(void *p, size_t n)
    {
        if (load) {
            if (pos + (int64_t) n > cap)
                return MMCA_ERR_ARG;
            memcpy(p, in + pos, n);
        } else if (buf) {
            if (pos + (int64_t) n > cap)
                return MMCA_ERR_ARG;
            memcpy(buf + pos, p, n);
        }
        pos += (int64_t) n;
        return MMCA_OK;
    }
    int dev(void *d, size_t n)
    {
        if (n == 0)
            return MMCA_OK;
        if (load) {
            if (pos + (int64_t) n > cap)
                return MMCA_ERR_ARG;
            HIPCHK(hipMemcpyAsync(d, in + pos, n, hipMemcpyHostToDevice, s));
        } else if (buf) {
            if (pos + (int64_t) n > cap)
                return MMCA_ERR_ARG;
            HIPCHK(hipMemcpyAsync(buf + pos, d, n, hipMemcpyDeviceToHost, s));
        }
        pos += (int64_t) n;
        return MMCA_OK;
    }
};

// Which grain table an image carries: its own (one block), none (a coupled
// rank after the first, whose table is rank 0's), or rank 0's of a coupled
// batch, with the creating rank of every grain.
enum TablePart { TABLE_OWN = 0, TABLE_SHARED = 1, TABLE_COUPLED_OWNER = 2 };

int state_body(mmca_ctx *ctx, StateIO &io, const StateHeader &h, Counters &hc, TablePart table)
{
    DevBufs &b = ctx->b;
    const size_t P = (size_t) h.P;
    const int cur = h.cur;
    TRY(io.host(&hc, sizeof(Counters)));
    GlibcRand g = io.load ? GlibcRand() : stream_at_consumption(ctx);
    TRY(io.host(&g, sizeof(GlibcRand)));
    if (io.load) {
        ctx->rng = g;
        ctx->rng_q0 = g;
        ctx->rng_now = g;
    }
    if (h.precision == MMCA_PREC_FP32DEV) {
        TRY(io.dev(b.ce32, P * sizeof(float)));
        TRY(io.dev(b.cl32[cur], P * sizeof(float)));
        TRY(io.dev(b.ledger, (size_t) (2 * h.nledger) * sizeof(double)));
    } else if (!(h.flags & MMCA_STATE_NO_FIELDS)) {
        TRY(io.dev(b.ce, P * sizeof(double)));
        TRY(io.dev(b.cl[cur], P * sizeof(double)));
    }
    if (!(h.flags & MMCA_STATE_NO_FIELDS)) {
        TRY(io.dev(b.fs[cur], P * sizeof(double)));
        if (io.load) {
            // the kernels index the grain table with gr
            if (io.cap - io.pos < (int64_t) (P * sizeof(int)))
                return MMCA_ERR_ARG;
            TRY(check_grain_ids(io.in + io.pos, P, ctx->p.max_total_grains));
        }
        TRY(io.dev(b.gr, P * sizeof(int)));
        TRY(io.dev(b.dc, 3 * P * sizeof(double)));
    }
    // interface list (slots 0..gindex) and the persistent dc_tmp slots 0..dc_tmp_size;
    // capture writes the cells the list names, so a loaded list must name interior cells
    if (io.load) {
        const int64_t nb = (int64_t) (hc.gindex + 1) * (int64_t) sizeof(int);
        if (io.cap - io.pos < nb)
            return MMCA_ERR_ARG;
        const Consts &c = ctx->c;
        for (int32_t s = 1; s <= hc.gindex; s++) {
            int32_t v;
            memcpy(&v, io.in + io.pos + (int64_t) s * (int64_t) sizeof(int), sizeof(int));
            const int64_t z = v / c.sxy, rem = v - z * c.sxy, y = rem / c.sx, x = rem - y * c.sx;
            if (v < 0 || x < 1 || x > c.nx || y < 1 || y > c.ny || z < 1 || z > c.nz)
                return MMCA_ERR_ARG;
        }
    }
    TRY(io.dev(b.diff_id, (size_t) (hc.gindex + 1) * sizeof(int)));
    const size_t nslot = (size_t) std::max(hc.dc_tmp_size, 0) + 1;
    TRY(io.dev(b.dtx, nslot * sizeof(double)));
    TRY(io.dev(b.dty, nslot * sizeof(double)));
    TRY(io.dev(b.dtz, nslot * sizeof(double)));
    const size_t ng = (size_t) hc.num_grains;
    if (table != TABLE_SHARED) {
        TRY(io.dev(b.g_rot, 9 * ng * sizeof(double)));
        TRY(io.dev(b.g_cell, ng * sizeof(int)));
        TRY(io.dev(b.g_step, ng * sizeof(unsigned long long)));
        TRY(io.dev(b.g_src, ng * sizeof(int)));
    }
    if (table == TABLE_COUPLED_OWNER)
        TRY(io.dev(b.g_dom, ng * sizeof(int)));
    TRY(io.dev(b.site_ls, (size_t) h.nsites * sizeof(int)));
    if (io.load)
        ctx->rand_angles.assign((size_t) (3 * h.n_angles), 0.0);
    TRY(io.host(ctx->rand_angles.data(), (size_t) (3 * h.n_angles) * sizeof(double)));
    if (io.load)
        ctx->ext_grains.assign((size_t) h.n_ext, mmca_grain());
    TRY(io.host(ctx->ext_grains.data(), (size_t) h.n_ext * sizeof(mmca_grain)));
    return MMCA_OK;
}

// Appends the image of one block (header and body) to io.
int save_block(mmca_ctx *ctx, StateIO &io, TablePart table, int flags)
{
    TRY(sync_counters(ctx));
    StateHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, STATE_MAGIC, 8);
    h.version = 1;
    h.precision = ctx->c.f32 ? MMCA_PREC_FP32DEV : MMCA_PREC_FP64;
    h.P = ctx->c.P;
    h.nx = ctx->c.nx;
    h.ny = ctx->c.ny;
    h.nz = ctx->c.nz;
    h.nsites = ctx->b.nsites;
    h.max_total_grains = ctx->p.max_total_grains;
    h.cur = ctx->cur;
    h.oriented = ctx->oriented;
    h.n_ext = (int32_t) ctx->ext_grains.size();
    h.nledger = ctx->c.f32 ? ctx->b.nledger : 0;
    h.n_angles = ctx->hc.random_used;
    // FP32-deviation fields stay in the image: FP64 copies would not give them back exactly
    h.flags = ctx->c.f32 ? 0 : (flags & MMCA_STATE_NO_FIELDS);
    if ((int64_t) ctx->rand_angles.size() < 3 * h.n_angles)
        return MMCA_ERR_STATE;
    Counters hc = ctx->hc;
    io.ctx = ctx;
    io.s = ctx->stream;
    TRY(io.host(&h, sizeof(h)));
    TRY(state_body(ctx, io, h, hc, table));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MMCA_OK;
}

// Restores the image of one block starting at io.pos; its header and counters
// must fit the block's parameters.  *out_hc receives the image's counters.
int load_block(mmca_ctx *ctx, StateIO &io, TablePart table, StateHeader *out_h, Counters *out_hc)
{
    TRY(sync_counters(ctx));
    if (io.cap - io.pos < (int64_t) (sizeof(StateHeader) + sizeof(Counters)))
        return MMCA_ERR_ARG;
    StateHeader h;
    memcpy(&h, io.in + io.pos, sizeof(h));
    if (memcmp(h.magic, STATE_MAGIC, 8) != 0 || h.version != 1 || h.P != ctx->c.P || h.nx != ctx->c.nx ||
        h.ny != ctx->c.ny || h.nz != ctx->c.nz || h.nsites != ctx->b.nsites ||
        h.max_total_grains != ctx->p.max_total_grains || h.oriented != ctx->oriented ||
        h.precision != (ctx->c.f32 ? MMCA_PREC_FP32DEV : MMCA_PREC_FP64) || h.cur < 0 || h.cur > 1 ||
        (ctx->c.f32 && h.nledger != ctx->b.nledger) || h.n_angles < 0 || h.n_angles > ctx->rand_capacity ||
        h.n_ext < 0 || (table != TABLE_OWN && h.n_ext != 0) || (h.flags & ~MMCA_STATE_NO_FIELDS) != 0 ||
        (ctx->c.f32 && h.flags != 0))
        return MMCA_ERR_ARG;
    // the counters come first after the header: check the sizes they imply
    Counters hc;
    memcpy(&hc, io.in + io.pos + sizeof(h), sizeof(Counters));
    if (hc.num_grains < 1 || (int64_t) hc.num_grains > ctx->p.max_total_grains + 1 || hc.gindex < 0 ||
        (int64_t) hc.gindex > (int64_t) ctx->c.nx * ctx->c.ny * ctx->c.nz || hc.dc_tmp_size < 0 ||
        (int64_t) hc.dc_tmp_size > (int64_t) ctx->c.nx * ctx->c.ny * ctx->c.nz + 100 ||
        hc.random_used != h.n_angles)
        return MMCA_ERR_ARG;
    io.ctx = ctx;
    io.s = ctx->stream;
    io.pos += sizeof(h);
    const int st = state_body(ctx, io, h, hc, table);
    if (st != MMCA_OK) {
        // device buffers may be partly overwritten: the block refuses to step from here
        (void) hipStreamSynchronize(ctx->stream);
        ctx->hc.status = MMCA_ERR_STATE;
        (void) push_counters(ctx);
        return st;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->cur = h.cur;
    ctx->q0 = hc.random_used;
    ctx->rand_filled = hc.random_used;
    ctx->rand_dirty = false;
    ctx->have_oce = false;
    ctx->view_dirty = true;
    ctx->hc = hc;
    TRY(push_counters(ctx));
    if (out_h)
        *out_h = h;
    if (out_hc)
        *out_hc = hc;
    return MMCA_OK;
}

}  // namespace

extern "C" {

int mmca_save_state(mmca_ctx *ctx, void *buf, int64_t capacity, int64_t *size)
{
    return mmca_save_state_ex(ctx, 0, buf, capacity, size);
}

int mmca_save_state_ex(mmca_ctx *ctx, int flags, void *buf, int64_t capacity, int64_t *size)
{
    if (!ctx || capacity < 0 || (buf && capacity == 0) || (flags & ~MMCA_STATE_NO_FIELDS) != 0)
        return MMCA_ERR_ARG;
    if (!ctx->peers.empty())
        return MMCA_ERR_STATE;   // a coupled rank: mmca_batch_save_state covers the whole batch
    if (!ctx->nucleation_ready)
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    StateIO io;
    io.buf = (char *) buf;
    io.in = nullptr;
    io.cap = capacity;
    io.load = false;
    TRY(save_block(ctx, io, TABLE_OWN, flags));
    if (size)
        *size = io.pos;
    return MMCA_OK;
}

int mmca_load_state(mmca_ctx *ctx, const void *buf, int64_t size)
{
    if (!ctx || !buf || size < (int64_t) sizeof(StateHeader))
        return MMCA_ERR_ARG;
    if (!ctx->peers.empty())
        return MMCA_ERR_STATE;
    if (!ctx->nucleation_ready)
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    StateIO io;
    io.buf = nullptr;
    io.in = (const char *) buf;
    io.cap = size;
    io.load = true;
    return load_block(ctx, io, TABLE_OWN, nullptr, nullptr);
}

int mmca_io_fields(mmca_ctx *ctx, float *fs_io, float *ce_io, float *temperature_io, int32_t *gr_io)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t P = (size_t) ctx->c.P;
    if (!ctx->io_buf)
        TRY(dev_alloc(&ctx->io_buf, 4 * P));
    float *d_fs = ctx->io_buf, *d_ce = d_fs + P, *d_t = d_ce + P;
    int *d_gr = (int *) (d_t + P);
    if (temperature_io)
        TRY(refresh_beam(ctx, 0, ctx->stream));
    HIPCHK(launch_io_fields(ctx->b, ctx->c, ctx->cur, fs_io ? d_fs : nullptr, ce_io ? d_ce : nullptr,
                            temperature_io ? d_t : nullptr, gr_io ? d_gr : nullptr, ctx->stream));
    if (fs_io)
        HIPCHK(hipMemcpyAsync(fs_io, d_fs, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (ce_io)
        HIPCHK(hipMemcpyAsync(ce_io, d_ce, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (temperature_io)
        HIPCHK(hipMemcpyAsync(temperature_io, d_t, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (gr_io)
        HIPCHK(hipMemcpyAsync(gr_io, d_gr, P * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return check_status(ctx);
}

int mmca_set_grains(mmca_ctx *ctx, const mmca_grain *grains, int32_t num_grains)
{
    if (!ctx || !grains || num_grains < 1 || (int64_t) num_grains > ctx->p.max_total_grains + 1)
        return MMCA_ERR_ARG;
    if (!ctx->peers.empty())
        return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(ctx->device));
    TRY(sync_counters(ctx));
    const size_t n = (size_t) num_grains;
    std::vector<double> rot(n * 9);
    std::vector<int> src(n, GRAIN_SET), cell(n, -1);
    std::vector<unsigned long long> step(n);
    for (size_t id = 0; id < n; id++) {
        for (int q = 0; q < 9; q++)
            rot[id * 9 + q] = grains[id].rotmat[q / 3][q % 3];
        step[id] = grains[id].nuc_timestep;
    }
    HIPCHK(hipMemcpyAsync(ctx->b.g_rot, rot.data(), rot.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->b.g_src, src.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->b.g_cell, cell.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->b.g_step, step.data(), n * sizeof(unsigned long long), hipMemcpyHostToDevice,
                          ctx->stream));
    ctx->ext_grains.assign(grains, grains + n);
    ctx->hc.num_grains = num_grains;
    return push_counters(ctx);
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

int mmca_set_tuning(mmca_ctx *ctx, int key, int value)
{
    if (!ctx)
        return MMCA_ERR_ARG;
    switch (key) {
    case MMCA_TUNE_ZM_PLANES:
        if (value < 0 || value > 1024)
            return MMCA_ERR_ARG;   // the LDS-staged kernel keeps one temperature per plane of a tile
        ctx->c.zm_planes = value;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_ZM_ROWS:
        if (value != 4 && value != 8 && value != 16)
            return MMCA_ERR_ARG;
        ctx->c.zm_by = value;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_XCD_REMAP:
        ctx->c.xcd_remap = value ? 1 : 0;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_ZM_LDS:
        ctx->c.zm_lds = value ? 1 : 0;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_CAPTURE_CHUNK:
        if (value < 0 || value > (1 << 24))
            return MMCA_ERR_ARG;
        ctx->c.capture_chunk = (value + 255) / 256 * 256;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_CAPTURE_BAND:
        if (value < 0 || value > (1 << 24))
            return MMCA_ERR_ARG;
        ctx->c.capture_band = value;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_TRACK_D:
        if (value)
            TRY(ensure_d(ctx));
        ctx->c.track_d = value ? 1 : 0;
        ctx->view_dirty = true;
        return MMCA_OK;
    case MMCA_TUNE_BEAM_SEP:
        if (value && !(ctx->c.am && ctx->c.f32))
            return MMCA_ERR_ARG;   // FP64 storage keeps the reference's quadrature
        if (value)
            TRY(ensure_beam(ctx));
        ctx->c.beam_sep = value ? 1 : 0;
        ctx->view_dirty = true;
        return MMCA_OK;
    default:
        return MMCA_ERR_ARG;
    }
}

int mmca_get_tuning(mmca_ctx *ctx, int key, int *value)
{
    if (!ctx || !value)
        return MMCA_ERR_ARG;
    switch (key) {
    case MMCA_TUNE_ZM_PLANES:
        *value = ctx->c.zm_planes;
        return MMCA_OK;
    case MMCA_TUNE_ZM_ROWS:
        *value = ctx->c.zm_by;
        return MMCA_OK;
    case MMCA_TUNE_XCD_REMAP:
        *value = ctx->c.xcd_remap;
        return MMCA_OK;
    case MMCA_TUNE_ZM_LDS:
        *value = ctx->c.zm_lds;
        return MMCA_OK;
    case MMCA_TUNE_CAPTURE_CHUNK:
        *value = ctx->c.capture_chunk;
        return MMCA_OK;
    case MMCA_TUNE_CAPTURE_BAND:
        *value = ctx->c.capture_band;
        return MMCA_OK;
    case MMCA_TUNE_TRACK_D:
        *value = ctx->c.track_d;
        return MMCA_OK;
    case MMCA_TUNE_BEAM_SEP:
        *value = ctx->c.beam_sep;
        return MMCA_OK;
    default:
        return MMCA_ERR_ARG;
    }
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

// ===========================================================================
// Batches: independent subblocks (domains) advanced together, each launch of
// a step covering every domain.
struct mmca_batch {
    std::vector<mmca_ctx *> dom;
    int device = 0;
    hipStream_t stream = nullptr;
    DomainDesc *d_desc = nullptr;
    long long *d_off = nullptr;
    int cap = 0;
    View view;
    bool view_dirty = true;
    int cur = 0;
    int zm_planes = 8, zm_by = 4, xcd_remap = 0, zm_lds = 1, capture_chunk = 0, capture_band = 0, track_d = 0;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double time_ms = 0.0;
    long long time_n = 0;
    bool use_graph = false;
    hipGraphExec_t graph[2] = { nullptr, nullptr };
    // the domains are the subblocks of one global domain (mmca_batch_couple)
    bool coupled = false;
    long long max_face = 0;      // largest face (cells) or halo-row word count
};

namespace {

int batch_view(mmca_batch *bt)
{
    if (!bt->view_dirty)
        return MMCA_OK;
    for (mmca_ctx *c : bt->dom) {
        c->c.zm_planes = bt->zm_planes;
        c->c.zm_by = bt->zm_by;
        c->c.xcd_remap = bt->xcd_remap;
        c->c.zm_lds = bt->zm_lds;
        c->c.capture_chunk = bt->capture_chunk;
        c->c.capture_band = bt->capture_band;
        c->c.track_d = bt->track_d;
        c->view_dirty = true;
    }
    TRY(build_view(bt->dom.data(), (int) bt->dom.size(), bt->xcd_remap, bt->zm_planes > 0 ? bt->zm_by : 0,
                   &bt->d_desc, &bt->d_off, &bt->cap, &bt->view, bt->stream));
    for (int p = 0; p < 2; p++)
        if (bt->graph[p]) {
            (void) hipGraphExecDestroy(bt->graph[p]);
            bt->graph[p] = nullptr;
        }
    bt->view_dirty = false;
    return MMCA_OK;
}

int batch_sync(mmca_batch *bt)
{
    for (mmca_ctx *c : bt->dom)
        HIPCHK(hipMemcpyAsync(&c->hc, c->b.cnt, sizeof(Counters), hipMemcpyDeviceToHost, bt->stream));
    HIPCHK(hipStreamSynchronize(bt->stream));
    return MMCA_OK;
}

// one doiteration() of every domain: seven launches whatever the batch size
// (ten for a coupled batch: two halo exchanges and the cross-rank activation)
int batch_enqueue(mmca_batch *bt, int cur)
{
    const Consts &c0 = bt->dom[0]->c;
    const int nxt = cur ^ 1;
    for (mmca_ctx *c : bt->dom)
        TRY(refresh_beam(c, 1, bt->stream));
    if (bt->coupled) {
        HIPCHK(launch_halo_exchange(bt->view, 0, cur, bt->max_face, bt->stream));
        HIPCHK(launch_nucleate_coupled(bt->view, cur, bt->stream));
    } else
        HIPCHK(launch_nucleate(bt->view, 7, cur, bt->stream));
    const size_t e0 = bt->ev_used;
    if (bt->timing) {
        while (bt->ev.size() < e0 + 2) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            bt->ev.push_back(e);
        }
        HIPCHK(hipEventRecord(bt->ev[e0], bt->stream));
    }
    HIPCHK(launch_solute(bt->view, c0.am && !c0.beam_sep, c0.f32, bt->zm_planes > 0 ? bt->zm_by : 0, bt->zm_lds, 1,
                         1, 0, cur, nxt, bt->stream));
    if (bt->timing) {
        HIPCHK(hipEventRecord(bt->ev[e0 + 1], bt->stream));
        bt->ev_used = e0 + 2;
    }
    if (bt->coupled)
        HIPCHK(launch_halo_exchange(bt->view, 1, nxt, bt->max_face, bt->stream));
    HIPCHK(launch_list(bt->view, nxt, 1, 1, bt->stream));
    return MMCA_OK;
}

// a pair of steps starting from buffer `cur`, captured once into a graph
int batch_graph_pair(mmca_batch *bt, int cur)
{
    if (!bt->graph[cur]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
        int st = batch_enqueue(bt, cur);
        if (st == MMCA_OK)
            st = batch_enqueue(bt, cur ^ 1);
        hipError_t e = hipStreamEndCapture(bt->stream, &g);
        if (st != MMCA_OK)
            return st;
        HIPCHK(e);
        HIPCHK(hipGraphInstantiate(&bt->graph[cur], g, nullptr, nullptr, 0));
        (void) hipGraphDestroy(g);
    }
    HIPCHK(hipGraphLaunch(bt->graph[cur], bt->stream));
    return MMCA_OK;
}

}  // namespace

extern "C" {

int mmca_batch_create(const mmca_params *params, const mmca_nucleation_cfg *nucleation, int ndomains,
                      int device, mmca_batch **out)
{
    if (!params || !out || ndomains < 1)
        return MMCA_ERR_ARG;
    *out = nullptr;
    for (int d = 1; d < ndomains; d++)
        if (params[d].precision != params[0].precision ||
            (params[d].temp_model == MMCA_TEMP_AM_BEAM) != (params[0].temp_model == MMCA_TEMP_AM_BEAM))
            return MMCA_ERR_ARG;   // one kernel instantiation serves the whole batch
    mmca_batch *bt = new mmca_batch();
    bt->device = device;
    *out = bt;
    for (int d = 0; d < ndomains; d++) {
        mmca_ctx *c = nullptr;
        int st = mmca_create(&params[d], device, &c);
        if (c)
            bt->dom.push_back(c);
        if (st != MMCA_OK)
            return st;
        if (nucleation)
            TRY(mmca_init_nucleation(c, &nucleation[d]));
    }
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreate(&bt->stream));
    for (mmca_ctx *c : bt->dom)
        HIPCHK(hipStreamSynchronize(c->stream));
    return MMCA_OK;
}

int mmca_batch_couple(mmca_batch *bt, const int32_t nsub[3])
{
    if (!bt || !nsub || nsub[0] < 1 || nsub[1] < 1 || nsub[2] < 1)
        return MMCA_ERR_ARG;
    const int n = (int) bt->dom.size();
    if ((long long) nsub[0] * nsub[1] * nsub[2] != n || bt->coupled)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(bt->device));
    TRY(batch_sync(bt));
    const mmca_params &p0 = bt->dom[0]->p;
    const int n3[3] = { p0.nx, p0.ny, p0.nz };
    for (int d = 0; d < n; d++) {
        const mmca_ctx *c = bt->dom[d];
        const mmca_params &p = c->p;
        // rank d of an MPI Cart grid ordered (z, y, x) (ca_main.c:197-212)
        const int rc[3] = { d % nsub[0], (d / nsub[0]) % nsub[1], d / (nsub[0] * nsub[1]) };
        for (int a = 0; a < 3; a++) {
            const int na = a == 0 ? p.nx : a == 1 ? p.ny : p.nz;
            if (na != n3[a] || p.sb_coords[a] != rc[a] || p.gdim[a] != nsub[a] * n3[a])
                return MMCA_ERR_ARG;
            if (!p.pad_bndy[a] != !p0.pad_bndy[a])
                return MMCA_ERR_ARG;     // one FaceCtrl for the whole domain
        }
        if (p.precision != p0.precision || p.max_total_grains != p0.max_total_grains ||
            c->hc.num_grains != 1 || c->hc.timestep != bt->dom[0]->hc.timestep || !c->nucleation_ready)
            return MMCA_ERR_ARG;     // fresh, initialised subblocks of one run
    }
    long long mf = 0;
    for (int d = 0; d < n; d++) {
        mmca_ctx *c = bt->dom[d];
        for (int a = 0; a < 3; a++)
            c->c.nsub[a] = nsub[a];
        mf = std::max({ mf, (long long) c->c.ny * c->c.nz, (long long) c->c.nx * c->c.nz,
                        (long long) c->c.nx * c->c.ny, (long long) c->c.nz * c->c.w64,
                        (long long) c->c.ny * c->c.w64 });
        c->peers = bt->dom;
        if (d > 0) {
            // one grain table for all ranks (the reference replicates it)
            DevBufs &b = c->b;
            void *own[] = { b.g_rot, b.g_cell, b.g_step, b.g_src, b.g_dom };
            for (void *q : own)
                HIPCHK(hipFree(q));
            const DevBufs &b0 = bt->dom[0]->b;
            b.g_rot = b0.g_rot;
            b.g_cell = b0.g_cell;
            b.g_step = b0.g_step;
            b.g_src = b0.g_src;
            b.g_dom = b0.g_dom;
            c->shares_grains = true;
        }
        c->view_dirty = true;
    }
    bt->max_face = mf;
    bt->coupled = true;
    bt->view_dirty = true;
    return MMCA_OK;
}

int mmca_batch_save_state(mmca_batch *bt, void *buf, int64_t capacity, int64_t *size)
{
    return mmca_batch_save_state_ex(bt, 0, buf, capacity, size);
}

int mmca_batch_save_state_ex(mmca_batch *bt, int flags, void *buf, int64_t capacity, int64_t *size)
{
    if (!bt || capacity < 0 || (buf && capacity == 0) || (flags & ~MMCA_STATE_NO_FIELDS) != 0)
        return MMCA_ERR_ARG;
    for (mmca_ctx *c : bt->dom)
        if (!c->nucleation_ready)
            return MMCA_ERR_STATE;
    HIPCHK(hipSetDevice(bt->device));
    TRY(batch_sync(bt));
    // the reference's error() ends every rank: a coupled batch that stopped
    // on an error has no state to continue from
    if (bt->coupled)
        for (mmca_ctx *c : bt->dom)
            if (c->hc.status != MMCA_OK)
                return MMCA_ERR_STATE;
    BatchStateHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, BATCH_MAGIC, 8);
    h.version = 1;
    h.ndom = (int32_t) bt->dom.size();
    h.coupled = bt->coupled ? 1 : 0;
    h.cur = bt->cur;
    for (int a = 0; a < 3; a++)
        h.nsub[a] = bt->coupled ? bt->dom[0]->c.nsub[a] : 0;
    StateIO io;
    io.buf = (char *) buf;
    io.in = nullptr;
    io.cap = capacity;
    io.load = false;
    TRY(io.host(&h, sizeof(h)));
    for (size_t d = 0; d < bt->dom.size(); d++) {
        const TablePart t = !bt->coupled ? TABLE_OWN : d == 0 ? TABLE_COUPLED_OWNER : TABLE_SHARED;
        TRY(save_block(bt->dom[d], io, t, flags));
    }
    if (size)
        *size = io.pos;
    return MMCA_OK;
}

int mmca_batch_load_state(mmca_batch *bt, const void *buf, int64_t size)
{
    if (!bt || !buf || size < (int64_t) sizeof(BatchStateHeader))
        return MMCA_ERR_ARG;
    for (mmca_ctx *c : bt->dom)
        if (!c->nucleation_ready)
            return MMCA_ERR_STATE;
    BatchStateHeader h;
    memcpy(&h, buf, sizeof(h));
    if (memcmp(h.magic, BATCH_MAGIC, 8) != 0 || h.version != 1 || h.ndom != (int32_t) bt->dom.size() ||
        h.coupled != (bt->coupled ? 1 : 0) || h.cur < 0 || h.cur > 1)
        return MMCA_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (h.nsub[a] != (bt->coupled ? bt->dom[0]->c.nsub[a] : 0))
            return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(bt->device));
    HIPCHK(hipStreamSynchronize(bt->stream));
    StateIO io;
    io.buf = nullptr;
    io.in = (const char *) buf;
    io.cap = size;
    io.load = true;
    io.pos = sizeof(h);
    Counters first;
    for (size_t d = 0; d < bt->dom.size(); d++) {
        const TablePart t = !bt->coupled ? TABLE_OWN : d == 0 ? TABLE_COUPLED_OWNER : TABLE_SHARED;
        StateHeader bh;
        Counters hc;
        TRY(load_block(bt->dom[d], io, t, &bh, &hc));
        if (d == 0)
            first = hc;
        // the domains of a batch step together; coupled ranks share one grain count
        if (bh.cur != h.cur || hc.timestep != first.timestep ||
            (bt->coupled && hc.num_grains != first.num_grains))
            return MMCA_ERR_ARG;
    }
    bt->cur = h.cur;
    bt->view_dirty = true;
    return MMCA_OK;
}

int mmca_batch_destroy(mmca_batch *bt)
{
    if (!bt)
        return MMCA_OK;
    (void) hipSetDevice(bt->device);
    if (bt->stream)
        (void) hipStreamSynchronize(bt->stream);
    for (int p = 0; p < 2; p++)
        if (bt->graph[p])
            (void) hipGraphExecDestroy(bt->graph[p]);
    for (auto e : bt->ev)
        (void) hipEventDestroy(e);
    if (bt->d_desc)
        (void) hipFree(bt->d_desc);
    if (bt->d_off)
        (void) hipFree(bt->d_off);
    for (mmca_ctx *c : bt->dom)
        mmca_destroy(c);
    if (bt->stream)
        (void) hipStreamDestroy(bt->stream);
    delete bt;
    return MMCA_OK;
}

int mmca_batch_size(mmca_batch *bt, int *ndomains)
{
    if (!bt || !ndomains)
        return MMCA_ERR_ARG;
    *ndomains = (int) bt->dom.size();
    return MMCA_OK;
}

int mmca_batch_domain(mmca_batch *bt, int i, mmca_ctx **ctx)
{
    if (!bt || !ctx || i < 0 || i >= (int) bt->dom.size())
        return MMCA_ERR_ARG;
    *ctx = bt->dom[i];
    return MMCA_OK;
}

int mmca_batch_set_tuning(mmca_batch *bt, int key, int value)
{
    if (!bt)
        return MMCA_ERR_ARG;
    switch (key) {
    case MMCA_TUNE_ZM_PLANES:
        if (value < 0 || value > 1024)
            return MMCA_ERR_ARG;
        bt->zm_planes = value;
        break;
    case MMCA_TUNE_ZM_ROWS:
        if (value != 4 && value != 8 && value != 16)
            return MMCA_ERR_ARG;
        bt->zm_by = value;
        break;
    case MMCA_TUNE_XCD_REMAP:
        bt->xcd_remap = value ? 1 : 0;
        break;
    case MMCA_TUNE_ZM_LDS:
        bt->zm_lds = value ? 1 : 0;
        break;
    case MMCA_TUNE_CAPTURE_CHUNK:
        if (value < 0 || value > (1 << 24))
            return MMCA_ERR_ARG;
        bt->capture_chunk = (value + 255) / 256 * 256;
        break;
    case MMCA_TUNE_CAPTURE_BAND:
        if (value < 0 || value > (1 << 24))
            return MMCA_ERR_ARG;
        bt->capture_band = value;
        break;
    case MMCA_TUNE_TRACK_D:
        for (mmca_ctx *d : bt->dom)
            TRY(mmca_set_tuning(d, MMCA_TUNE_TRACK_D, value));
        bt->track_d = value ? 1 : 0;
        break;
    case MMCA_TUNE_BEAM_SEP:
        for (mmca_ctx *d : bt->dom)
            TRY(mmca_set_tuning(d, MMCA_TUNE_BEAM_SEP, value));
        break;
    case MMCA_TUNE_GRAPH:
        bt->use_graph = value != 0;
        return MMCA_OK;
    default:
        return MMCA_ERR_ARG;
    }
    bt->view_dirty = true;
    return MMCA_OK;
}

int mmca_batch_set_kernel_timing(mmca_batch *bt, int enable)
{
    if (!bt)
        return MMCA_ERR_ARG;
    bt->timing = enable != 0;
    return MMCA_OK;
}

int mmca_batch_kernel_time_ms(mmca_batch *bt, double *avg_ms, int64_t *launches)
{
    if (!bt)
        return MMCA_ERR_ARG;
    if (avg_ms)
        *avg_ms = bt->time_ms;
    if (launches)
        *launches = bt->time_n;
    return MMCA_OK;
}

int mmca_batch_step(mmca_batch *bt, int nsteps, int want_sums, mmca_report *reports)
{
    if (!bt || nsteps < 0)
        return MMCA_ERR_ARG;
    HIPCHK(hipSetDevice(bt->device));
    TRY(batch_view(bt));
    TRY(batch_sync(bt));
    for (mmca_ctx *c : bt->dom) {
        TRY(replan_random(c));
        c->hc.new_grains = 0;
        HIPCHK(hipMemcpyAsync(&c->b.cnt->new_grains, &c->hc.new_grains, sizeof(int), hipMemcpyHostToDevice,
                              bt->stream));
        int st = courant_check(c);
        if (st != MMCA_OK && nsteps > 0 && c->hc.status == MMCA_OK) {
            c->hc.status = st;
            c->hc.failed_step = c->hc.timestep + 1;
            HIPCHK(hipMemcpyAsync(c->b.cnt, &c->hc, sizeof(Counters), hipMemcpyHostToDevice, bt->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(bt->stream));
    // a domain that stops on an error keeps the buffers of its last step
    // (its kernels skip every later step while the batch's parity turns)
    std::vector<std::pair<int, uint64_t>> start(bt->dom.size());
    for (size_t d = 0; d < bt->dom.size(); d++)
        start[d] = { bt->dom[d]->cur, bt->dom[d]->hc.timestep };
    int done = 0;
    while (done < nsteps) {
        // chunk bound: drawn orientations cover every grain the chunk can create
        long long k = nsteps - done;
        for (mmca_ctx *c : bt->dom) {
            if (c->b.nsites == 0 || c->rand_filled >= c->rand_capacity || c->hc.status != MMCA_OK)
                continue;
            long long spare = (long long) c->rand_filled - c->hc.random_used;
            if (spare < (long long) c->b.nsites * 16) {
                long long want = (long long) c->hc.random_used +
                                 std::max((long long) c->b.nsites * 256, 2 * spare + 65536);
                TRY(ensure_random(c, want));
                spare = (long long) c->rand_filled - c->hc.random_used;
            }
            if (c->rand_filled < c->rand_capacity)
                k = std::min(k, spare / c->b.nsites);
        }
        if (k < 1)
            return MMCA_ERR_STATE;
        long long n = 0;
        if (bt->use_graph && !bt->timing)
            for (; n + 2 <= k; n += 2)
                TRY(batch_graph_pair(bt, bt->cur));   // two steps: cur returns to itself
        for (; n < k; n++) {
            TRY(batch_enqueue(bt, bt->cur));
            bt->cur ^= 1;
        }
        done += (int) k;
        if (done < nsteps)
            TRY(batch_sync(bt));
    }
    TRY(batch_sync(bt));
    for (size_t d = 0; d < bt->dom.size(); d++) {
        mmca_ctx *c = bt->dom[d];
        c->cur = c->hc.status == MMCA_OK ? bt->cur : start[d].first ^ (int) ((c->hc.timestep - start[d].second) & 1);
        if (want_sums)
            HIPCHK(launch_reduce(c->b, c->c, c->cur, c->d_red, bt->stream));
    }
    TRY(batch_sync(bt));
    if (bt->timing) {
        double tot = 0.0;
        long long n = 0;
        for (size_t i = 0; i + 1 < bt->ev_used; i += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, bt->ev[i], bt->ev[i + 1]));
            tot += ms;
            n++;
        }
        bt->ev_used = 0;
        bt->time_ms = n ? tot / n : 0.0;
        bt->time_n = n;
    }
    int worst = MMCA_OK;
    for (size_t d = 0; d < bt->dom.size(); d++) {
        mmca_ctx *c = bt->dom[d];
        if (c->hc.status != MMCA_OK && worst == MMCA_OK)
            worst = c->hc.status;
        if (!reports)
            continue;
        mmca_report *r = &reports[d];
        memset(r, 0, sizeof(*r));
        r->timestep = c->hc.timestep;
        r->num_grains = c->hc.num_grains;
        r->new_grains = c->hc.new_grains;
        r->gindex = c->hc.gindex;
        r->status = c->hc.status;
        r->failed_step = c->hc.failed_step;
        if (want_sums && c->hc.status == MMCA_OK) {
            double s2[2];
            HIPCHK(hipMemcpy(s2, c->d_red, sizeof(s2), hipMemcpyDeviceToHost));
            r->solid_sum = s2[0];
            r->solute_sum = s2[1];
            if (c->c.f32) {
                r->solute_sum += (double) c->c.nx * c->c.ny * c->c.nz * c->c.cinit;
                TRY(read_ledger(c, &r->solute_ledger));
            }
        }
    }
    return worst;
}

int mmca_batch_autotune(const mmca_params *params, const mmca_nucleation_cfg *nucleation, int ndomains,
                        int device, int warmup, int trial_steps, int *best_planes, int *best_rows,
                        double *best_ms)
{
    if (!params || ndomains < 1 || trial_steps < 1 || warmup < 0)
        return MMCA_ERR_ARG;
    // solute-kernel shapes tried: (planes per workgroup, rows); planes 0 is
    // the cell-parallel kernel
    static const int cand[][2] = { { 0, 4 }, { 8, 4 }, { 16, 4 }, { 8, 8 }, { 16, 8 }, { 32, 8 },
                                   { 64, 8 }, { 16, 16 }, { 32, 16 }, { 64, 16 } };
    mmca_batch *bt = nullptr;
    int st = mmca_batch_create(params, nucleation, ndomains, device, &bt);
    if (st != MMCA_OK) {
        mmca_batch_destroy(bt);
        return st;
    }
    int bp = 16, br = 8;
    double bms = 1e300;
    if (warmup > 0)
        st = mmca_batch_step(bt, warmup, 0, nullptr);
    for (const auto &cd : cand) {
        if (st != MMCA_OK)
            break;
        mmca_batch_set_tuning(bt, MMCA_TUNE_ZM_PLANES, cd[0]);
        mmca_batch_set_tuning(bt, MMCA_TUNE_ZM_ROWS, cd[1]);
        mmca_batch_set_kernel_timing(bt, 1);
        st = mmca_batch_step(bt, trial_steps, 0, nullptr);
        double ms = 0;
        mmca_batch_kernel_time_ms(bt, &ms, nullptr);
        if (st == MMCA_OK && ms > 0 && ms < bms) {
            bms = ms;
            bp = cd[0];
            br = cd[1];
        }
    }
    mmca_batch_destroy(bt);
    if (best_planes)
        *best_planes = bp;
    if (best_rows)
        *best_rows = br;
    if (best_ms)
        *best_ms = bms;
    // a failing domain (e.g. a negative ce) still leaves valid timings
    return (bms < 1e300) ? MMCA_OK : st;
}

}  // extern "C"
