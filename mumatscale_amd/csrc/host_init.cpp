// Host-only part of the CA step library (see host_init.h).
#include "host_init.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <math.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif

#define TRY(x)                 \
    do {                       \
        int st_ = (x);         \
        if (st_ != MMCA_OK)    \
            return st_;        \
    } while (0)

namespace mmca {

void GlibcRand::seed(unsigned int s)
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

int32_t GlibcRand::next()
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

float GlibcRand::scale() { return (float) next() / (float) 2147483647; }

float GlibcRand::gaussian()
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
    // the reference is C: log() of the float s is the double log (in C++
    // log(float) would be logf)
    float multiplier = sqrtf(-2.0 * log((double) s) / s);
    float val = n0 * multiplier;
    gauss_next = n1 * multiplier;
    gauss_have = true;
    return val;
}

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

void draw_orientations(GlibcRand &rng, int oriented, int n, double *angles3, double *rot9)
{
    for (int k = 0; k < n; k++) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (oriented) {
            float acon = 4.0;
            gx = M_PI / acon - 2 * (M_PI / acon * rng.scale());
            gy = M_PI / acon - 2 * (M_PI / acon * rng.scale());
            gz = M_PI / acon - 2 * (M_PI / acon * rng.scale());
        }
        angles3[3 * k + 0] = gx;
        angles3[3 * k + 1] = gy;
        angles3[3 * k + 2] = gz;
        rotmat_of(gx, gy, gz, rot9 + 9 * (size_t) k);
    }
}

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
    c.w64 = (c.nx + 63) / 64;          // mask words per row (interior x)
    c.zm_planes = 64;                   // z-marching solute kernel by default (FP32 storage:
    c.zm_by = 16;                       // shorter tiles, below)
    c.xcd_remap = 1;
    c.zm_lds = 1;                       // planes staged in LDS (k_solute_zl)
    c.capture_chunk = 0;
    c.capture_band = 0;
    c.track_d = 0;
    if (p.precision != MMCA_PREC_FP64 && p.precision != MMCA_PREC_FP32DEV)
        return MMCA_ERR_ARG;
    c.f32 = (p.precision == MMCA_PREC_FP32DEV) ? 1 : 0;
    if (c.f32) {
        // FP32 storage streams faster in short tiles of 64 x 4 cells x 8 planes
        // (several workgroups per CU); FP64 storage runs as fast or faster in
        // the long 16-row tiles over a long run
        c.zm_planes = 8;
        c.zm_by = 4;
        c.xcd_remap = 0;
    }
    const int n3[3] = { p.nx, p.ny, p.nz };
    for (int a = 0; a < 3; a++) {
        // a face is a mold wall where the global domain ends (one subblock:
        // both faces of a padded axis)
        c.pad_lo[a] = (p.pad_bndy[a] && p.sb_coords[a] == 0) ? 1 : 0;
        c.pad_hi[a] = (p.pad_bndy[a] && (long long) (p.sb_coords[a] + 1) * n3[a] >= p.gdim[a]) ? 1 : 0;
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
    c.beam_sep = (c.am && c.f32) ? 1 : 0;   // FP64 storage keeps the reference's quadrature
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


// init_nucleation_sites (grain.c:603-713) + init_sb_nucleation (grain.c:931-1010)
int plan_nucleation(const mmca_params &params, const Consts &c, const mmca_nucleation_cfg *cfg,
                    uint64_t timestep, GlibcRand &rng, SitePlan &out)
{
    if (!cfg)
        return MMCA_ERR_ARG;
    if ((cfg->n_fixed > 0 && !cfg->fixed) || (cfg->n_volumes > 0 && !cfg->volumes))
        return MMCA_ERR_ARG;
    const mmca_params &p = params;
    // init_nucleation_sites runs on rank 0, after srand(0 + 1 + seed)
    // (ca_main.c:215, 244-245); each rank then samples its own subblock
    // (init_sb_nucleation) from its own stream, srand(rank + 1 + seed)
    rng.seed((unsigned int) (0 + 1 + cfg->random_seed));
    const int oriented = cfg->oriented;
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
        if (oriented) {
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
    if (cfg->rank != 0)
        rng.seed((unsigned int) (cfg->rank + 1 + cfg->random_seed));
    // points of this subblock (grain.c:695-711)
    out.nuc_pt.clear();
    for (const NucPt &n : g) {
        size_t ix, iy, iz;
        real2cell(n.x, n.y, n.z, &ix, &iy, &iz);
        if ((int) (ix / p.nx) == p.sb_coords[0] && (int) (iy / p.ny) == p.sb_coords[1] &&
            (int) (iz / p.nz) == p.sb_coords[2])
            out.nuc_pt.push_back(n);
    }

    // init_sb_nucleation: dense thresholds while sampling
    const long long P = c.P;
    std::vector<float> thr((size_t) P, INFINITY);
    auto is_mold = [&](long x, long y, long z) {
        return (c.pad_lo[0] && x == 0) || (c.pad_hi[0] && x == c.nx + 1) || (c.pad_lo[1] && y == 0) ||
               (c.pad_hi[1] && y == c.ny + 1) || (c.pad_lo[2] && z == 0) || (c.pad_hi[2] && z == c.nz + 1);
    };
    for (const NucPt &n : out.nuc_pt) {
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
    out.site_idx.clear();
    out.site_thr.clear();
    out.site_fixed.clear();
    std::vector<double> &fixed_rot = out.fixed_rot;
    fixed_rot.assign(out.nuc_pt.size() * 9 + 9, 0.0);
    std::vector<int> &site_ls = out.site_ls;
    site_ls.clear();
    for (size_t j = 0; j < out.nuc_pt.size(); j++)
        rotmat_of(out.nuc_pt[j].gx, out.nuc_pt[j].gy, out.nuc_pt[j].gz, &fixed_rot[j * 9]);
    const double t0 = timestep * c.dt;
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
        for (size_t j = 0; j < out.nuc_pt.size(); j++)
            if (out.nuc_pt[j].cx == gx && out.nuc_pt[j].cy == gy && out.nuc_pt[j].cz == gz) {
                fixed = (int) j;
                break;
            }
        out.site_idx.push_back((int) i);
        out.site_thr.push_back(thr[(size_t) i]);
        out.site_fixed.push_back(fixed);
        int ls = 0;
        if (c.am)   // tempPrepareSB_int (temperature.c:183-189)
            ls = host_temp(c, t0, x, y, z) >= c.Tliq ? 1 : 0;
        site_ls.push_back(ls);
    }
    return MMCA_OK;
}

}  // namespace mmca

// ---------------------------------------------------------------------------
// Device-free entry points of the host set-up (C ABI).
extern "C" {

int mmca_nucleation_sites(const mmca_params *params, const mmca_nucleation_cfg *cfg,
                          int32_t *idx_out, float *thr_out, int32_t *fixed_out, int64_t max,
                          int64_t *n_sites)
{
    using namespace mmca;
    if (!params || !cfg)
        return MMCA_ERR_ARG;
    Consts c;
    int st = build_consts(*params, c);
    if (st != MMCA_OK)
        return st;
    GlibcRand rng;
    SitePlan plan;
    st = plan_nucleation(*params, c, cfg, params->timestep, rng, plan);
    if (st != MMCA_OK)
        return st;
    const int64_t n = (int64_t) plan.site_idx.size();
    if (n_sites)
        *n_sites = n;
    for (int64_t i = 0; i < n && i < max; i++) {
        if (idx_out)
            idx_out[i] = plan.site_idx[i];
        if (thr_out)
            thr_out[i] = plan.site_thr[i];
        if (fixed_out)
            fixed_out[i] = plan.site_fixed[i];
    }
    return MMCA_OK;
}

int mmca_orientation_draws(const mmca_params *params, const mmca_nucleation_cfg *cfg, int32_t n,
                           double *angles_out, double *rotmat_out)
{
    using namespace mmca;
    if (!params || !cfg || n < 0)
        return MMCA_ERR_ARG;
    Consts c;
    int st = build_consts(*params, c);
    if (st != MMCA_OK)
        return st;
    GlibcRand rng;
    SitePlan plan;
    st = plan_nucleation(*params, c, cfg, params->timestep, rng, plan);
    if (st != MMCA_OK)
        return st;
    std::vector<double> a((size_t) n * 3 + 3), r((size_t) n * 9 + 9);
    draw_orientations(rng, cfg->oriented, n, a.data(), r.data());
    if (angles_out)
        memcpy(angles_out, a.data(), (size_t) n * 3 * sizeof(double));
    if (rotmat_out)
        memcpy(rotmat_out, r.data(), (size_t) n * 9 * sizeof(double));
    return MMCA_OK;
}

}  // extern "C"
