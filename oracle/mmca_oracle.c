/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Scalar CPU restatement of muMatScale's per-timestep CA core, used by the
 * tests, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the
 * checker.  The product library (mumatscale_amd/csrc) never links, loads or
 * calls this file.
 *
 * Every function follows one reference function; the reference file:line is
 * given at each.  Arithmetic is written in the reference's operation order
 * and built with -ffp-contract=off so that results are bit-identical with
 * the reference CPU build (gcc, default x86-64 -march, no FMA).
 *
 * Pinning (what checks each part of this file):
 *   - rng_*:          glibc rand() in the same process (test_oracle_rng.py)
 *   - whole step,
 *     nucleation init: per-step FNV-1a hashes and full padded fields written
 *                      by the reference itself (oracle/_ref, fixtures under
 *                      tests/golden/), bit for bit.
 *   - diffusion:       mass conservation, uniform fixed point, 1-D closed form
 *   - fs_change:       per-cell closed forms of each branch
 *   - capture corner:  truth table of the 5-test chain
 *
 * Loops whose cells are independent (temperature, diffusion, fs_change,
 * octahedra, capture over list slots) run with OpenMP; list building,
 * nucleation order and sums stay serial in raster order, so results do not
 * depend on the thread count.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mmca.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#define O_MIN(x, y) (((x) < (y)) ? (x) : (y))
#define O_MAX(x, y) (((x) > (y)) ? (x) : (y))

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif

/* ------------------------------------------------------------------------ */
/* glibc rand(): TYPE_3 additive feedback generator, x[i] = x[i-3] + x[i-31]
 * (published algorithm of glibc random_r/srandom_r).  The reference calls
 * srand(iproc + 1 + seed) (ca_main.c:215) and rand() (functions.c:387-445). */
typedef struct orc_rng {
    int32_t r[31];
    int f, b;           /* front (fptr) and rear (rptr) indices */
    int gauss_have;     /* functions.c:380 static have_one_stored */
    float gauss_next;   /* functions.c:381 static next_val        */
} orc_rng;

static int32_t
rng_next(orc_rng *g)
{
    uint32_t val = (uint32_t) g->r[g->f] + (uint32_t) g->r[g->b];
    g->r[g->f] = (int32_t) val;
    int32_t out = (int32_t) (val >> 1);
    if (++g->f >= 31)
    {
        g->f = 0;
        ++g->b;
    }
    else if (++g->b >= 31)
        g->b = 0;
    return out;
}

static void
rng_seed(orc_rng *g, unsigned int seed)
{
    if (seed == 0)
        seed = 1;
    int32_t word = (int32_t) seed;
    g->r[0] = word;
    for (int i = 1; i < 31; i++)
    {
        long hi = word / 127773;
        long lo = word % 127773;
        word = (int32_t) (16807 * lo - 2836 * hi);
        if (word < 0)
            word += 2147483647;
        g->r[i] = word;
    }
    g->f = 3;
    g->b = 0;
    for (int i = 0; i < 310; i++)
        (void) rng_next(g);
    g->gauss_have = 0;
    g->gauss_next = 0.0f;
}

/* getRandScale (functions.c:441-445) */
static float
rand_scale(orc_rng *g)
{
    return (float) rng_next(g) / (float) 2147483647;
}

/* getRandomGaussian (functions.c:377-405): polar method with cached pair */
static float
rand_gaussian(orc_rng *g)
{
    if (g->gauss_have)
    {
        g->gauss_have = 0;
        return g->gauss_next;
    }
    float n0, n1, s;
    do
    {
        n0 = ((float) rng_next(g) / (float) (2147483647 / 2.0)) - 1.0;
        n1 = ((float) rng_next(g) / (float) (2147483647 / 2.0)) - 1.0;
        s = n0 * n0 + n1 * n1;
    }
    while (s >= 1.0 || s == 0);
    float multiplier = sqrtf(-2.0 * log(s) / s);
    float val = n0 * multiplier;
    g->gauss_next = n1 * multiplier;
    g->gauss_have = 1;
    return val;
}

/* exported for the RNG test */
void
orc_rng_test(unsigned int seed, int n, int32_t *out)
{
    orc_rng g;
    rng_seed(&g, seed);
    for (int i = 0; i < n; i++)
        out[i] = rng_next(&g);
}

/* ------------------------------------------------------------------------ */
typedef struct orc_nuc {        /* nucleation_t, globals.h:147-160 */
    double x, y, z;
    size_t cx, cy, cz;
    double gx, gy, gz;
    double threshold;
} orc_nuc;

typedef struct orc_state {
    mmca_params p;
    int nx, ny, nz;
    size_t P;
    double rev_fsgrow;
    uint64_t timestep;
    /* fields over the padded box */
    double *ce, *oce, *cl, *fs, *d, *temperature;
    double *dc;                 /* decentered_t, packed x,y,z */
    int *gr, *ogr, *diff_id, *lsindex, *nuc_id2;
    float *nuc_threshold;
    int8_t *mold;
    int gindex;
    /* growth.c:644-645 static dc_tmp.  The reference allocates dc_tmp_size
     * slots but uses slots 1..gindex with gindex <= dc_tmp_size, so slot
     * dc_tmp_size lies one element past its heap block (undefined
     * behaviour).  Here that slot exists and is zeroed with the others;
     * dc_taint marks cells whose dc was copied from it while it held no
     * value written in the same call, and taint_nucleations counts
     * nucleations of such cells (where the reference reads heap bytes). */
    double *dc_tmp;
    int dc_tmp_size;
    uint8_t *dc_taint;
    int taint_nucleations;
    /* grain.c:49 grain_cache and bp->num_grains */
    mmca_grain *grains;
    int num_grains;
    /* sb->nuc_pt */
    orc_nuc *nuc_pt;
    int nnuc;
    /* grain.c:55-59 */
    size_t numNewGrains;
    uint32_t *newGrainLocs;     /* x,y,z triples */
    size_t newGrainSize;
    orc_rng rng;
    int oriented;
    int status;
    uint64_t failed_step;
    /* when set, orc_fs_change writes each interior cell's MMCA_BRANCH_* code
     * here (padded layout, one step) */
    uint8_t *rec;
} orc_state;

#define SBIDX(s, x, y, z) \
    ((size_t) (x) + (size_t) (y) * ((s)->nx + 2) + (size_t) (z) * ((s)->nx + 2) * ((s)->ny + 2))

static void *
zalloc(size_t n)
{
    void *p = calloc(n ? n : 1, 1);
    if (!p)
    {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* CELLCOORD2REALCOORDn (functions.h:284-286) */
static double
cell2real(const orc_state *s, int axis, long i)
{
    return s->p.origin[axis] + (i + 0.5) * s->p.cell_size;
}

/* realcoord2scoord (functions.c:92-106) */
static void
real2cell(const orc_state *s, double x, double y, double z, size_t *i, size_t *j, size_t *k)
{
    x = x - s->p.origin[0] + s->p.cell_size * 0.1;
    y = y - s->p.origin[1] + s->p.cell_size * 0.1;
    z = z - s->p.origin[2] + s->p.cell_size * 0.1;
    *i = (size_t) (x / s->p.cell_size);
    *j = (size_t) (y / s->p.cell_size);
    *k = (size_t) (z / s->p.cell_size);
}

/* direct_temp_calc (temperature.c:24-126), x,y,z padded indices */
static double
direct_temp_calc(const orc_state *s, int x, int y, int z)
{
    const mmca_params *p = &s->p;
    double sim_time = s->timestep * p->dt;
    double T;
    int am = (p->temp_model == MMCA_TEMP_AM_BEAM);
    if (!am && p->gradient == 0)
    {
        T = p->t_init - sim_time * p->cool_rate;
    }
    else
    {
        int cx = p->sb_coords[0] * s->nx;
        int cy = p->sb_coords[1] * s->ny;
        int cz = p->sb_coords[2] * s->nz;
        double rx = cell2real(s, 0, cx + x);
        double ry = cell2real(s, 1, cy + y);
        double rz = cell2real(s, 2, cz + z);
        if (!am)
        {
            double axis = ((p->gdim[2] == 1) ? ry : rz) - p->cell_size * 0.5;
            double tgrad = p->gradient + p->grad_coef * sim_time;
            double v = p->velocity + p->velo_coef * sim_time;
            double vt = 0.5 * (v + p->velocity) * sim_time;
            T = p->t_init + tgrad * (axis - vt);
        }
        else
        {
            double lv = p->laser_speed;
            double invlv = 1. / lv;
            double ls = p->laser_spot / (2. * 1.414);
            double xlen = p->gnsbx * s->nx * p->cell_size;
            xlen = xlen + p->laser_spot;
            double tline = xlen * invlv;
            double tpause = 1.5e-4 * invlv;
            tline = tline + tpause;
            double trackno = 0;
            if (p->am_tracks > 1)
                trackno = floor(sim_time / tline);
            double xnow = p->laser_start[0] + (sim_time - trackno * tline) * lv;
            double ynow = p->laser_start[1] + trackno * p->laser_hatch;
            double znow = p->laser_start[2];
            double lpow = p->laser_power;
            double lx = rx - xnow;
            double ly = ry - ynow;
            double lz = rz - znow;
            if (rz <= znow + p->cell_size)
            {
                int nt = 400;
                double lt;
                double Tinteg = 0.0;
                double Tlocal, Tfront;
                double pi15 = 5.568328;
                double lalpha = p->th_conductivity / (p->density * p->specific_heat);
                double tsteady = 20 * lalpha * invlv * invlv;
                double ldt = tsteady / nt;
                if (sim_time < 1e-4)
                    lpow = 0;
                for (int i = 0; i < nt; i++)
                {
                    lt = (0.5 + i) * ldt;
                    Tlocal = sqrt(lalpha * lt) * (ls * ls + 4. * lalpha * lt);
                    Tinteg +=
                        ldt *
                        exp(-((lx + lv * lt) * (lx + lv * lt) + ly * ly) /
                            (ls * ls + 4. * lalpha * lt) -
                            (lz * lz / (4. * lalpha * lt))) / Tlocal;
                }
                Tfront = p->laser_absorp * lpow * lalpha / (pi15 * p->th_conductivity);
                T = p->sub_init_temp + Tfront * Tinteg;
            }
            else
            {
                T = 0.0;
            }
        }
    }
    return T;
}

/* tempPrepareSB_int + blank_moldface (temperature.c:156-215, 133-152) */
static void
temp_prepare(orc_state *s)
{
    int am = (s->p.temp_model == MMCA_TEMP_AM_BEAM);
    for (int z = 0; z < s->nz + 2; z++)
        for (int y = 0; y < s->ny + 2; y++)
            for (int x = 0; x < s->nx + 2; x++)
            {
                size_t i = SBIDX(s, x, y, z);
                s->temperature[i] = direct_temp_calc(s, x, y, z);
                if (am)
                    s->lsindex[i] = (s->temperature[i] >= s->p.liquidus_temp) ? 1 : 0;
            }
    int X = s->nx + 2, Y = s->ny + 2, Z = s->nz + 2;
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++)
            {
                int m = 0;
                /* one subblock: a padded axis has no neighbour (-1) on
                 * either side; a wrapped axis has itself as neighbour */
                if (s->p.pad_bndy[2] && (z == Z - 1 || z == 0)) m = 1;
                if (s->p.pad_bndy[0] && (x == 0 || x == X - 1)) m = 1;
                if (s->p.pad_bndy[1] && (y == 0 || y == Y - 1)) m = 1;
                if (m)
                    s->mold[SBIDX(s, x, y, z)] = 1;
            }
}

/* tempUpdateSB_int (temperature.c:219-272) */
void
orc_temp_update(orc_state *s)
{
    int ana = (s->p.temp_model == MMCA_TEMP_AM_BEAM);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < s->nz + 2; z++)
        for (int y = 0; y < s->ny + 2; y++)
            for (int x = 0; x < s->nx + 2; x++)
            {
                size_t i = SBIDX(s, x, y, z);
                s->temperature[i] = direct_temp_calc(s, x, y, z);
                if (ana && s->temperature[i] > s->p.liquidus_temp)
                    s->lsindex[i] = 1;
            }
}

/* allocateFields + SetupInitialConditions (setup.c:113-236, 42-96) */
orc_state *
orc_create(const mmca_params *p)
{
    if (p->nx < 1 || p->ny < 1 || p->nz < 1)
        return NULL;
    orc_state *s = (orc_state *) zalloc(sizeof(orc_state));
    s->p = *p;
    s->nx = p->nx;
    s->ny = p->ny;
    s->nz = p->nz;
    s->P = (size_t) (s->nx + 2) * (s->ny + 2) * (s->nz + 2);
    s->rev_fsgrow = 1.0 / p->fsgrow;   /* read_ctrl.c:551 */
    s->timestep = p->timestep;
    size_t P = s->P;
    s->ce = zalloc(P * sizeof(double));
    s->oce = zalloc(P * sizeof(double));
    s->cl = zalloc(P * sizeof(double));
    s->fs = zalloc(P * sizeof(double));
    s->d = zalloc(P * sizeof(double));
    s->temperature = zalloc(P * sizeof(double));
    s->dc = zalloc(3 * P * sizeof(double));
    s->gr = zalloc(P * sizeof(int));
    s->ogr = zalloc(P * sizeof(int));
    s->diff_id = zalloc(P * sizeof(int));
    s->nuc_id2 = zalloc(P * sizeof(int));
    s->lsindex = zalloc(P * sizeof(int));
    s->nuc_threshold = zalloc(P * sizeof(float));
    s->mold = zalloc(P);
    s->dc_taint = zalloc(P);
    s->grains = zalloc((size_t) (p->max_total_grains + 1) * sizeof(mmca_grain));
    s->num_grains = 1;          /* read_ctrl.c:52: the "zero" grain */
    temp_prepare(s);
    for (size_t i = 0; i < P; i++)
    {
        s->nuc_threshold[i] = INFINITY;
        s->ce[i] = s->cl[i] = p->c_init;
    }
    s->gindex = 0;
    return s;
}

void
orc_destroy(orc_state *s)
{
    if (!s)
        return;
    free(s->ce); free(s->oce); free(s->cl); free(s->fs); free(s->d);
    free(s->temperature); free(s->dc); free(s->gr); free(s->ogr);
    free(s->diff_id); free(s->nuc_id2); free(s->lsindex);
    free(s->nuc_threshold); free(s->mold); free(s->grains); free(s->nuc_pt);
    free(s->newGrainLocs); free(s->dc_tmp); free(s->dc_taint);
    free(s);
}

/* ---------------- nucleation set-up -------------------------------------- */

/* setNucOrientation (grain.c:118-129) */
static void
set_nuc_orientation(orc_state *s, orc_nuc *nuc)
{
    nuc->gx = nuc->gy = nuc->gz = 0.0;
    if (s->oriented)
    {
        float acon = 4.0;
        nuc->gx = M_PI / acon - 2 * (M_PI / acon * rand_scale(&s->rng));
        nuc->gy = M_PI / acon - 2 * (M_PI / acon * rand_scale(&s->rng));
        nuc->gz = M_PI / acon - 2 * (M_PI / acon * rand_scale(&s->rng));
    }
}

/* genRandomThreshold (grain.c:141-154) with a bound on the rejection loop */
static float
gen_random_threshold(orc_state *s, double center, double sigma)
{
    float t = 0.0;
    long tries = 0;
    do
    {
        if (++tries > 100000000L)
        {
            s->status = MMCA_ERR_NUC_THRESHOLD;
            return INFINITY;
        }
        t = (float) (sigma * rand_gaussian(&s->rng) + center);
    }
    while ((t < 1.0e-5) || (fabs(t - center) > (6.0 * sigma)));
    return t;
}

typedef void (*select_fn)(orc_state *s, size_t index, void *ud);

/* distribute_nucleations: Knuth Algorithm S (grain.c:251-296) */
static void
distribute_nucleations(orc_state *s, size_t nc, size_t num_to_select, select_fn fn, void *ud)
{
    size_t n = O_MIN(num_to_select, nc);
    size_t N = nc;
    size_t t = 0, m = 0;
    size_t idx = 0;
    if (n == 0)
        return;
    do
    {
        double U = rand_scale(&s->rng);
        if ((N - t) * U >= (n - m))
        {
            t++;
        }
        else
        {
            fn(s, idx, ud);
            t++;
            m++;
        }
        idx++;
    }
    while (m < n && s->status == MMCA_OK);
}

typedef struct vol_ud {
    size_t minCellX, minCellY, minCellZ, maxCellX, maxCellY, maxCellZ;
    double threshold;
    orc_nuc *out;
    size_t *count;
} vol_ud;

/* select_nuc_volume (grain.c:168-201) */
static void
select_nuc_volume(orc_state *s, size_t index, void *ud)
{
    vol_ud *v = (vol_ud *) ud;
    size_t nx = v->maxCellX - v->minCellX;
    size_t ny = v->maxCellY - v->minCellY;
    size_t vx = index % nx;
    size_t vy = (index % (nx * ny)) / (nx);
    size_t vz = index / (nx * ny);
    orc_nuc nuc;
    nuc.cx = v->minCellX + vx;
    nuc.cy = v->minCellY + vy;
    nuc.cz = v->minCellZ + vz;
    nuc.threshold = v->threshold;
    set_nuc_orientation(s, &nuc);
    nuc.x = cell2real(s, 0, (long) nuc.cx);
    nuc.y = cell2real(s, 1, (long) nuc.cy);
    nuc.z = cell2real(s, 2, (long) nuc.cz);
    v->out[(*v->count)++] = nuc;
}

typedef struct sel_ud {
    size_t *cells;
    double center, sigma;
} sel_ud;

/* select_if_not_fixed (grain.c:226-237) */
static void
select_if_not_fixed(orc_state *s, size_t index, void *ud)
{
    sel_ud *d = (sel_ud *) ud;
    size_t myindex = d->cells[index];
    if (s->nuc_threshold[myindex] == INFINITY)
        s->nuc_threshold[myindex] = gen_random_threshold(s, d->center, d->sigma);
}

/*
 * init_nucleation_sites (grain.c:603-713) for the sites of this subblock,
 * then init_sb_nucleation (grain.c:931-1010).  One subblock (rank 0).
 */
int
orc_init_nucleation(orc_state *s, const mmca_nucleation_cfg *cfg)
{
    rng_seed(&s->rng, (unsigned int) (cfg->rank + 1 + cfg->random_seed));
    s->oriented = cfg->oriented;
    const mmca_params *p = &s->p;
    int gdx = p->gdim[0], gdy = p->gdim[1], gdz = p->gdim[2];

    /* g_nuc_pt: fixed nucs first, in linked-list (reverse file) order */
    size_t cap = (size_t) cfg->n_fixed + 16;
    orc_nuc *g = zalloc(cap * sizeof(orc_nuc));
    size_t gn = 0;
    for (int i = cfg->n_fixed - 1; i >= 0; i--)
    {
        const mmca_fixed_nuc *f = &cfg->fixed[i];
        orc_nuc n;
        n.x = f->x;
        n.y = f->y;
        n.z = f->z;
        n.gx = (M_PI / 180.0) * f->ax;     /* grain.c:532-535 */
        n.gy = (M_PI / 180.0) * f->ay;
        n.gz = (M_PI / 180.0) * f->az;
        n.threshold = f->threshold;
        real2cell(s, n.x, n.y, n.z, &n.cx, &n.cy, &n.cz);
        g[gn++] = n;
    }
    /* NucVolumes (grain.c:644-693), list order = reverse file order */
    for (int vi = cfg->n_volumes - 1; vi >= 0; vi--)
    {
        mmca_nuc_volume v = cfg->volumes[vi];
        vol_ud u;
        v.minx = O_MAX(v.minx, p->origin[0]);
        v.maxx = O_MIN(v.maxx, p->origin[0] + (gdx * p->cell_size));
        v.miny = O_MAX(v.miny, p->origin[1]);
        v.maxy = O_MIN(v.maxy, p->origin[1] + (gdy * p->cell_size));
        v.minz = O_MAX(v.minz, p->origin[2]);
        v.maxz = O_MIN(v.maxz, p->origin[2] + (gdz * p->cell_size));
        real2cell(s, v.minx, v.miny, v.minz, &u.minCellX, &u.minCellY, &u.minCellZ);
        real2cell(s, v.maxx, v.maxy, v.maxz, &u.maxCellX, &u.maxCellY, &u.maxCellZ);
        size_t volume_count = (u.maxCellX - u.minCellX) * (u.maxCellY - u.minCellY) *
            (u.maxCellZ - u.minCellZ);
        size_t max_nucs = volume_count * v.density;
        g = realloc(g, (gn + max_nucs + 16) * sizeof(orc_nuc));
        u.threshold = v.threshold;
        u.out = g;
        u.count = &gn;
        distribute_nucleations(s, volume_count, max_nucs, select_nuc_volume, &u);
    }
    /* keep the points inside this subblock (grain.c:695-711) */
    s->nuc_pt = zalloc((gn + 1) * sizeof(orc_nuc));
    s->nnuc = 0;
    for (size_t i = 0; i < gn; i++)
    {
        size_t ix, iy, iz;
        real2cell(s, g[i].x, g[i].y, g[i].z, &ix, &iy, &iz);
        if ((int) (ix / s->nx) == p->sb_coords[0] && (int) (iy / s->ny) == p->sb_coords[1] &&
            (int) (iz / s->nz) == p->sb_coords[2])
            s->nuc_pt[s->nnuc++] = g[i];
    }
    free(g);

    /* init_sb_nucleation (grain.c:931-1010) */
    int dimx = s->nx, dimy = s->ny;
    for (int nidx = 0; nidx < s->nnuc; nidx++)
    {
        orc_nuc *n = &s->nuc_pt[nidx];
        size_t x = 1 + n->cx % s->nx;
        size_t y = 1 + n->cy % s->ny;
        size_t z = 1 + n->cz % s->nz;
        size_t index = z * (dimy + 2) * (dimx + 2) + y * (dimx + 2) + x;
        if (!s->mold[index])
            s->nuc_threshold[index] = n->threshold;
    }
    size_t N = (size_t) s->nx * s->ny * s->nz;
    size_t *suf_cells = zalloc(N * sizeof(size_t));
    size_t *vol_cells = zalloc(N * sizeof(size_t));
    size_t sc = 0, vc = 0;
    for (int z = 1; z <= s->nz; z++)
        for (int y = 1; y <= s->ny; y++)
            for (int x = 1; x <= s->nx; x++)
            {
                size_t i = SBIDX(s, x, y, z);
                if (!s->mold[i])
                {
                    vol_cells[vc++] = i;
                    int surf = 0;
                    surf += s->mold[SBIDX(s, x - 1, y, z)];
                    surf += s->mold[SBIDX(s, x + 1, y, z)];
                    surf += s->mold[SBIDX(s, x, y - 1, z)];
                    surf += s->mold[SBIDX(s, x, y + 1, z)];
                    surf += s->mold[SBIDX(s, x, y, z - 1)];
                    surf += s->mold[SBIDX(s, x, y, z + 1)];
                    if (surf)
                        suf_cells[sc++] = i;
                }
            }
    sel_ud ud;
    ud.cells = vol_cells;
    ud.center = cfg->bulk_mean;
    ud.sigma = cfg->bulk_sigma;
    size_t gr_count = cfg->bulk_density * (vc * pow(p->cell_size, 3.0));
    distribute_nucleations(s, vc, gr_count, select_if_not_fixed, &ud);
    if (cfg->do_surface && (cfg->surf_density > 0.0) && s->status == MMCA_OK)
    {
        ud.cells = suf_cells;
        ud.center = cfg->surf_mean;
        ud.sigma = cfg->surf_sigma;
        gr_count = cfg->surf_density * (sc * pow(p->cell_size, 2.0));
        distribute_nucleations(s, sc, gr_count, select_if_not_fixed, &ud);
    }
    free(suf_cells);
    free(vol_cells);
    return s->status;
}

/* ---------------- per-step operators -------------------------------------- */

/* cell_nucleation, NUC_SEP branch (grain.c:1222-1318) */
int
orc_cell_nucleation(orc_state *s)
{
    const mmca_params *p = &s->p;
    int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    double a = p->m_solute[0], b = p->m_solute[1], c = p->m_solute[2];
    double cinit = p->c_init;
    int nindex2 = 0;
    double sim_time = s->timestep * p->dt;
    int am = (p->temp_model == MMCA_TEMP_AM_BEAM);
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
            {
                int idx = k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i;
                double Ttemp = s->temperature[idx];
                int layerno = 1;
                int initnuc = 1;
                if (am)
                {
                    if (p->am_layers == 1)
                        layerno = 1;
                    else if (p->am_layers > 1 && s->lsindex[idx] == 1)
                        layerno = 1;
                    else
                        layerno = 0;
                    if (sim_time <= 1e-4)
                        initnuc = 1;
                    else if (Ttemp >= p->solidus_temp)
                        initnuc = 1;
                    else
                        initnuc = 0;
                }
                if (s->gr[idx] <= 0 && layerno && initnuc)
                {
                    double Tunder = p->liquidus_temp - Ttemp;
                    double m_solute0 = a * Ttemp * Ttemp + b * Ttemp + c;
                    Tunder += m_solute0 * (s->cl[idx] - cinit);
                    if (Tunder > s->nuc_threshold[idx])
                    {
                        nindex2 = nindex2 + 1;
                        s->nuc_id2[nindex2] = idx;
                    }
                }
            }
    int coordx = p->sb_coords[0] * dimx;
    int coordy = p->sb_coords[1] * dimy;
    int coordz = p->sb_coords[2] * dimz;
    for (int i = 1; i <= nindex2; i++)
    {
        int idx = s->nuc_id2[i];
        int dimxy = (dimx + 2) * (dimy + 2);
        int kk = idx / dimxy;
        int jj = (idx - kk * dimxy) / (dimx + 2);
        int ii = idx - kk * dimxy - jj * (dimx + 2);
        uint32_t gx = (ii - 1) + coordx;
        uint32_t gy = (jj - 1) + coordy;
        uint32_t gz = (kk - 1) + coordz;
        /* addNewGrain (grain.c:89-109) */
        if (s->numNewGrains + 1 > s->newGrainSize)
        {
            s->newGrainSize = O_MAX(8, 2 * s->newGrainSize);
            s->newGrainLocs = realloc(s->newGrainLocs, s->newGrainSize * 3 * sizeof(uint32_t));
        }
        if ((s->num_grains + s->numNewGrains) < (uint64_t) (p->max_total_grains + 1))
        {
            s->newGrainLocs[3 * s->numNewGrains + 0] = gx;
            s->newGrainLocs[3 * s->numNewGrains + 1] = gy;
            s->newGrainLocs[3 * s->numNewGrains + 2] = gz;
            s->numNewGrains++;
        }
    }
    s->numNewGrains = nindex2;   /* grain.c:1318 */
    return nindex2;
}

/* nucleate_grain (grain.c:301-344) */
static void
nucleate_grain(orc_state *s, const orc_nuc *nuc, mmca_grain *grain)
{
    double c0, s0, c1, s1, c2, s2;
    double a0 = nuc->gx, a1 = nuc->gy, a2 = nuc->gz;
    /* the reference's gcc build turns each cos/sin pair into one sincos() */
    sincos(a0, &s0, &c0);
    sincos(a1, &s1, &c1);
    sincos(a2, &s2, &c2);
    grain->rotang[0] = nuc->gx;
    grain->rotang[1] = nuc->gy;
    grain->rotang[2] = nuc->gz;
    grain->rotmat[0][0] = c2 * c0 - s2 * s1 * s0;
    grain->rotmat[1][0] = c2 * s0 + s2 * c1 * c0;
    grain->rotmat[2][0] = s2 * s1;
    grain->rotmat[0][1] = -1. * s2 * c0 - c2 * c1 * s0;
    grain->rotmat[1][1] = -1. * s2 * s0 + c2 * c1 * c0;
    grain->rotmat[2][1] = c2 * s1;
    grain->rotmat[0][2] = s1 * s0;
    grain->rotmat[1][2] = -1. * s1 * c0;
    grain->rotmat[2][2] = c1;
    grain->nuc_x = nuc->x;
    grain->nuc_y = nuc->y;
    grain->nuc_z = nuc->z;
    grain->nuc_timestep = s->timestep;
    grain->nuc_threshold = nuc->threshold;
}

/* activateNewGrains + createGrains on one rank (grain.c:841-916, 348-425) */
int
orc_activate_new_grains(orc_state *s)
{
    const mmca_params *p = &s->p;
    int grains = (int) s->numNewGrains;
    int new_activations = grains;
    int count = grains;
    if (s->num_grains + new_activations > (p->max_total_grains + 1))
    {
        new_activations = (int) ((p->max_total_grains + 1) - s->num_grains);
        if ((size_t) count > (size_t) p->max_total_grains)
            count = (int) p->max_total_grains;
    }
    if (new_activations)
    {
        int base = s->num_grains;
        for (int g = 0; g < count; g++)
        {
            int gr_num = base + g;
            if (gr_num > p->max_total_grains)
            {
                s->status = MMCA_ERR_GRAIN_CAPACITY;   /* grain.c:361-366 */
                s->failed_step = s->timestep;
                return s->status;
            }
            int x = s->newGrainLocs[3 * g + 0];
            int y = s->newGrainLocs[3 * g + 1];
            int z = s->newGrainLocs[3 * g + 2];
            int sx = 1 + (x % s->nx);
            int sy = 1 + (y % s->ny);
            int sz = 1 + (z % s->nz);
            int idx = sx + sy * (s->nx + 2) + sz * (s->nx + 2) * (s->ny + 2);
            s->gr[idx] = gr_num;
            if (s->dc_taint[idx])
                s->taint_nucleations++;
            int found = 0;
            orc_nuc nuc;
            for (int j = 0; j < s->nnuc; j++)
            {
                if ((s->nuc_pt[j].cx == (size_t) x) && (s->nuc_pt[j].cy == (size_t) y) &&
                    (s->nuc_pt[j].cz == (size_t) z))
                {
                    nuc = s->nuc_pt[j];
                    found = 1;
                    break;
                }
            }
            if (!found)
            {
                nuc.cx = x;
                nuc.cy = y;
                nuc.cz = z;
                nuc.threshold = s->nuc_threshold[idx];
                nuc.x = cell2real(s, 0, x);
                nuc.y = cell2real(s, 1, y);
                nuc.z = cell2real(s, 2, z);
                set_nuc_orientation(s, &nuc);
            }
            nucleate_grain(s, &nuc, &s->grains[gr_num]);
        }
    }
    s->num_grains += new_activations;
    s->numNewGrains = 0;
    return MMCA_OK;
}

/* oce = ce (calculate.c:267-280) */
void
orc_copy_oce(orc_state *s)
{
    memcpy(s->oce, s->ce, s->P * sizeof(double));
}

/* sb_diffuse_alloy_decentered (growth.c:30-170) */
int
orc_diffuse(orc_state *s)
{
    const mmca_params *p = &s->p;
    int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    double dtx = p->dt / (p->cell_size * p->cell_size);
    double rs = (p->d_sol * dtx);
    double rl = (p->d_liq * dtx);
    if (rs > p->courant_limit || rl > p->courant_limit)
    {
        s->status = MMCA_ERR_COURANT;
        s->failed_step = s->timestep;
        return s->status;
    }
    int8_t *mold = s->mold;
    double *fs = s->fs, *cl = s->cl, *ce = s->ce;
    int negative = 0;
    /* cells are independent (ce from oce-time cl and fs); a negative ce stops
     * the step as the reference's abort does (growth.c:163-165) */
#pragma omp parallel for schedule(static) reduction(|:negative)
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
            {
                int idx = k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i;
                double nbsum = 0;
                double conc = cl[idx];
                double lspfs = fs[idx];
                int idxp = idx + 1;
                int idxm = idx - 1;
                int idyp = idx + dimx + 2;
                int idym = idx - (dimx + 2);
                int idzp = idx + (dimy + 2) * (dimx + 2);
                int idzm = idx - (dimy + 2) * (dimx + 2);
                double fs_av, r;
                fs_av = 0.5 * (lspfs + fs[idxp]);
                r = (rs * fs_av + rl * (1 - fs_av)) * (1 - mold[idxp]);
                nbsum += r * (cl[idxp] - conc);
                fs_av = 0.5 * (lspfs + fs[idxm]);
                r = (rs * fs_av + rl * (1 - fs_av)) * (1 - mold[idxm]);
                nbsum += r * (cl[idxm] - conc);
                fs_av = 0.5 * (lspfs + fs[idyp]);
                r = (rs * fs_av + rl * (1 - fs_av)) * (1 - mold[idyp]);
                nbsum += r * (cl[idyp] - conc);
                fs_av = 0.5 * (lspfs + fs[idym]);
                r = (rs * fs_av + rl * (1 - fs_av)) * (1 - mold[idym]);
                nbsum += r * (cl[idym] - conc);
                fs_av = 0.5 * (lspfs + fs[idzp]);
                r = (rs * fs_av + rl * (1 - fs_av)) * (1 - mold[idzp]);
                nbsum += r * (cl[idzp] - conc);
                fs_av = 0.5 * (lspfs + fs[idzm]);
                r = (rs * fs_av + rl * (1 - fs_av)) * (1 - mold[idzm]);
                nbsum += r * (cl[idzm] - conc);
                ce[idx] += nbsum;
                if (p->check_negative_ce && ce[idx] < 0)
                    negative = 1;
            }
    if (negative)
    {
        s->status = MMCA_ERR_NEGATIVE_CE;
        s->failed_step = s->timestep;
        return s->status;
    }
    return MMCA_OK;
}

/* fs_change_diffuse (growth.c:177-276), tip_curv = 0 */
void
orc_fs_change(orc_state *s)
{
    const mmca_params *p = &s->p;
    const double liq_slope = p->m_solute[2];
    const double k_inv = 1.0 / p->part_coef;
    const double km = 1.0 - p->part_coef;
    const double km_inv = 1.0 / km;
    const double m_inv = 1.0 / liq_slope;
    const double D_x = p->d_liq / p->cell_size / p->cell_size;
    const double D_xt = D_x * p->dt;
    const int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    const double cinit = p->c_init;
    const double Tliq = p->liquidus_temp;
    double *oce = s->oce, *ce = s->ce, *fs = s->fs, *cl = s->cl, *temperature = s->temperature;
    int *gr = s->gr;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
            {
                int idx = k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i;
                double Tcell = temperature[idx];
                double c_int = cinit + m_inv * (Tcell - Tliq);
                double Tunder = Tliq - Tcell;
                Tunder += liq_slope * (cl[idx] - cinit);
                if (gr[idx] <= 0 || Tcell < 1.0 || Tcell > Tliq)
                {
                    cl[idx] = ce[idx];
                    gr[idx] = 0;
                    fs[idx] = 0;
                    if (s->rec)
                        s->rec[idx] = MMCA_BRANCH_LIQUID;
                    continue;
                }
                if (fs[idx] == 1.0 && (Tunder >= 0.0))
                {
                    cl[idx] = ce[idx] * k_inv;
                    if (s->rec)
                        s->rec[idx] = MMCA_BRANCH_SOLID;
                    continue;
                }
                double dce = ce[idx] - oce[idx];
                double dcl = (c_int - cl[idx]) * D_xt * 2 + dce / (1 - km * fs[idx]);
                double dfs = (dcl * (1. - km * fs[idx]) - dce) * km_inv / cl[idx];
                cl[idx] += dcl;
                fs[idx] += dfs;
                int branch = (1.0 < fs[idx]) ? MMCA_BRANCH_FULL : MMCA_BRANCH_GROW;
                fs[idx] = O_MIN(1.0, fs[idx]);
                if (fs[idx] < 0.)
                {
                    fs[idx] = 0.;
                    gr[idx] = 0;
                    cl[idx] = ce[idx];
                    branch = MMCA_BRANCH_REMELT;
                }
                if (s->rec)
                    s->rec[idx] = (uint8_t) branch;
            }
}

/* grow_octahedra (growth.c:339-407) */
void
orc_grow_octahedra(orc_state *s)
{
    const int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    const double fsgrow = s->rev_fsgrow;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
            {
                int idx = k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i;
                int grid = s->gr[idx];
                if (grid > 0 && grid < s->p.max_total_grains)
                {
                    const mmca_grain *g = &s->grains[grid];
                    double g00 = g->rotmat[0][0], g10 = g->rotmat[1][0], g20 = g->rotmat[2][0];
                    double g01 = g->rotmat[0][1], g11 = g->rotmat[1][1], g21 = g->rotmat[2][1];
                    double g02 = g->rotmat[0][2], g12 = g->rotmat[1][2], g22 = g->rotmat[2][2];
                    double rx0 = 0.0 - s->dc[3 * idx + 0];
                    double ry0 = 0.0 - s->dc[3 * idx + 1];
                    double rz0 = 0.0 - s->dc[3 * idx + 2];
                    double gx0 = g00 * rx0 + g10 * ry0 + g20 * rz0;
                    double gy0 = g01 * rx0 + g11 * ry0 + g21 * rz0;
                    double gz0 = g02 * rx0 + g12 * ry0 + g22 * rz0;
                    double ds1 = fabs(gx0) + fabs(gy0) + fabs(gz0);
                    ds1 = O_MIN(ds1, 1.5);
                    s->d[idx] = ds1 + s->fs[idx] * fsgrow;
                }
            }
}

/* grow_cell_reduction (growth.c:461-538) */
int
orc_grow_cell_reduction(orc_state *s)
{
    int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    int gindex = 0;
    int *gr = s->gr;
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
            {
                int idx = k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i;
                int idxp = idx + 1;
                int idxm = idx - 1;
                int idyp = idx + dimx + 2;
                int idym = idx - (dimx + 2);
                int idzp = idx + (dimy + 2) * (dimx + 2);
                int idzm = idx - (dimy + 2) * (dimx + 2);
                if (s->mold[idx])
                    continue;
                if (gr[idx] <= 0)
                {
                    int nbindex = gr[idxp] + gr[idxm] + gr[idyp] + gr[idym] + gr[idzp] + gr[idzm];
                    if (nbindex > 0)
                    {
                        gindex = gindex + 1;
                        s->diff_id[gindex] = idx;
                    }
                }
            }
    s->gindex = gindex;
    return gindex;
}

/* capture_octahedra_diffuse, INDEX_SEP branch (growth.c:574-807) */
void
orc_capture(orc_state *s)
{
    static const short nn[6][3] = { {1, 0, 0}, {0, 1, 0}, {-1, 0, 0},
        {0, -1, 0}, {0, 0, 1}, {0, 0, -1} };
    int dimx = s->nx, dimy = s->ny;
    double *d = s->d, *dc = s->dc, *ce = s->ce, *cl = s->cl, *fs = s->fs;
    int *gr = s->gr, *ogr = s->ogr;
    memcpy(ogr, gr, s->P * sizeof(int));
    int gindex = s->gindex;
    if (gindex > s->dc_tmp_size)
    {
        free(s->dc_tmp);
        s->dc_tmp_size = gindex + 100;
        s->dc_tmp = zalloc((size_t) (s->dc_tmp_size + 1) * 3 * sizeof(double));
    }
    double *dc_tmp = s->dc_tmp;
    int *diff_id = s->diff_id;
    /* listed cells are independent: each reads solid neighbours through the
     * ogr snapshot and writes only itself and its slot */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= gindex; i++)
    {
        int idx = diff_id[i];
        if (ogr[idx] > 0)
            continue;
        const int ndxs[6] = { idx + 1, idx + dimx + 2, idx - 1, idx - (dimx + 2),
            idx + (dimy + 2) * (dimx + 2), idx - (dimy + 2) * (dimx + 2) };
        double g00 = 0, g10 = 0, g20 = 0, g01 = 0, g11 = 0, g21 = 0, g02 = 0, g12 = 0, g22 = 0;
        double dx = 0, dy = 0, dz = 0;
        int ndx = 0;
        short found = -1;
        for (short n = 0; n < 6; n++)
        {
            ndx = ndxs[n];
            int gid = ogr[ndx];
            if (gid > 0 && gid < s->p.max_total_grains)
            {
                const mmca_grain *g = &s->grains[gid];
                g00 = g->rotmat[0][0]; g10 = g->rotmat[1][0]; g20 = g->rotmat[2][0];
                g01 = g->rotmat[0][1]; g11 = g->rotmat[1][1]; g21 = g->rotmat[2][1];
                g02 = g->rotmat[0][2]; g12 = g->rotmat[1][2]; g22 = g->rotmat[2][2];
                double rx1 = -1.0 * nn[n][0] - dc[3 * ndx + 0];
                double ry1 = -1.0 * nn[n][1] - dc[3 * ndx + 1];
                double rz1 = -1.0 * nn[n][2] - dc[3 * ndx + 2];
                dx = g00 * rx1 + g10 * ry1 + g20 * rz1;
                dy = g01 * rx1 + g11 * ry1 + g21 * rz1;
                dz = g02 * rx1 + g12 * ry1 + g22 * rz1;
                double dfs = d[ndx] - (fabs(dx) + fabs(dy) + fabs(dz));
                if (dfs > 0)
                {
                    found = n;
                    break;
                }
            }
        }
        if (found > -1)
        {
            double ncx = 0., ncz = 0., ncy = 0.;
            double ncd = O_MIN(d[ndx], 1.0);
            double dd = d[ndx] - ncd;
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
            dc_tmp[3 * i + 0] = dc[3 * ndx + 0] + nn[found][0] + g00 * ncx + g01 * ncy + g02 * ncz;
            dc_tmp[3 * i + 1] = dc[3 * ndx + 1] + nn[found][1] + g10 * ncx + g11 * ncy + g12 * ncz;
            dc_tmp[3 * i + 2] = dc[3 * ndx + 2] + nn[found][2] + g20 * ncx + g21 * ncy + g22 * ncz;
            gr[idx] = ogr[ndx];
            fs[idx] = 0.;
            cl[idx] = ce[idx];
        }
    }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= gindex; i++)
    {
        int idx = diff_id[i];
        s->dc_taint[idx] = (i == s->dc_tmp_size && gr[idx] == ogr[idx]);
        dc[3 * idx + 0] = dc_tmp[3 * i + 0];
        dc[3 * idx + 1] = dc_tmp[3 * i + 1];
        dc[3 * idx + 2] = dc_tmp[3 * i + 2];
    }
}

/* reduce_fs (growth.c:541-566): serial sum of min(1,fs) */
double
orc_reduce_fs(const orc_state *s)
{
    double sum_fs = 0.0;
    int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
            {
                int idx = k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i;
                sum_fs += O_MIN(1.0, s->fs[idx]);
            }
    return sum_fs;
}

/* sum of ce over the interior (solute ledger; no reference counterpart) */
double
orc_reduce_ce(const orc_state *s)
{
    double sum = 0.0;
    int dimx = s->nx, dimy = s->ny, dimz = s->nz;
    for (int k = 1; k <= dimz; k++)
        for (int j = 1; j <= dimy; j++)
            for (int i = 1; i <= dimx; i++)
                sum += s->ce[k * (dimy + 2) * (dimx + 2) + j * (dimx + 2) + i];
    return sum;
}

/* doiteration (calculate.c:193-352) on one subblock, no halo exchange */
int
orc_doiteration(orc_state *s)
{
    if (s->status != MMCA_OK)
        return s->status;
    orc_temp_update(s);
    orc_cell_nucleation(s);
    if (orc_activate_new_grains(s) != MMCA_OK)
        return s->status;
    orc_copy_oce(s);
    if (orc_diffuse(s) != MMCA_OK)
        return s->status;
    orc_fs_change(s);
    orc_grow_octahedra(s);
    orc_grow_cell_reduction(s);
    orc_capture(s);
    return MMCA_OK;
}

/* loop() body without IO (loop.c:222-227) */
int
orc_step(orc_state *s, int nsteps)
{
    for (int n = 0; n < nsteps; n++)
    {
        s->timestep++;
        if (orc_doiteration(s) != MMCA_OK)
            return s->status;
    }
    return MMCA_OK;
}

/* ---------------- accessors for the ctypes wrapper ------------------------ */
void *
orc_field(orc_state *s, int which)
{
    switch (which)
    {
        case MMCA_FIELD_CE: return s->ce;
        case MMCA_FIELD_CL: return s->cl;
        case MMCA_FIELD_FS: return s->fs;
        case MMCA_FIELD_GR: return s->gr;
        case MMCA_FIELD_DC: return s->dc;
        case MMCA_FIELD_D: return s->d;
        case MMCA_FIELD_TEMP: return s->temperature;
        case MMCA_FIELD_THRESHOLD: return s->nuc_threshold;
        case MMCA_FIELD_MOLD: return s->mold;
        case MMCA_FIELD_OCE: return s->oce;
        default: return NULL;
    }
}

void orc_set_record(orc_state *s, uint8_t *rec) { s->rec = rec; }
/* n raw rand() values of the block's stream (the XDMF writer's grain
 * renumbering draws them between steps, io_xdmf.c:345-376) */
void
orc_rand(orc_state *s, int n, int32_t *out)
{
    for (int i = 0; i < n; i++)
        out[i] = rng_next(&s->rng);
}

/* threads the parallel loops use (1 without OpenMP) */
int
orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
/* grain_cache records 0..n-1 and bp->num_grains, as a restart restores them
 * (addGrainCache, checkpoint.c:676-745) */
int
orc_set_grains(orc_state *s, const mmca_grain *g, int n)
{
    if (n < 1 || n > s->p.max_total_grains + 1)
        return MMCA_ERR_ARG;
    memcpy(s->grains, g, sizeof(mmca_grain) * (size_t) n);
    s->num_grains = n;
    return MMCA_OK;
}
int orc_num_grains(const orc_state *s) { return s->num_grains; }
int orc_gindex(const orc_state *s) { return s->gindex; }
int orc_status(const orc_state *s) { return s->status; }
uint64_t orc_failed_step(const orc_state *s) { return s->failed_step; }
uint64_t orc_timestep(const orc_state *s) { return s->timestep; }
void orc_set_timestep(orc_state *s, uint64_t t) { s->timestep = t; }
int *orc_diff_id(orc_state *s) { return s->diff_id; }
mmca_grain *orc_grains(orc_state *s) { return s->grains; }
int orc_nnuc(const orc_state *s) { return s->nnuc; }
int orc_dc_tmp_size(const orc_state *s) { return s->dc_tmp_size; }
int orc_taint_nucleations(const orc_state *s) { return s->taint_nucleations; }
uint8_t *orc_dc_taint(orc_state *s) { return s->dc_taint; }
double *orc_dc_tmp(orc_state *s) { return s->dc_tmp; }
/* y[i] = exp(x[i]) through the host C library: the checker of the device's
 * restatement of it (temperature.c:108-112 calls this exp) */
void
orc_exp_array(const double *x, double *y, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        y[i] = exp(x[i]);
}

size_t orc_sizeof_grain(void) { return sizeof(mmca_grain); }
size_t orc_sizeof_params(void) { return sizeof(mmca_params); }
size_t orc_sizeof_nuc_cfg(void) { return sizeof(mmca_nucleation_cfg); }

/* FNV-1a 64 over a byte range (same hash the reference dump hook writes) */
uint64_t
orc_fnv1a(const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *) p;
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; i++)
    {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}
