/*
 * mmca_run — C host driver of the CA step: reads a muMatScale control file,
 * runs the reference's main loop on one GPU and writes its grain summary.
 *
 *   loop()       loop.c:186-272   timestep++, doiteration(), screen line
 *                                 every ScreenPFreq steps, checkDone
 *   checkDone    loop.c:148-170   FinishTime (every step) and FsFinish
 *                                 (fs refreshed on ScreenPFreq steps)
 *   computeFS    loop.c:173-182   solid_volume / totalNonMoldVolume
 *   output_grains grain.c:752-833 <BaseFileName>_grains.csv and the
 *                                 "N grains. Min ... Average" line
 *   writeData    loop.c:116-127   DataFormat XDMF or VTK: at timestep 0,
 *                                 every DataWriteFreq steps from
 *                                 DataWriteStart on and after the last step
 *                                 (mmca_xdmf.h, mmca_vtk.h; the writers live
 *                                 in libmmca_io.so, loaded only when such
 *                                 output is asked for)
 *   writeCheckpoint loop.c:129-145 every checkpointfreq steps (DataWriteStart,
 *                                 as the reference reads it, or
 *                                 --checkpoint-freq): the restart image of
 *                                 the whole run in cp_main.mmca; --restart
 *                                 continues from it bit for bit
 *
 * With NSubblocks x y z the domain runs as the reference runs it on x*y*z
 * MPI ranks: one subblock per rank, each with its own random stream, halos
 * exchanged between them and grain ids given in rank order (a coupled
 * batch, mmca_batch_couple), all on one GPU.  --one-block runs the whole
 * domain as a single block with rank 0's stream instead.  Steps between two
 * screen or finish checks are issued as one step call.
 *
 * Usage: mmca_run <ctrl-file> [--device N] [--max-steps N] [--fp32]
 *                 [--no-ce-check] [--one-block] [--checkpoint-freq N]
 *                 [--restart] [--dump-ctrl]
 */
#include <dlfcn.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "mmca.h"
#include "mmca_ctrl.h"
#include "mmca_ckpt.h"
#include "mmca_vtk.h"
#include "mmca_xdmf.h"

/* The output writers (XDMF with libhdf5 behind it, VTK with zlib), from
 * libmmca_io.so beside this program; runs without data output do not need
 * HDF5 at all. */
typedef struct io_lib {
    void *so;
    mmca_xdmf *(*open)(const mmca_run_cfg *, const int *, const int *);
    int (*write)(mmca_xdmf *, mmca_ctx *const *, const mmca_params *, uint64_t);
    void (*close)(mmca_xdmf *);
    int (*flush)(mmca_xdmf *);
    int64_t (*get_renumber)(const mmca_xdmf *, uint64_t *, int64_t);
    int (*set_renumber)(mmca_xdmf *, const uint64_t *, int64_t);
    mmca_vtk *(*vopen)(const mmca_run_cfg *, const int *, const int *);
    int (*vwrite)(mmca_vtk *, mmca_ctx *const *, const mmca_params *, uint64_t);
    void (*vclose)(mmca_vtk *);
    int (*cwrite)(const mmca_run_cfg *, const int *, mmca_batch *, mmca_ctx *const *, const mmca_params *,
                  const mmca_ckpt_snap *const *, uint64_t, const uint64_t *, int64_t);
    int (*cread)(const mmca_run_cfg *, const int *, mmca_batch *, mmca_ctx *const *, const mmca_params *,
                 uint64_t *, int64_t, int64_t *);
    mmca_xdmf *x;                  /* DataFormat XDMF */
    mmca_vtk *v;                   /* DataFormat VTK */
} io_lib;

/* libmmca_io.so beside this program */
static void *
io_open(io_lib *w, char *path, size_t n)
{
    ssize_t k = readlink("/proc/self/exe", path, n - 32);
    if (k <= 0)
        return NULL;
    path[k] = 0;
    char *slash = strrchr(path, '/');
    strcpy(slash ? slash + 1 : path, "libmmca_io.so");
    if (!w->so)
        w->so = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    return w->so;
}

/* the checkpoint writer of libmmca_io.so (mmca_ckpt.h); -1 without it */
static int
ckpt_load(io_lib *w)
{
    char path[PATH_MAX];
    if (!io_open(w, path, sizeof(path)))
        return -1;
    *(void **) &w->cwrite = dlsym(w->so, "mmca_ckpt_write");
    *(void **) &w->cread = dlsym(w->so, "mmca_ckpt_read");
    return (w->cwrite && w->cread) ? 0 : -1;
}

static int
io_load(io_lib *w, int vtk)
{
    char path[PATH_MAX];
    if (!io_open(w, path, sizeof(path)))
    {
        fprintf(stderr, "DataFormat %s needs %s: %s\n", vtk ? "VTK" : "XDMF", path, dlerror());
        return -1;
    }
    int ok;
    if (vtk)
    {
        *(void **) &w->vopen = dlsym(w->so, "mmca_vtk_open");
        *(void **) &w->vwrite = dlsym(w->so, "mmca_vtk_write");
        *(void **) &w->vclose = dlsym(w->so, "mmca_vtk_close");
        ok = w->vopen && w->vwrite && w->vclose;
    }
    else
    {
        *(void **) &w->open = dlsym(w->so, "mmca_xdmf_open");
        *(void **) &w->write = dlsym(w->so, "mmca_xdmf_write");
        *(void **) &w->close = dlsym(w->so, "mmca_xdmf_close");
        *(void **) &w->flush = dlsym(w->so, "mmca_xdmf_flush");
        *(void **) &w->get_renumber = dlsym(w->so, "mmca_xdmf_get_renumber");
        *(void **) &w->set_renumber = dlsym(w->so, "mmca_xdmf_set_renumber");
        ok = w->open && w->write && w->close && w->flush && w->get_renumber && w->set_renumber;
    }
    if (!ok)
    {
        fprintf(stderr, "%s lacks the writer's entry points\n", path);
        return -1;
    }
    return 0;
}

/* writeData (loop.c:116-127) in the configured format */
static int
data_write(io_lib *w, mmca_ctx *const *doms, const mmca_params *ps, uint64_t t)
{
    return w->x ? w->write(w->x, doms, ps, t) : w->vwrite(w->v, doms, ps, t);
}

static double
wall_now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1.0e-6;
}

static int
check(int st, const char *what)
{
    if (st != MMCA_OK)
        fprintf(stderr, "%s: %s\n", what, mmca_status_string(st));
    return st;
}

/* First timestep >= t with timestep * dt >= finish_time (loop.c:154-155). */
static uint64_t
finish_step(uint64_t t, double dt, double finish_time)
{
    double g = floor(finish_time / dt) - 2.0;
    uint64_t s = (g > (double) t) ? (uint64_t) g : t;
    while (s * dt < finish_time)
        s++;
    return s;
}

/*
 * writeCheckpoint / mainRestart (loop.c:129-145, checkpoint.c:867-1238) with
 * this build's restart image: the whole run (every rank, its random stream
 * and dc_tmp slots, the grain table) plus the XDMF writer's grain
 * renumbering, so that a restarted run continues bit for bit and writes the
 * same files.  Written as cp_main_<step>.mmca and renamed to cp_main.mmca,
 * as clean_checkpoint (checkpoint.c:165-177) renames its file.
 *
 *   "MMCARUN1", int64 image bytes, int64 renumber entries, image, entries
 */
static const char CP_MAGIC[8] = { 'M', 'M', 'C', 'A', 'R', 'U', 'N', '1' };

static int
save_checkpoint(mmca_batch *bt, mmca_ctx *ctx, const io_lib *out, uint64_t t)
{
    int64_t n = 0;
    if (check(bt ? mmca_batch_save_state(bt, NULL, 0, &n) : mmca_save_state(ctx, NULL, 0, &n), "save state"))
        return -1;
    const int64_t nren = out->x ? out->get_renumber(out->x, NULL, 0) : 0;
    char *buf = malloc((size_t) n);
    uint64_t *ren = malloc(sizeof(uint64_t) * (size_t) (nren > 0 ? nren : 1));
    if (!buf || !ren)
    {
        fprintf(stderr, "out of host memory\n");
        return -1;
    }
    if (check(bt ? mmca_batch_save_state(bt, buf, n, &n) : mmca_save_state(ctx, buf, n, &n), "save state"))
        return -1;
    if (out->x)
        out->get_renumber(out->x, ren, nren);
    char tmp[64];
    snprintf(tmp, sizeof(tmp), "cp_main_%05lu.mmca", (unsigned long) t);
    FILE *f = fopen(tmp, "wb");
    int ok = f && fwrite(CP_MAGIC, 8, 1, f) == 1 && fwrite(&n, 8, 1, f) == 1 && fwrite(&nren, 8, 1, f) == 1 &&
             fwrite(buf, 1, (size_t) n, f) == (size_t) n &&
             (nren == 0 || fwrite(ren, sizeof(uint64_t), (size_t) nren, f) == (size_t) nren);
    if (f && fclose(f) != 0)
        ok = 0;
    free(buf);
    free(ren);
    if (!ok || rename(tmp, "cp_main.mmca") != 0)
    {
        fprintf(stderr, "Unable to write checkpoint %s\n", tmp);
        return -1;
    }
    printf("Checkpoint at timestep %lu \n", (unsigned long) t);
    return 0;
}

static int
load_checkpoint(mmca_batch *bt, mmca_ctx *ctx, io_lib *out)
{
    FILE *f = fopen("cp_main.mmca", "rb");
    char magic[8];
    int64_t n = 0, nren = 0;
    if (!f || fread(magic, 8, 1, f) != 1 || memcmp(magic, CP_MAGIC, 8) != 0 || fread(&n, 8, 1, f) != 1 ||
        fread(&nren, 8, 1, f) != 1 || n <= 0 || nren < 0)
    {
        fprintf(stderr, "cp_main.mmca is missing or not a checkpoint of this driver\n");
        if (f)
            fclose(f);
        return -1;
    }
    char *buf = malloc((size_t) n);
    uint64_t *ren = malloc(sizeof(uint64_t) * (size_t) (nren > 0 ? nren : 1));
    int ok = buf && ren && fread(buf, 1, (size_t) n, f) == (size_t) n &&
             (nren == 0 || fread(ren, sizeof(uint64_t), (size_t) nren, f) == (size_t) nren);
    fclose(f);
    if (!ok)
    {
        fprintf(stderr, "cp_main.mmca is truncated\n");
        return -1;
    }
    if (check(bt ? mmca_batch_load_state(bt, buf, n) : mmca_load_state(ctx, buf, n), "load state"))
        return -1;
    if (out->x && out->set_renumber(out->x, ren, nren))
        return -1;
    free(buf);
    free(ren);
    return 0;
}

/*
 * writeCheckpoint (loop.c:129-145) in the reference's file layout
 * (writeTaskCheckpoint, checkpoint.c:955-1238): one HDF5 file per rank,
 * renamed to cp_main.h5 / cp_task_<rank>.h5, with the restart image that
 * lacks the fields and the XDMF renumbering in rank 0's file (mmca_ckpt.h).
 */
static int
ckpt_write(const io_lib *out, const mmca_run_cfg *cfg, const int nsub[3], mmca_batch *bt, mmca_ctx *const *doms,
           const mmca_params *ps, const mmca_ckpt_snap *const *hist, uint64_t t)
{
    const int64_t nren = out->x ? out->get_renumber(out->x, NULL, 0) : 0;
    uint64_t *ren = malloc(sizeof(uint64_t) * (size_t) (nren > 0 ? nren : 1));
    if (!ren)
        return -1;
    if (out->x)
        out->get_renumber(out->x, ren, nren);
    const int rc = out->cwrite(cfg, nsub, bt, doms, ps, hist, t, ren, nren);
    free(ren);
    if (rc)
    {
        fprintf(stderr, "Unable to write checkpoint at timestep %lu\n", (unsigned long) t);
        return -1;
    }
    printf("Checkpoint at timestep %lu \n", (unsigned long) t);
    return 0;
}

/* Fields kept from the steps before a checkpoint (mmca_ckpt_write): ce one
 * step before it (its oce) and, with several ranks, cl, fs, dc, gr and d of
 * the last MMCA_CKPT_HISTORY steps (its halo planes).  Slot k holds step
 * checkpoint - MMCA_CKPT_HISTORY + k. */
typedef struct ckpt_hist {
    int depth;                                  /* steps kept: 1 (one rank) or MMCA_CKPT_HISTORY */
    mmca_ckpt_snap *slot[MMCA_CKPT_HISTORY];
    uint64_t step[MMCA_CKPT_HISTORY];
} ckpt_hist;

static int
hist_alloc(ckpt_hist *h, const mmca_params *ps, int nr)
{
    h->depth = nr > 1 ? MMCA_CKPT_HISTORY : 1;
    for (int k = 0; k < MMCA_CKPT_HISTORY; k++)
    {
        h->step[k] = UINT64_MAX;
        h->slot[k] = NULL;
        if (k < MMCA_CKPT_HISTORY - h->depth)
            continue;
        if (!(h->slot[k] = calloc((size_t) nr, sizeof(mmca_ckpt_snap))))
            return -1;
        for (int r = 0; r < nr; r++)
        {
            const size_t P = (size_t) (ps[r].nx + 2) * (ps[r].ny + 2) * (ps[r].nz + 2);
            mmca_ckpt_snap *c = &h->slot[k][r];
            if (!(c->ce = malloc(P * sizeof(double))))
                return -1;
            if (nr > 1 && (!(c->cl = malloc(P * sizeof(double))) || !(c->fs = malloc(P * sizeof(double))) ||
                           !(c->dc = malloc(3 * P * sizeof(double))) || !(c->d = malloc(P * sizeof(double))) ||
                           !(c->gr = malloc(P * sizeof(int32_t)))))
                return -1;
        }
    }
    return 0;
}

static void
hist_free(ckpt_hist *h, int nr)
{
    for (int k = 0; k < MMCA_CKPT_HISTORY; k++)
    {
        for (int r = 0; h->slot[k] && r < nr; r++)
        {
            mmca_ckpt_snap *c = &h->slot[k][r];
            free(c->ce);
            free(c->cl);
            free(c->fs);
            free(c->dc);
            free(c->d);
            free(c->gr);
        }
        free(h->slot[k]);
        h->slot[k] = NULL;
    }
}

/* keeps the fields after step t when a checkpoint follows within the depth */
static int
hist_take(ckpt_hist *h, mmca_ctx *const *doms, int nr, long long freq, uint64_t t)
{
    if (freq <= 0)
        return 0;
    const uint64_t cp = (t / (uint64_t) freq + 1) * (uint64_t) freq;
    if (cp - t > (uint64_t) h->depth)
        return 0;
    const int k = MMCA_CKPT_HISTORY - (int) (cp - t);
    for (int r = 0; r < nr; r++)
    {
        mmca_ckpt_snap *c = &h->slot[k][r];
        if (mmca_download_field(doms[r], MMCA_FIELD_CE, c->ce) != MMCA_OK ||
            (c->cl && mmca_download_field(doms[r], MMCA_FIELD_CL, c->cl) != MMCA_OK) ||
            (c->fs && mmca_download_field(doms[r], MMCA_FIELD_FS, c->fs) != MMCA_OK) ||
            (c->dc && mmca_download_field(doms[r], MMCA_FIELD_DC, c->dc) != MMCA_OK) ||
            (c->d && mmca_download_field(doms[r], MMCA_FIELD_D, c->d) != MMCA_OK) ||
            (c->gr && mmca_download_field(doms[r], MMCA_FIELD_GR, c->gr) != MMCA_OK))
        {
            fprintf(stderr, "download of the fields before a checkpoint failed\n");
            return -1;
        }
    }
    h->step[k] = t;
    return 0;
}

/* the next step before `next` at which hist_take keeps fields */
static uint64_t
hist_stop(const ckpt_hist *h, long long freq, uint64_t t, uint64_t next)
{
    if (freq <= 0)
        return next;
    const uint64_t cp = (t / (uint64_t) freq + 1) * (uint64_t) freq;
    for (int j = h->depth; j >= 1; j--)
        if (cp > (uint64_t) j && cp - (uint64_t) j > t)
            return cp - (uint64_t) j < next ? cp - (uint64_t) j : next;
    return next;
}

static int
ckpt_read(io_lib *out, const mmca_run_cfg *cfg, const int nsub[3], mmca_batch *bt, mmca_ctx *const *doms,
          const mmca_params *ps)
{
    const int64_t max = cfg->max_total_grains + 2;
    uint64_t *ren = malloc(sizeof(uint64_t) * (size_t) max);
    int64_t nren = 0;
    int rc = !ren || out->cread(cfg, nsub, bt, doms, ps, ren, max, &nren);
    if (rc == 0 && out->x && out->set_renumber(out->x, ren, nren))
        rc = 1;
    free(ren);
    return rc ? -1 : 0;
}

static void
dump_ctrl(const mmca_run_cfg *c, const mmca_params *p, const mmca_nucleation_cfg *n)
{
    printf("basefilename %s\n", c->basefilename);
    printf("nsub %d %d %d\n", c->nsub[0], c->nsub[1], c->nsub[2]);
    printf("finish_time %.17g\nfs_finish %.17g\nscreenpfreq %d\n", c->finish_time, c->fs_finish, c->screenpfreq);
    printf("data_write_freq %d\ndata_write_start %d\ndata_format %s\nrandomize_grains %d\n", c->data_write_freq,
           c->data_write_start, c->vis_format == MMCA_VIS_XDMF ? "XDMF" : c->vis_format == MMCA_VIS_VTK ? "VTK" : "NONE",
           c->randomize_grains);
    printf("nx %d\nny %d\nnz %d\n", p->nx, p->ny, p->nz);
    printf("pad_bndy %d %d %d\n", p->pad_bndy[0], p->pad_bndy[1], p->pad_bndy[2]);
    printf("gdim %d %d %d\n", p->gdim[0], p->gdim[1], p->gdim[2]);
    printf("cell_size %.17g\ndt %.17g\ncourant_limit %.17g\n", p->cell_size, p->dt, p->courant_limit);
    printf("c_init %.17g\nliquidus_temp %.17g\nsolidus_temp %.17g\n", p->c_init, p->liquidus_temp,
           p->solidus_temp);
    printf("m_solute %.17g %.17g %.17g\n", p->m_solute[0], p->m_solute[1], p->m_solute[2]);
    printf("part_coef %.17g\nd_sol %.17g\nd_liq %.17g\nfsgrow %.17g\n", p->part_coef, p->d_sol, p->d_liq,
           p->fsgrow);
    printf("temp_model %d\nam_layers %d\nam_tracks %d\ngnsbx %d\n", p->temp_model, p->am_layers, p->am_tracks,
           p->gnsbx);
    printf("t_init %.17g\ncool_rate %.17g\ngradient %.17g\ngrad_coef %.17g\nvelocity %.17g\nvelo_coef %.17g\n",
           p->t_init, p->cool_rate, p->gradient, p->grad_coef, p->velocity, p->velo_coef);
    printf("laser %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p->laser_power, p->laser_speed,
           p->laser_spot, p->laser_absorp, p->laser_hatch, p->sub_init_temp, p->th_conductivity, p->density,
           p->specific_heat);
    printf("laser_start %.17g %.17g %.17g\n", p->laser_start[0], p->laser_start[1], p->laser_start[2]);
    printf("max_total_grains %lld\n", (long long) p->max_total_grains);
    printf("random_seed %u\noriented %d\ndo_surface %d\n", n->random_seed, n->oriented, n->do_surface);
    printf("bulk %.17g %.17g %.17g\nsurf %.17g %.17g %.17g\n", n->bulk_density, n->bulk_mean, n->bulk_sigma,
           n->surf_density, n->surf_mean, n->surf_sigma);
    for (int i = 0; i < n->n_fixed; i++)
        printf("fixed %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", n->fixed[i].x, n->fixed[i].y, n->fixed[i].z,
               n->fixed[i].ax, n->fixed[i].ay, n->fixed[i].az, n->fixed[i].threshold);
    for (int i = 0; i < n->n_volumes; i++)
        printf("volume %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", n->volumes[i].density,
               n->volumes[i].minx, n->volumes[i].miny, n->volumes[i].minz, n->volumes[i].maxx,
               n->volumes[i].maxy, n->volumes[i].maxz, n->volumes[i].threshold);
}

/* output_grains (grain.c:752-833) on the downloaded gr field and grain table */
/* output_grains (grain.c:752-833) over the subblocks doms[0..ndom) of one
 * domain (params ps); the grain table is the one of doms[0], shared by a
 * coupled batch */
static int
output_grains(mmca_ctx *const *doms, const mmca_params *ps, int ndom, const mmca_run_cfg *c)
{
    int32_t ng = 0;
    if (check(mmca_get_grains(doms[0], NULL, 0, &ng), "mmca_get_grains"))
        return 1;
    mmca_grain *g = calloc((size_t) (ng > 0 ? ng : 1), sizeof(mmca_grain));
    size_t *sizes = calloc((size_t) (ng > 0 ? ng : 1), sizeof(size_t));
    if (!g || !sizes)
    {
        fprintf(stderr, "out of host memory\n");
        return 1;
    }
    if (check(mmca_get_grains(doms[0], g, ng, NULL), "mmca_get_grains"))
        return 1;
    for (int d = 0; d < ndom; d++)        /* calculateGrainSizes (grain.c:715-745), summed over ranks */
    {
        int64_t P = 0;
        const mmca_params q = ps[d];
        if (check(mmca_padded_cells(doms[d], &P), "mmca_padded_cells"))
            return 1;
        int32_t *gr = malloc(sizeof(int32_t) * (size_t) P);
        if (!gr)
        {
            fprintf(stderr, "out of host memory\n");
            return 1;
        }
        if (check(mmca_download_field(doms[d], MMCA_FIELD_GR, gr), "mmca_download_field"))
            return 1;
        const int sx = q.nx + 2, sxy = (q.nx + 2) * (q.ny + 2);
        for (int z = 1; z <= q.nz; z++)
            for (int y = 1; y <= q.ny; y++)
                for (int x = 1; x <= q.nx; x++)
                {
                    int id = gr[(size_t) z * sxy + (size_t) y * sx + x];
                    if (id > 0 && id < ng)
                        sizes[id]++;
                }
        free(gr);
    }
    char fname[512];
    snprintf(fname, sizeof(fname), "%s_grains.csv", c->basefilename);
    FILE *fp = fopen(fname, "w");
    if (!fp)
        fprintf(stderr, "Unable to open grain file %s for writing!\n", fname);
    else
        fprintf(fp, "Number,Nuc Timestep,Nuc Time,Nuc Threshold,X,Y,Z,CellX,CellY,CellZ,Rot1,Rot2,Rot3,"
                    "Rotation Matrix,Size mm^3\n");
    const double cell_volume = pow(c->cell_size * 1000, 3.0);       /* mm^3 */
    size_t min_size = ng > 1 ? sizes[1] : 0, max_size = min_size, sum = 0, zero = 0;
    for (int i = 1; i < ng; i++)
    {
        const mmca_grain *q = &g[i];
        if (fp)
        {
            /* realcoord2scoord (functions.c:92-107) with a zero origin offset */
            size_t cx = (size_t) ((q->nuc_x + c->cell_size * 0.1) / c->cell_size);
            size_t cy = (size_t) ((q->nuc_y + c->cell_size * 0.1) / c->cell_size);
            size_t cz = (size_t) ((q->nuc_z + c->cell_size * 0.1) / c->cell_size);
            fprintf(fp, "%d,%lu,%lg,%lg,%lg,%lg,%lg,%lu,%lu,%lu,%lg,%lg,%lg,[[%lg %lg %lg][%lg %lg %lg][%lg %lg %lg]],"
                        "%lg\n",
                    i, (unsigned long) q->nuc_timestep, q->nuc_timestep * c->dt, q->nuc_threshold, q->nuc_x,
                    q->nuc_y, q->nuc_z, cx, cy, cz, q->rotang[0], q->rotang[1], q->rotang[2], q->rotmat[0][0],
                    q->rotmat[0][1], q->rotmat[0][2], q->rotmat[1][0], q->rotmat[1][1], q->rotmat[1][2],
                    q->rotmat[2][0], q->rotmat[2][1], q->rotmat[2][2], sizes[i] * cell_volume);
        }
        if (sizes[i] < min_size)
            min_size = sizes[i];
        if (sizes[i] > max_size)
            max_size = sizes[i];
        sum += sizes[i];
        if (sizes[i] == 0)
            zero++;
    }
    if (fp)
        fclose(fp);
    printf("%d grains.  Min: %lg mm^3   Max: %lg mm^3   Average:  %lg mm^3\n", ng - 1, min_size * cell_volume,
           max_size * cell_volume, ((double) sum / ((double) ((ng - 1) - (int) zero)) * cell_volume));
    free(g);
    free(sizes);
    return 0;
}

int
main(int argc, char **argv)
{
    const char *path = NULL;
    int device = 0, fp32 = 0, ce_check = 1, dump = 0, one_block = 0, restart = 0;
    long long max_steps = -1, cp_freq = -1;
    for (int i = 1; i < argc; i++)
    {
        if (!strcmp(argv[i], "--device") && i + 1 < argc)
            device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--max-steps") && i + 1 < argc)
            max_steps = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--fp32"))
            fp32 = 1;
        else if (!strcmp(argv[i], "--no-ce-check"))
            ce_check = 0;
        else if (!strcmp(argv[i], "--one-block"))
            one_block = 1;
        else if (!strcmp(argv[i], "--dump-ctrl"))
            dump = 1;
        else if (!strcmp(argv[i], "--checkpoint-freq") && i + 1 < argc)
            cp_freq = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--restart"))
            restart = 1;
        else if (argv[i][0] != '-' && !path)
            path = argv[i];
        else
        {
            fprintf(stderr, "usage: %s <ctrl-file> [--device N] [--max-steps N] [--fp32] [--no-ce-check] "
                            "[--one-block] [--checkpoint-freq N] [--restart] [--dump-ctrl]\n", argv[0]);
            return 2;
        }
    }
    if (!path)
    {
        fprintf(stderr, "No Control file specified.\n");
        return 2;
    }
    mmca_run_cfg cfg;
    mmca_ctrl_defaults(&cfg);
    if (mmca_ctrl_read(path, &cfg))
        return 1;
    const int prec = fp32 ? MMCA_PREC_FP32DEV : MMCA_PREC_FP64;
    mmca_params p;
    mmca_nucleation_cfg nuc;
    mmca_ctrl_params(&cfg, ce_check, prec, &p);
    mmca_ctrl_nucleation(&cfg, &nuc);
    if (dump)
    {
        dump_ctrl(&cfg, &p, &nuc);
        mmca_ctrl_free(&cfg);
        return 0;
    }
    /* one subblock per MPI rank of the reference, or the whole domain */
    const int nr = one_block ? 1 : cfg.nsub[0] * cfg.nsub[1] * cfg.nsub[2];
    mmca_params *ps = calloc((size_t) nr, sizeof(mmca_params));
    mmca_nucleation_cfg *ns = calloc((size_t) nr, sizeof(mmca_nucleation_cfg));
    mmca_ctx **doms = calloc((size_t) nr, sizeof(mmca_ctx *));
    mmca_report *reps = calloc((size_t) nr, sizeof(mmca_report));
    if (!ps || !ns || !doms || !reps)
    {
        fprintf(stderr, "out of host memory\n");
        return 1;
    }
    for (int r = 0; r < nr; r++)
    {
        if (nr > 1)
            mmca_ctrl_params_rank(&cfg, ce_check, prec, r, &ps[r]);
        else
            ps[r] = p;
        ns[r] = nuc;
        ns[r].rank = r;
    }
    const double t_start = wall_now();
    mmca_batch *bt = NULL;
    if (nr > 1)
    {
        if (check(mmca_batch_create(ps, ns, nr, device, &bt), "mmca_batch_create") ||
            check(mmca_batch_couple(bt, cfg.nsub), "mmca_batch_couple"))
            return 1;
        for (int r = 0; r < nr; r++)
            if (check(mmca_batch_domain(bt, r, &doms[r]), "mmca_batch_domain"))
                return 1;
    }
    else if (check(mmca_create(&ps[0], device, &doms[0]), "mmca_create") ||
             check(mmca_init_nucleation(doms[0], &ns[0]), "mmca_init_nucleation"))
        return 1;
    /* writeData's output (loop.c:116-127): XDMF or VTK files, none by default */
    io_lib out = { 0 };
    const int write_freq = cfg.vis_format != MMCA_VIS_NONE ? cfg.data_write_freq : 0;
    if (write_freq > 0)
    {
        const int one[3] = { 1, 1, 1 };
        const int vtk = cfg.vis_format == MMCA_VIS_VTK;
        if (io_load(&out, vtk))
            return 1;
        if (vtk)
            out.v = out.vopen(&cfg, nr > 1 ? cfg.nsub : one, nr > 1 ? cfg.sbdim : p.gdim);
        else
            out.x = out.open(&cfg, nr > 1 ? cfg.nsub : one, nr > 1 ? cfg.sbdim : p.gdim);
        if (!out.x && !out.v)
            return 1;
    }
    /* checkpointfreq: the reference reads DataWriteStart into it (read_ctrl.c:594-599) */
    const long long checkpoint_freq = cp_freq >= 0 ? cp_freq : cfg.data_write_start;
    /* checkpoints in the reference's layout when libmmca_io.so is there, else
     * the restart image cp_main.mmca; the layout keeps the reference's d array
     * and its oce (ce before the last step, fetched one step before each
     * checkpoint) */
    const int one_sub[3] = { 1, 1, 1 };
    const int *lay = nr > 1 ? cfg.nsub : one_sub;
    const int h5cp = (checkpoint_freq > 0 || restart) && ckpt_load(&out) == 0;
    ckpt_hist hist;
    memset(&hist, 0, sizeof(hist));
    if (h5cp)
    {
        if (check(bt ? mmca_batch_set_tuning(bt, MMCA_TUNE_TRACK_D, 1) : mmca_set_tuning(doms[0], MMCA_TUNE_TRACK_D, 1),
                  "track d"))
            return 1;
        if (checkpoint_freq > 0 && hist_alloc(&hist, ps, nr))
        {
            fprintf(stderr, "out of host memory\n");
            return 1;
        }
    }
    else
        out.cwrite = NULL;
    uint64_t t = 0;
    if (restart)
    {
        FILE *h5 = h5cp ? fopen("cp_main.h5", "rb") : NULL;
        if (h5)
            fclose(h5);
        if ((h5 ? ckpt_read(&out, &cfg, lay, bt, doms, ps) : load_checkpoint(bt, doms[0], &out)) ||
            check(mmca_get_timestep(doms[0], &t), "mmca_get_timestep"))
            return 1;
        printf("Restarting from timestep %lu\n", (unsigned long) t);
    }
    /* the fields the next checkpoint needs from the steps before it */
    const int keep_hist = h5cp && checkpoint_freq > 0;
    if (keep_hist && hist_take(&hist, doms, nr, checkpoint_freq, t))
        return 1;
    printf("Initial Subblocks assigned.  Starting calculations.\n");
    const double total_volume = (p.gdim[0] * p.gdim[1] * p.gdim[2]) * pow(cfg.cell_size, 3);
    const double cell_vol = pow(cfg.cell_size, 3.0);
    if (cfg.screenpfreq > 0)
        printf("%10g  %10g  %8.4lf%% %8d\n", wall_now() - t_start, 0.0, 0.0, 0);
    double fs = 0.0;
    const uint64_t fin = cfg.finish_time ? finish_step(t + 1, cfg.dt, cfg.finish_time) : 0;
    int status = MMCA_OK;
    if (write_freq > 0 && !restart && data_write(&out, doms, ps, 0))
        status = MMCA_ERR_ARG;
    while (status == MMCA_OK)
    {
        uint64_t next = UINT64_MAX;
        if (cfg.screenpfreq > 0)
            next = (t / cfg.screenpfreq + 1) * cfg.screenpfreq;
        if (write_freq > 0)
        {
            /* next output step: a multiple of DataWriteFreq at or after DataWriteStart */
            uint64_t w = (t / write_freq + 1) * write_freq;
            if (cfg.data_write_start > 0 && w < (uint64_t) cfg.data_write_start)
                w = ((uint64_t) cfg.data_write_start + write_freq - 1) / write_freq * write_freq;
            if (w < next)
                next = w;
        }
        if (checkpoint_freq > 0 && (t / checkpoint_freq + 1) * checkpoint_freq < next)
            next = (t / checkpoint_freq + 1) * checkpoint_freq;
        if (keep_hist)
            next = hist_stop(&hist, checkpoint_freq, t, next);   /* stop before the checkpoint for its history */
        if (fin && (fin > t ? fin : t + 1) < next)
            next = fin > t ? fin : t + 1;
        if (max_steps >= 0 && (uint64_t) max_steps < next)
            next = (uint64_t) max_steps > t ? (uint64_t) max_steps : t + 1;
        if (next == UINT64_MAX)
            next = t + 1000;
        const int screen = cfg.screenpfreq > 0 && next % cfg.screenpfreq == 0;
        if (bt)
            status = mmca_batch_step(bt, (int) (next - t), screen, reps);
        else
            status = mmca_step(doms[0], (int) (next - t), screen, &reps[0]);
        if (status != MMCA_OK)
        {
            uint64_t failed = 0;
            for (int r = 0; r < nr; r++)
                if (reps[r].status != MMCA_OK && (failed == 0 || reps[r].failed_step < failed))
                    failed = reps[r].failed_step;
            fprintf(stderr, "step %llu: %s\n", (unsigned long long) failed, mmca_status_string(status));
            break;
        }
        t = next;
        if (screen)
        {
            double solid = 0.0;                       /* computeFS: sum over the ranks */
            for (int r = 0; r < nr; r++)
                solid += reps[r].solid_sum;
            fs = solid * cell_vol / total_volume;
            printf("%10g  %10g  %8.4lf%% %8d\n", wall_now() - t_start, t * cfg.dt, 100.0 * fs,
                   reps[0].num_grains - 1);
            fflush(stdout);
        }
        if (write_freq > 0 && (long long) t >= cfg.data_write_start && t % write_freq == 0 &&
            data_write(&out, doms, ps, t))
        {
            status = MMCA_ERR_ARG;
            break;
        }
        if (checkpoint_freq > 0 && t % checkpoint_freq == 0)
        {
            const mmca_ckpt_snap *hp[MMCA_CKPT_HISTORY];
            for (int k = 0; k < MMCA_CKPT_HISTORY; k++)
                hp[k] = (t + (uint64_t) k >= MMCA_CKPT_HISTORY &&
                         hist.step[k] == t + (uint64_t) k - MMCA_CKPT_HISTORY) ? hist.slot[k] : NULL;
            if (h5cp ? ckpt_write(&out, &cfg, lay, bt, doms, ps, hp, t) : save_checkpoint(bt, doms[0], &out, t))
            {
                status = MMCA_ERR_ARG;
                break;
            }
        }
        if (keep_hist && hist_take(&hist, doms, nr, checkpoint_freq, t))
            status = MMCA_ERR_ARG;
        if (status != MMCA_OK)
            break;
        int done = 0;
        if (cfg.finish_time && t * cfg.dt >= cfg.finish_time)
        {
            printf("\nAuto-Finish time met.\n");
            done = 1;
        }
        if (fs >= cfg.fs_finish)
        {
            printf("\nSolid Fraction termination condition met.\n");
            done = 1;
        }
        if (max_steps >= 0 && t >= (uint64_t) max_steps)
            done = 1;
        if (done)
            break;
    }
    /* the final solution, unless this step was written already */
    if (status == MMCA_OK && write_freq > 0 && t % write_freq != 0 && data_write(&out, doms, ps, t))
        status = MMCA_ERR_ARG;
    if (out.x)
    {
        if (out.flush(out.x) && status == MMCA_OK)
            status = MMCA_ERR_ARG;
        out.close(out.x);
    }
    if (out.v)
        out.vclose(out.v);
    int rc = status == MMCA_OK ? output_grains(doms, ps, nr, &cfg) : 1;
    if (bt)
        mmca_batch_destroy(bt);
    else
        mmca_destroy(doms[0]);
    hist_free(&hist, nr);
    free(ps);
    free(ns);
    free(doms);
    free(reps);
    mmca_ctrl_free(&cfg);
    if (rc == 0)
        printf("\nSimulation complete.  Shutting down.\n");
    return rc;
}
