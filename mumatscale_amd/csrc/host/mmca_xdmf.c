/*
 * XDMF + HDF5 output of the C host driver (the reference's io_xdmf.c
 * layout).  The fields come from the device through mmca_io_fields
 * (cache_io_data, loop.c:74-113); the grain renumbering draws from rank 0's
 * random stream through mmca_rand, so later orientations stay those of a
 * reference run that writes the same files.
 */
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include <hdf5.h>
#include <zlib.h>

#include "mmca_xdmf.h"

struct mmca_xdmf {
    const mmca_run_cfg *cfg;
    int nsub[3], sbdim[3];     /* subblocks per axis, cells per subblock */
    int nranks;
    FILE *ts_fp;               /* time series, open for the whole run */
    size_t *renumber;          /* getRenumberArray (io_xdmf.c:344-375) */
    size_t renumber_sz;
    void *io[4];               /* host copies of all ranks: Temperature, CE, Grain, FracSolid */
    size_t io_cells;           /* padded cells each of them holds */
    float *ang;                /* grain angles X, Y, Z of the pending write, ng each */
    size_t ang_cap;
    /* the pending write, finished by a host thread while the run goes on */
    const mmca_params *ps;
    uint64_t ts;
    int ng;
    pthread_t th;
    int running, rc;
};

/* Cart coordinates of a rank (x fastest), the reference's (cx, cy, cz) */
static void
rank_coords(const mmca_xdmf *x, int r, int *cx, int *cy, int *cz)
{
    *cx = r % x->nsub[0];
    *cy = (r / x->nsub[0]) % x->nsub[1];
    *cz = r / (x->nsub[0] * x->nsub[1]);
}

/* BaseFileName without its directories (get_local_rank_file, io_xdmf.c:82-95) */
static const char *
local_base(const mmca_run_cfg *c)
{
    const char *p = strrchr(c->basefilename, '/');
    return p ? p + 1 : c->basefilename;
}

mmca_xdmf *
mmca_xdmf_open(const mmca_run_cfg *cfg, const int nsub[3], const int sbdim[3])
{
    mmca_xdmf *x = calloc(1, sizeof(*x));
    if (!x)
        return NULL;
    x->cfg = cfg;
    for (int d = 0; d < 3; d++)
    {
        x->nsub[d] = nsub[d];
        x->sbdim[d] = sbdim[d];
    }
    x->nranks = nsub[0] * nsub[1] * nsub[2];
    for (int r = 0; r < x->nranks; r++)
    {
        char dir[600];
        snprintf(dir, sizeof(dir), "%s_r%d", cfg->basefilename, r);
        if (mkdir(dir, 0755) != 0 && errno != EEXIST)
        {
            fprintf(stderr, "Failed to create directory %s:  %d\n", dir, errno);
            free(x);
            return NULL;
        }
    }
    return x;
}

/* getRenumberArray: a new random permutation of the ids whenever the grain
 * count has grown (grain 0 stays 0), from rank 0's stream */
static int
renumber_array(mmca_xdmf *x, mmca_ctx *dom0, int ng)
{
    if ((size_t) ng > x->renumber_sz)
    {
        free(x->renumber);
        x->renumber = malloc(sizeof(size_t) * (size_t) ng);
        if (!x->renumber)
            return -1;
        x->renumber_sz = (size_t) ng;
        for (int i = 0; i < ng; i++)
            x->renumber[i] = (size_t) i;
        for (int i = 1; i < ng; i++)
        {
            int picidx;
            do
            {
                int32_t r;
                if (mmca_rand(dom0, &r, 1) != MMCA_OK)
                    return -1;
                picidx = r % ng;
            }
            while (picidx == 0);
            size_t tmp = x->renumber[picidx];
            x->renumber[picidx] = x->renumber[i];
            x->renumber[i] = tmp;
        }
    }
    if (!x->renumber)
    {
        x->renumber = malloc(sizeof(size_t));
        if (!x->renumber)
            return -1;
        x->renumber[0] = 0;
    }
    return 0;
}

/* writeXDMFGrids (io_xdmf.c:141-304) */
static void
write_grids(mmca_xdmf *x, FILE *fp, uint64_t ts, const char *in)
{
    const mmca_run_cfg *c = x->cfg;
    const int nx = x->sbdim[0], ny = x->sbdim[1], nz = x->sbdim[2];
    const char *b = local_base(c);
    fprintf(fp, "%s<Grid Name=\"Full Volume at time %g\" GridType=\"Collection\" CollectionType=\"Spatial\">\n", in,
            ts * c->dt);
    fprintf(fp, "%s\t<Time Value=\"%g\" />\n", in, ts * c->dt);
    for (int i = 0; i < x->nranks; i++)
    {
        int cx, cy, cz;
        rank_coords(x, i, &cx, &cy, &cz);
        char file[700], sub[64];
        snprintf(file, sizeof(file), "%s_r%d/%05lu.h5", b, i, (unsigned long) ts);
        snprintf(sub, sizeof(sub), "/SubBlock_%dx%dx%d", cx, cy, cz);
        fprintf(fp, "%s\t<Grid Name=\"SubBlock %dx%dx%d (%d)\" GridType=\"Uniform\">\n", in, cx, cy, cz, i);
        fprintf(fp, "%s\t\t<Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%u %u %u\" />\n", in, nz + 1, ny + 1,
                nx + 1);
        fprintf(fp, "%s\t\t<Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n", in);
        fprintf(fp, "%s\t\t\t<DataItem Format=\"XML\" Dimensions=\"3\">\n", in);
        /* getSBRealBounds (functions.c:153-173): lower corner of the subblock */
        const double rx = 0.0 + (cx * nx + 0.5) * c->cell_size - c->cell_size / 2.0;
        const double ry = 0.0 + (cy * ny + 0.5) * c->cell_size - c->cell_size / 2.0;
        const double rz = 0.0 + (cz * nz + 0.5) * c->cell_size - c->cell_size / 2.0;
        fprintf(fp, "%s\t\t\t\t%e %e %e\n", in, rz, ry, rx);
        fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t\t<DataItem Format=\"XML\" Dimensions=\"3\">\n", in);
        fprintf(fp, "%s\t\t\t\t%e %e %e\n", in, c->cell_size, c->cell_size, c->cell_size);
        fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t</Geometry>\n", in);
        /* the reference writes Precision="8" (sizeof(double)) for its float data */
        fprintf(fp, "%s\t\t<Attribute Name=\"Temperature\" Center=\"Cell\">\n", in);
        fprintf(fp, "%s\t\t\t<DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"%lu\" Dimensions=\"%u %u %u\">\n",
                in, (unsigned long) sizeof(double), nz, ny, nx);
        fprintf(fp, "%s\t\t\t\tSERIAL:%s:%s/Temperature\n", in, file, sub);
        fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t</Attribute>\n", in);
        fprintf(fp, "%s\t\t<Attribute Name=\"Solute Composition\" Center=\"Cell\">\n", in);
        fprintf(fp, "%s\t\t\t<DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"%lu\" Dimensions=\"%u %u %u\">\n",
                in, (unsigned long) sizeof(double), nz, ny, nx);
        fprintf(fp, "%s\t\t\t\tSERIAL:%s:%s/CE\n", in, file, sub);
        fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t</Attribute>\n", in);
        fprintf(fp, "%s\t\t<Attribute Name=\"Grain\" Center=\"Cell\">\n", in);
        if (c->randomize_grains)
        {
            fprintf(fp, "%s\t\t\t<DataItem ItemType=\"Function\" NumberType=\"UInt\" Dimensions=\"%u %u %u\" "
                        "Function=\"$0[$1]\">\n", in, nz, ny, nx);
            fprintf(fp, "%s\t\t\t\t<DataItem Reference=\"XML\" Dimensions=\"%u %u %u\">\n", in, nz, ny, nx);
            fprintf(fp, "%s\t\t\t\t\t/Xdmf/Domain/DataItem[@Name=\"Grain Renumber\"]\n", in);
            fprintf(fp, "%s\t\t\t\t</DataItem>\n", in);
        }
        fprintf(fp, "%s\t\t\t\t<DataItem Format=\"HDF\" NumberType=\"UInt\" Dimensions=\"%u %u %u\">\n", in, nz, ny,
                nx);
        fprintf(fp, "%s\t\t\t\t\tSERIAL:%s:%s/Grain\n", in, file, sub);
        fprintf(fp, "%s\t\t\t\t</DataItem>\n", in);
        if (c->randomize_grains)
            fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t</Attribute>\n", in);
        fprintf(fp, "%s\t\t<Attribute Name=\"Grain Angle\" Center=\"Cell\" AttributeType=\"Vector\">\n", in);
        fprintf(fp, "%s\t\t\t<DataItem ItemType=\"Function\" NumberType=\"Float\" Dimensions=\"%u %u %u 3\" "
                    "Function=\"$1[$0], $2[$0], $3[$0]\">\n", in, nz, ny, nx);
        fprintf(fp, "%s\t\t\t\t<DataItem Format=\"HDF\" NumberType=\"UInt\" Dimensions=\"%u %u %u\">\n", in, nz, ny,
                nx);
        fprintf(fp, "%s\t\t\t\t\tSERIAL:%s:%s/Grain\n", in, file, sub);
        fprintf(fp, "%s\t\t\t\t</DataItem>\n", in);
        const char lbl[3] = { 'X', 'Y', 'Z' };
        for (int d = 0; d < 3; d++)
        {
            fprintf(fp, "%s\t\t\t\t<DataItem Reference=\"XML\" Dimensions=\"%u %u %u\">\n", in, nz, ny, nx);
            fprintf(fp, "%s\t\t\t\t\t/Xdmf/Domain/DataItem[@Name=\"Grain Angle %c\"]\n", in, lbl[d]);
            fprintf(fp, "%s\t\t\t\t</DataItem>\n", in);
        }
        fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t</Attribute>\n", in);
        fprintf(fp, "%s\t\t<Attribute Name=\"Fraction Solid\" Center=\"Cell\">\n", in);
        fprintf(fp, "%s\t\t\t<DataItem Format=\"HDF\" NumberType=\"Float\" Precision=\"%lu\" Dimensions=\"%u %u %u\">\n",
                in, (unsigned long) sizeof(double), nz, ny, nx);
        fprintf(fp, "%s\t\t\t\tSERIAL:%s:%s/FracSolid\n", in, file, sub);
        fprintf(fp, "%s\t\t\t</DataItem>\n", in);
        fprintf(fp, "%s\t\t</Attribute>\n", in);
        fprintf(fp, "%s\t</Grid>\n", in);
    }
    fprintf(fp, "%s</Grid>\n", in);
}

/* writeGlobalData (io_xdmf.c:377-405) */
static void
write_global_data(mmca_xdmf *x, FILE *fp, const char *in, int ng)
{
    const char *b = local_base(x->cfg);
    if (x->cfg->randomize_grains)
    {
        fprintf(fp, "%s<DataItem Name=\"Grain Renumber\" Format=\"HDF\" Dimensions=\"%d\" NumberType=\"UInt\">\n", in,
                ng);
        fprintf(fp, "%s\tSERIAL:%s_gbl.h5:/GrainRenumber", in, b);
        fprintf(fp, "\n%s</DataItem>\n", in);
    }
    const char lbl[3] = { 'X', 'Y', 'Z' };
    for (int d = 0; d < 3; d++)
    {
        fprintf(fp, "%s<DataItem Name=\"Grain Angle %c\" Format=\"HDF\" Dimensions=\"%d\" NumberType=\"Float\" "
                    "Precision=\"8\">\n", in, lbl[d], ng);
        fprintf(fp, "%s\tSERIAL:%s_gbl.h5:/GrainAngle%c\n", in, b, lbl[d]);
        fprintf(fp, "%s</DataItem>\n", in);
    }
}

/* writeXDMFTimeSeries (io_xdmf.c:307-339, 482-498): the entry is written,
 * then the closing tags, and the file position is put back before them so
 * that the next entry replaces them */
static int
write_time_series(mmca_xdmf *x, uint64_t ts, int ng)
{
    if (!x->ts_fp)
    {
        char fn[600];
        snprintf(fn, sizeof(fn), "%s_ts.xmf", x->cfg->basefilename);
        x->ts_fp = fopen(fn, "w");
        if (!x->ts_fp)
        {
            fprintf(stderr, "Unable to open file %s for writing.\n", fn);
            return -1;
        }
        FILE *f = x->ts_fp;
        fprintf(f, "<?xml version=\"1.0\" ?>\n");
        fprintf(f, "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
        fprintf(f, "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"2.0\">\n");
        fprintf(f, "\t<Domain>\n");
        fprintf(f, "\t\t<Information Name=\"Description\">\n");
        fprintf(f, "\t\t\tSimulation time series data\n");
        fprintf(f, "\t\t</Information>\n");
        fprintf(f, "\t\t<Information Name=\"Simulation Date\">\n");
        time_t now = time(NULL);
        fprintf(f, "\t\t\t%s", ctime(&now));
        fprintf(f, "\t\t</Information>\n");
        fprintf(f, "\t\t<Grid Name=\"Time Series\" GridType=\"Collection\" CollectionType=\"Temporal\">\n");
    }
    FILE *f = x->ts_fp;
    write_grids(x, f, ts, "\t\t\t");
    fpos_t pos;
    fgetpos(f, &pos);
    fprintf(f, "\t\t</Grid>\n");
    write_global_data(x, f, "\t\t", ng);
    fprintf(f, "\t</Domain>\n");
    fprintf(f, "</Xdmf>\n");
    fflush(f);
    fsetpos(f, &pos);
    return 0;
}

/* dataset creation as io_xdmf.c:416-421 and 580-595: chunked, never filled,
 * shuffled, deflate level 9 */
static hid_t
chunked(int rank, const hsize_t *dims)
{
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(pl, rank, dims);
    H5Pset_fill_time(pl, H5D_FILL_TIME_NEVER);
    H5Pset_shuffle(pl);
    H5Pset_deflate(pl, 9);
    return pl;
}

/* one dataset of a fresh file: created with the reference's properties and
 * written from memory space `mem` (H5S_ALL: the whole dataset) */
static int
put(hid_t f, const char *path, hid_t ftype, hid_t mtype, hid_t space, hid_t mem, hid_t link, hid_t pl,
    const void *data)
{
    hid_t ds = H5Dcreate(f, path, ftype, space, link, pl, H5P_DEFAULT);
    if (ds < 0)
        return -1;
    const herr_t st = H5Dwrite(ds, mtype, mem, H5S_ALL, H5P_DEFAULT, data);
    return (H5Dclose(ds) < 0 || st < 0) ? -1 : 0;
}

/* writeGlobalHeavyData (io_xdmf.c:407-468), from the pending write's
 * renumbering and angles */
static int
write_global_heavy(mmca_xdmf *x, int ng)
{
    char fn[600];
    snprintf(fn, sizeof(fn), "%s_gbl.h5", x->cfg->basefilename);
    hid_t f = H5Fcreate(fn, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0)
    {
        fprintf(stderr, "Unable to open HDF5 file %s\n", fn);
        return -1;
    }
    hsize_t n = (hsize_t) ng;
    hid_t space = H5Screate_simple(1, &n, NULL);
    hid_t pl = chunked(1, &n);
    int rc = 0;
    if (x->cfg->randomize_grains)
        rc = put(f, "/GrainRenumber", H5T_NATIVE_UINT, sizeof(size_t) == 8 ? H5T_NATIVE_UINT64 : H5T_NATIVE_UINT32,
                 space, H5S_ALL, H5P_DEFAULT, pl, x->renumber);
    const char lbl[3] = { 'X', 'Y', 'Z' };
    for (int d = 0; d < 3 && rc == 0; d++)
    {
        char name[32];
        snprintf(name, sizeof(name), "/GrainAngle%c", lbl[d]);
        rc = put(f, name, H5T_NATIVE_FLOAT, H5T_NATIVE_FLOAT, space, H5S_ALL, H5P_DEFAULT, pl,
                 x->ang + (size_t) d * (size_t) ng);
    }
    H5Pclose(pl);
    H5Sclose(space);
    if (H5Fclose(f) < 0)
        rc = -1;
    if (rc)
        fprintf(stderr, "Unable to write HDF5 file %s\n", fn);
    return rc;
}

/*
 * One dataset of a subblock file, filtered off the HDF5 call: the interior of
 * a padded FP32/int32 host copy, shuffled and deflated exactly as the
 * dataset's filter pipeline does it (H5Z shuffle: byte b of element i goes to
 * b * n + i; H5Z deflate: compress2 at level 9 into ceil(1.001 n) + 12
 * bytes, and an optional filter that fails leaves its bit set in the chunk's
 * filter mask), so that H5Dwrite_chunk stores the bytes H5Dwrite would.  The
 * datasets of all ranks are compressed at once on host threads, as the
 * reference's ranks compress their subblocks at the same time.
 */
typedef struct ztask {
    const unsigned char *padded;   /* 4-byte elements over the padded box */
    int nx, ny, nz;
    unsigned char *out;            /* the stored chunk */
    size_t out_n;
    uint32_t mask;                 /* filters skipped (bit 1: deflate) */
    int failed;
} ztask;

typedef struct zpool {
    ztask *t;
    int n;
    int next;
} zpool;

static void
ztask_run(ztask *z)
{
    const size_t nx = (size_t) z->nx, ny = (size_t) z->ny, nz = (size_t) z->nz;
    const size_t n = nx * ny * nz, nb = 4 * n;
    unsigned char *shuf = malloc(nb ? nb : 1);
    uLongf cap = (uLongf) ceil((double) nb * (double) 1.001f) + 12;
    z->out = malloc(cap);
    if (!shuf || !z->out)
    {
        free(shuf);
        z->failed = 1;
        return;
    }
    /* interior rows of the padded box, x fastest, straight into shuffled order */
    size_t i = 0;
    for (size_t k = 1; k <= nz; k++)
        for (size_t j = 1; j <= ny; j++)
        {
            const unsigned char *row = z->padded + 4 * ((k * (ny + 2) + j) * (nx + 2) + 1);
            for (size_t x = 0; x < nx; x++, i++)
                for (int b = 0; b < 4; b++)
                    shuf[(size_t) b * n + i] = row[4 * x + b];
        }
    const int st = compress2(z->out, &cap, shuf, nb, 9);
    if (st == Z_OK)
    {
        z->out_n = cap;
        z->mask = 0;
        free(shuf);
    }
    else if (st == Z_BUF_ERROR)
    {
        free(z->out);              /* deflate is optional: the shuffled bytes are stored */
        z->out = shuf;
        z->out_n = nb;
        z->mask = 1u << 1;
    }
    else
    {
        free(shuf);
        z->failed = 1;
    }
}

static void *
zpool_worker(void *arg)
{
    zpool *p = arg;
    for (;;)
    {
        const int i = __atomic_fetch_add(&p->next, 1, __ATOMIC_RELAXED);
        if (i >= p->n)
            return NULL;
        ztask_run(&p->t[i]);
    }
}

static int
io_threads(int tasks)
{
    const char *e = getenv("MMCA_IO_THREADS");
    long n = e ? atol(e) : sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1)
        n = 1;
    if (n > 64)
        n = 64;
    return (int) (n < tasks ? n : tasks);
}

/* cache_io_data (loop.c:74-113) of every rank into the host copies (field
 * order of the files: Temperature, CE, Grain, FracSolid) */
static int
fetch_fields(mmca_xdmf *x, mmca_ctx *const *doms, const mmca_params *ps)
{
    const int nr = x->nranks;
    const size_t P = (size_t) (ps[0].nx + 2) * (ps[0].ny + 2) * (ps[0].nz + 2);
    if (P * (size_t) nr > x->io_cells)
    {
        for (int v = 0; v < 4; v++)
        {
            free(x->io[v]);
            x->io[v] = malloc(4 * P * (size_t) nr);
            if (!x->io[v])
            {
                fprintf(stderr, "out of host memory\n");
                x->io_cells = 0;
                return -1;
            }
        }
        x->io_cells = P * (size_t) nr;
    }
    for (int r = 0; r < nr; r++)
    {
        float *t = (float *) x->io[0] + P * r, *ce = (float *) x->io[1] + P * r;
        int32_t *gr = (int32_t *) x->io[2] + P * r;
        float *fs = (float *) x->io[3] + P * r;
        if (mmca_io_fields(doms[r], fs, ce, t, gr) != MMCA_OK)
            return -1;
    }
    return 0;
}

/* xdmf_writeSubblocks (io_xdmf.c:559-673) for every rank, from the host
 * copies: one file each with its four datasets, created with the
 * reference's properties */
static int
write_subblocks(mmca_xdmf *x, const mmca_params *ps, uint64_t ts)
{
    const int nr = x->nranks;
    if (nr < 1)
        return -1;
    const size_t P = (size_t) (ps[0].nx + 2) * (ps[0].ny + 2) * (ps[0].nz + 2);
    const int nt = 4 * nr;
    ztask *tk = calloc((size_t) nt, sizeof(ztask));
    if (!tk)
        return -1;
    for (int r = 0; r < nr; r++)
        for (int v = 0; v < 4; v++)
        {
            ztask *z = &tk[4 * r + v];
            z->padded = (const unsigned char *) x->io[v] + 4 * P * (size_t) r;
            z->nx = ps[r].nx;
            z->ny = ps[r].ny;
            z->nz = ps[r].nz;
        }
    zpool pool = { tk, nt, 0 };
    const int nth = io_threads(nt);
    pthread_t th[64];
    int started = 0;
    for (; started < nth - 1; started++)
        if (pthread_create(&th[started], NULL, zpool_worker, &pool) != 0)
            break;
    zpool_worker(&pool);
    for (int i = 0; i < started; i++)
        pthread_join(th[i], NULL);
    int rc = 0;
    for (int i = 0; i < nt; i++)
        if (tk[i].failed)
            rc = -1;
    if (rc)
        fprintf(stderr, "out of host memory\n");
    for (int r = 0; r < nr && rc == 0; r++)
    {
        char fn[700];
        snprintf(fn, sizeof(fn), "%s_r%d/%05lu.h5", x->cfg->basefilename, r, (unsigned long) ts);
        hid_t f = H5Fcreate(fn, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        if (f < 0)
        {
            fprintf(stderr, "Unable to open HDF5 file %s\n", fn);
            rc = -1;
            break;
        }
        const hsize_t count[3] = { (hsize_t) ps[r].nz, (hsize_t) ps[r].ny, (hsize_t) ps[r].nx };
        const hsize_t origin[3] = { 0, 0, 0 };
        hid_t space = H5Screate_simple(3, count, NULL);
        hid_t link = H5Pcreate(H5P_LINK_CREATE);
        H5Pset_create_intermediate_group(link, 1);
        hid_t pl = chunked(3, count);
        int cx, cy, cz;
        rank_coords(x, r, &cx, &cy, &cz);
        const char *names[4] = { "Temperature", "CE", "Grain", "FracSolid" };
        for (int v = 0; v < 4 && rc == 0; v++)
        {
            char path[128];
            snprintf(path, sizeof(path), "/SubBlock_%dx%dx%d/%s", cx, cy, cz, names[v]);
            const ztask *z = &tk[4 * r + v];
            hid_t ds = H5Dcreate(f, path, v == 2 ? H5T_NATIVE_INT : H5T_NATIVE_FLOAT, space, link, pl, H5P_DEFAULT);
            if (ds < 0 || H5Dwrite_chunk(ds, H5P_DEFAULT, z->mask, origin, z->out_n, z->out) < 0)
                rc = -1;
            if (ds >= 0 && H5Dclose(ds) < 0)
                rc = -1;
        }
        H5Pclose(pl);
        H5Pclose(link);
        H5Sclose(space);
        if (H5Fclose(f) < 0)
            rc = -1;
        if (rc)
            fprintf(stderr, "Unable to write HDF5 file %s\n", fn);
    }
    for (int i = 0; i < nt; i++)
        free(tk[i].out);
    free(tk);
    return rc;
}

/* the part of a write that needs neither the device nor the random stream:
 * writeMain's time-series entry and global file, then the subblock files */
static void *
finish_write(void *arg)
{
    mmca_xdmf *x = arg;
    x->rc = (write_time_series(x, x->ts, x->ng) || write_global_heavy(x, x->ng) || write_subblocks(x, x->ps, x->ts))
                ? -1 : 0;
    return NULL;
}

int
mmca_xdmf_flush(mmca_xdmf *x)
{
    if (x->running)
    {
        pthread_join(x->th, NULL);
        x->running = 0;
    }
    return x->rc;
}

int
mmca_xdmf_write(mmca_xdmf *x, mmca_ctx *const *doms, const mmca_params *ps, uint64_t ts)
{
    if (mmca_xdmf_flush(x))
        return -1;
    int32_t ng = 0;
    if (mmca_get_grains(doms[0], NULL, 0, &ng) != MMCA_OK || ng < 1)
        return -1;
    /* at the write step: the renumbering's draws from rank 0's stream, the
     * grain angles and the fields */
    if (x->cfg->randomize_grains && renumber_array(x, doms[0], ng))
        return -1;
    if ((size_t) ng > x->ang_cap)
    {
        free(x->ang);
        x->ang = malloc(3 * sizeof(float) * (size_t) ng);
        x->ang_cap = x->ang ? (size_t) ng : 0;
    }
    mmca_grain *g = calloc((size_t) ng, sizeof(mmca_grain));
    if (!x->ang || !g)
    {
        free(g);
        fprintf(stderr, "out of host memory\n");
        return -1;
    }
    if (mmca_get_grains(doms[0], g, ng, NULL) != MMCA_OK)
    {
        free(g);
        return -1;
    }
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < ng; i++)
            x->ang[(size_t) d * (size_t) ng + i] = (float) g[i].rotang[d];
    free(g);
    if (fetch_fields(x, doms, ps))
        return -1;
    x->ps = ps;
    x->ts = ts;
    x->ng = ng;
    x->rc = 0;
    if (pthread_create(&x->th, NULL, finish_write, x) != 0)
    {
        finish_write(x);
        return x->rc;
    }
    x->running = 1;
    return 0;
}

void
mmca_xdmf_close(mmca_xdmf *x)
{
    if (!x)
        return;
    mmca_xdmf_flush(x);
    if (x->ts_fp)
        fclose(x->ts_fp);
    free(x->renumber);
    for (int v = 0; v < 4; v++)
        free(x->io[v]);
    free(x->ang);
    free(x);
}

int64_t
mmca_xdmf_get_renumber(const mmca_xdmf *x, uint64_t *out, int64_t max)
{
    const int64_t n = x->renumber ? (int64_t) x->renumber_sz : 0;
    for (int64_t i = 0; i < n && i < max; i++)
        out[i] = x->renumber[i];
    return n;
}

int
mmca_xdmf_set_renumber(mmca_xdmf *x, const uint64_t *in, int64_t n)
{
    free(x->renumber);
    x->renumber = NULL;
    x->renumber_sz = 0;
    if (n <= 0)
        return 0;
    x->renumber = malloc(sizeof(size_t) * (size_t) n);
    if (!x->renumber)
        return -1;
    for (int64_t i = 0; i < n; i++)
        x->renumber[i] = (size_t) in[i];
    x->renumber_sz = (size_t) n;
    return 0;
}
