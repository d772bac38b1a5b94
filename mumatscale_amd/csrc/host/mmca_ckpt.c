/*
 * Checkpoint files in the layout of the reference's writeTaskCheckpoint
 * (checkpoint.c:955-1238), written and read through the C ABI (mmca.h).
 * See mmca_ckpt.h.
 */
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <hdf5.h>
#include <zlib.h>

#include "mmca_ckpt.h"

/* FACE_* order of the neighbour table (globals.h:46-51) */
enum { F_TOP, F_BOTTOM, F_LEFT, F_RIGHT, F_FRONT, F_BACK };

/* The CoordTemp record: COMMON_SUBBLOCK_DEFS, last_active_ts and mold_size
 * of the reference's taskSubblock (checkpoint.c:966-974, globals.h:170-175),
 * in its memory layout (the dataset's type is that layout). */
typedef struct coord_rec {
    int32_t neighbors[6][2];
    uint32_t coords[3];
    int32_t subblockid;
    int32_t nnuc;
    void *nuc_pt;
    int32_t last_active_ts;
    uint64_t mold_size;
} coord_rec;

/* decentered_t (globals.h:141-144) */
typedef struct dc_rec {
    double x, y, z;
} dc_rec;

static int
fail(const char *what)
{
    fprintf(stderr, "checkpoint: %s\n", what);
    return -1;
}

/* Cart coordinates of rank r (x fastest, as sbID_from_sbcoords, functions.c:272-311) */
static void
rank_xyz(const int nsub[3], int r, int c[3])
{
    c[0] = r % nsub[0];
    c[1] = (r / nsub[0]) % nsub[1];
    c[2] = r / (nsub[0] * nsub[1]);
}

/* sbID_from_sbcoords (functions.c:272-311): -1 past a padded face, the
 * subblock at the other end of a wrapped axis */
static int
sb_id(const int nsub[3], const int pad[3], int x, int y, int z)
{
    int v[3] = { x, y, z };
    for (int a = 0; a < 3; a++)
    {
        if (pad[a])
        {
            if (v[a] < 0 || v[a] >= nsub[a])
                return -1;
        }
        else
            v[a] = (v[a] + nsub[a]) % nsub[a];
    }
    return v[0] + v[1] * nsub[0] + v[2] * nsub[0] * nsub[1];
}

/* determine_3dneighbors (functions.c:27-48) */
static void
neighbours(const int nsub[3], const int pad[3], int r, int32_t nb[6][2])
{
    int c[3];
    rank_xyz(nsub, r, c);
    memset(nb, 0, sizeof(int32_t) * 12);
    nb[F_TOP][0] = sb_id(nsub, pad, c[0], c[1], c[2] + 1);
    nb[F_BOTTOM][0] = sb_id(nsub, pad, c[0], c[1], c[2] - 1);
    nb[F_BACK][0] = sb_id(nsub, pad, c[0], c[1] + 1, c[2]);
    nb[F_FRONT][0] = sb_id(nsub, pad, c[0], c[1] - 1, c[2]);
    nb[F_RIGHT][0] = sb_id(nsub, pad, c[0] + 1, c[1], c[2]);
    nb[F_LEFT][0] = sb_id(nsub, pad, c[0] - 1, c[1], c[2]);
}

/* ---- halo planes as the reference's face exchanges leave them -------------
 * The reference's arrays carry in their halo planes what its face exchanges
 * last put there (ExchangeFacesForVar / FinishExchangeForVar,
 * calculate.c:55-191): for every face with a neighbour rank other than
 * itself, in face order, the neighbour's whole face layer, edge and corner
 * halo cells included (computeHaloInfo / computeFaceInfo, packing.c:217-300),
 * as it was when all ranks sent.  cl, fs and dc are sent at the start of a
 * step, gr after partitioning, d after grow_octahedra; kernels write no halo
 * cell.  This build exchanges face cells only, at its own points of the step,
 * so the writer rebuilds those planes from the fields of the last steps: a
 * face cell needs the neighbour's layer one exchange back, an edge cell two
 * (through the neighbour's halo), a corner three. */

/* halo plane of face f (receiver) and the neighbour's layer it receives */
static void
face_planes(int f, const int n[3], int *axis, int *halo, int *layer)
{
    switch (f)
    {
    case F_TOP: *axis = 2; *halo = n[2] + 1; *layer = 1; break;
    case F_BOTTOM: *axis = 2; *halo = 0; *layer = n[2]; break;
    case F_LEFT: *axis = 0; *halo = 0; *layer = n[0]; break;
    case F_RIGHT: *axis = 0; *halo = n[0] + 1; *layer = 1; break;
    case F_FRONT: *axis = 1; *halo = 0; *layer = n[1]; break;
    default: *axis = 1; *halo = n[1] + 1; *layer = 1; break;
    }
}

/* One exchange over all ranks: dst[r] receives, in its halo planes, the
 * face layers of src[neighbour] (all ranks send before any unpacks, so src
 * and dst are different arrays).  esize bytes per cell. */
static void
exchange_halos(const int nsub[3], const int pad[3], const int n[3], size_t esize, unsigned char *const *dst,
               unsigned char *const *src)
{
    const int nr = nsub[0] * nsub[1] * nsub[2];
    const size_t sx = (size_t) n[0] + 2, sxy = sx * ((size_t) n[1] + 2);
    const size_t stride[3] = { 1, sx, sxy };
    for (int r = 0; r < nr; r++)
    {
        int32_t nb[6][2];
        neighbours(nsub, pad, r, nb);
        for (int f = 0; f < 6; f++)
        {
            const int q = nb[f][0];
            if (q < 0 || q == r)
                continue;
            int axis, halo, layer;
            face_planes(f, n, &axis, &halo, &layer);
            const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;   /* the plane's two axes */
            for (int b = 0; b < n[v] + 2; b++)
                for (int a = 0; a < n[u] + 2; a++)
                {
                    const size_t base = (size_t) a * stride[u] + (size_t) b * stride[v];
                    memcpy(dst[r] + (base + (size_t) halo * stride[axis]) * esize,
                           src[q] + (base + (size_t) layer * stride[axis]) * esize, esize);
                }
        }
    }
}

/* interior of a (1 for cells with no coordinate in a halo layer) */
static int
is_interior(const int n[3], int x, int y, int z)
{
    return x >= 1 && x <= n[0] && y >= 1 && y <= n[1] && z >= 1 && z <= n[2];
}

/* number of coordinates of (x, y, z) in a halo layer */
static int
halo_axes(const int n[3], int x, int y, int z)
{
    return (x == 0 || x == n[0] + 1) + (y == 0 || y == n[1] + 1) + (z == 0 || z == n[2] + 1);
}

/* Rebuilds the halo of one field over all ranks.  snap[k][r] is rank r's
 * padded array after the k-th step of the history (oldest first, k <
 * nsnap; NULL where not kept), cur[r] the array after the checkpoint's step,
 * into which the result goes.  sent_at_start: the field is sent at the start
 * of a step (cl, fs, dc: the previous step's values); otherwise during the
 * step, with that step's interior (d).  own_faces: the face cells of this
 * build's arrays already hold what the reference's exchange put there (gr,
 * exchanged at the same point of the step), only edges and corners are
 * rebuilt. */
static int
rebuild_halos(const int nsub[3], const int pad[3], const int n[3], size_t esize, unsigned char *const *const *snap,
              int nsnap, unsigned char *const *cur, int sent_at_start, int own_faces)
{
    const int nr = nsub[0] * nsub[1] * nsub[2];
    const size_t P = ((size_t) n[0] + 2) * ((size_t) n[1] + 2) * ((size_t) n[2] + 2);
    int k0 = 0;
    while (k0 < nsnap && !snap[k0])
        k0++;
    if (k0 == nsnap)
        return 0;                                   /* no history: the arrays stay as they are */
    unsigned char **model = calloc((size_t) nr, sizeof(*model));
    unsigned char **next = calloc((size_t) nr, sizeof(*next));
    int rc = (model && next) ? 0 : -1;
    for (int r = 0; r < nr && rc == 0; r++)
    {
        model[r] = malloc(P * esize);
        next[r] = malloc(P * esize);
        if (!model[r] || !next[r])
            rc = -1;
        else
            memcpy(model[r], snap[k0][r], P * esize);
    }
    /* each later step: the exchange of that step, then that step's interior */
    for (int k = k0 + 1; k <= nsnap && rc == 0; k++)
    {
        unsigned char *const *after = k < nsnap ? snap[k] : cur;
        if (!after)
            break;
        for (int r = 0; r < nr; r++)
            memcpy(next[r], model[r], P * esize);
        if (!sent_at_start)
            /* sent during the step: the sender's interior is already this step's */
            for (int r = 0; r < nr; r++)
                for (int z = 1; z <= n[2]; z++)
                    for (int y = 1; y <= n[1]; y++)
                    {
                        const size_t i = (((size_t) z * (n[1] + 2) + y) * (n[0] + 2) + 1) * esize;
                        memcpy(model[r] + i, after[r] + i, (size_t) n[0] * esize);
                    }
        exchange_halos(nsub, pad, n, esize, next, model);
        for (int r = 0; r < nr; r++)
            for (int z = 0; z < n[2] + 2; z++)
                for (int y = 0; y < n[1] + 2; y++)
                    for (int x = 0; x < n[0] + 2; x++)
                    {
                        const int h = halo_axes(n, x, y, z);
                        if (h == 0 || (own_faces && h == 1))
                        {
                            const size_t i = (((size_t) z * (n[1] + 2) + y) * (n[0] + 2) + x) * esize;
                            memcpy(next[r] + i, after[r] + i, esize);
                        }
                    }
        unsigned char **t = model;
        model = next;
        next = t;
    }
    if (rc == 0)
        for (int r = 0; r < nr; r++)
            for (int z = 0; z < n[2] + 2; z++)
                for (int y = 0; y < n[1] + 2; y++)
                    for (int x = 0; x < n[0] + 2; x++)
                        if (!is_interior(n, x, y, z))
                        {
                            const size_t i = (((size_t) z * (n[1] + 2) + y) * (n[0] + 2) + x) * esize;
                            memcpy(cur[r] + i, model[r] + i, esize);
                        }
    for (int r = 0; r < nr; r++)
    {
        if (model)
            free(model[r]);
        if (next)
            free(next[r]);
    }
    free(model);
    free(next);
    return rc ? fail("out of host memory") : 0;
}

/* ---- chunks compressed off the HDF5 call ---------------------------------
 * A dataset here is one chunk (the whole array).  Its stored bytes are what
 * the dataset's filters make of the array: H5Z shuffle (byte b of element i
 * to b * n + i; nothing for 1-byte elements), then H5Z deflate (compress2
 * at level 9 into ceil(1.001 * bytes) + 12; deflate is optional, so when it
 * does not fit the shuffled bytes are stored with the filter's mask bit
 * set).  H5Dwrite_chunk then stores exactly what H5Dwrite would, and the
 * arrays of all ranks are compressed at once on host threads. */
typedef struct zjob {
    const unsigned char *data;
    size_t n, esize;
    unsigned char *out;
    size_t out_n;
    uint32_t mask;
    int failed;
} zjob;

typedef struct zqueue {
    zjob *j;
    int n, next;
} zqueue;

static void
zjob_run(zjob *z)
{
    const size_t nb = z->n * z->esize;
    unsigned char *shuf = malloc(nb ? nb : 1);
    uLongf cap = (uLongf) ceil((double) nb * (double) 1.001f) + 12;
    z->out = malloc(cap);
    if (!shuf || !z->out)
    {
        free(shuf);
        z->failed = 1;
        return;
    }
    if (z->esize > 1)
    {
        for (size_t i = 0; i < z->n; i++)
            for (size_t b = 0; b < z->esize; b++)
                shuf[b * z->n + i] = z->data[i * z->esize + b];
    }
    else
        memcpy(shuf, z->data, nb);
    const int st = compress2(z->out, &cap, shuf, nb, 9);
    if (st == Z_OK)
    {
        z->out_n = cap;
        z->mask = 0;
        free(shuf);
    }
    else if (st == Z_BUF_ERROR)
    {
        free(z->out);
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
zworker(void *arg)
{
    zqueue *q = arg;
    for (;;)
    {
        const int i = __atomic_fetch_add(&q->next, 1, __ATOMIC_RELAXED);
        if (i >= q->n)
            return NULL;
        zjob_run(&q->j[i]);
    }
}

static int
compress_all(zjob *jobs, int n)
{
    const char *e = getenv("MMCA_IO_THREADS");
    long nt = e ? atol(e) : sysconf(_SC_NPROCESSORS_ONLN);
    if (nt < 1)
        nt = 1;
    if (nt > 64)
        nt = 64;
    if (nt > n)
        nt = n;
    zqueue q = { jobs, n, 0 };
    pthread_t th[64];
    int started = 0;
    for (; started < nt - 1; started++)
        if (pthread_create(&th[started], NULL, zworker, &q) != 0)
            break;
    zworker(&q);
    for (int i = 0; i < started; i++)
        pthread_join(th[i], NULL);
    for (int i = 0; i < n; i++)
        if (jobs[i].failed)
            return -1;
    return 0;
}

/* ---- datatypes of the reference's writer --------------------------------- */
static hid_t
dc_type(void)
{
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(dc_rec));
    H5Tinsert(t, "x", HOFFSET(dc_rec, x), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "y", HOFFSET(dc_rec, y), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "z", HOFFSET(dc_rec, z), H5T_NATIVE_DOUBLE);
    return t;
}

/* storeGlobalGrainInfo (checkpoint.c:619-672): grain_t's first fields */
static hid_t
grain_type(void)
{
    hsize_t rd[2] = { 3, 3 };
    hid_t rot = H5Tarray_create(H5T_NATIVE_DOUBLE, 2, rd);
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(mmca_grain));
    H5Tinsert(t, "nuc_x", offsetof(mmca_grain, nuc_x), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "nuc_y", offsetof(mmca_grain, nuc_y), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "nuc_z", offsetof(mmca_grain, nuc_z), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "nuc_timestep", offsetof(mmca_grain, nuc_timestep), H5T_NATIVE_UINT64);
    H5Tinsert(t, "rotmat", offsetof(mmca_grain, rotmat), rot);
    H5Tclose(rot);
    return t;
}

static hid_t
coord_type(void)
{
    hid_t loc = H5Tcreate(H5T_COMPOUND, 3 * sizeof(uint32_t));
    H5Tinsert(loc, "x", 0, H5T_NATIVE_UINT32);
    H5Tinsert(loc, "y", 4, H5T_NATIVE_UINT32);
    H5Tinsert(loc, "z", 8, H5T_NATIVE_UINT32);
    hsize_t nd[2] = { 6, 2 };
    hid_t nb = H5Tarray_create(H5T_NATIVE_INT, 2, nd);
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(coord_rec));
    H5Tinsert(t, "last_active", offsetof(coord_rec, last_active_ts), H5T_NATIVE_INT);
    H5Tinsert(t, "coord", offsetof(coord_rec, coords), loc);
    H5Tinsert(t, "subblockid", offsetof(coord_rec, subblockid), H5T_NATIVE_INT);
    H5Tinsert(t, "neighbor", offsetof(coord_rec, neighbors), nb);
    H5Tinsert(t, "nnuc", offsetof(coord_rec, nnuc), H5T_NATIVE_INT);
    H5Tinsert(t, "mold_size", offsetof(coord_rec, mold_size), H5T_NATIVE_UINT64);
    H5Tclose(nb);
    H5Tclose(loc);
    return t;
}

/* nucleation_t (globals.h:147-160) */
static hid_t
nuc_type(void)
{
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(mmca_nuc_point));
    H5Tinsert(t, "x", offsetof(mmca_nuc_point, x), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "y", offsetof(mmca_nuc_point, y), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "z", offsetof(mmca_nuc_point, z), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "cx", offsetof(mmca_nuc_point, cx), H5T_NATIVE_UINT64);
    H5Tinsert(t, "cy", offsetof(mmca_nuc_point, cy), H5T_NATIVE_UINT64);
    H5Tinsert(t, "cz", offsetof(mmca_nuc_point, cz), H5T_NATIVE_UINT64);
    H5Tinsert(t, "gx", offsetof(mmca_nuc_point, gx), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "gy", offsetof(mmca_nuc_point, gy), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "gz", offsetof(mmca_nuc_point, gz), H5T_NATIVE_DOUBLE);
    H5Tinsert(t, "threshold", offsetof(mmca_nuc_point, threshold), H5T_NATIVE_DOUBLE);
    return t;
}

static hid_t
whole_chunk(int rank, const hsize_t *dims)
{
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(pl, rank, dims);
    H5Pset_fill_time(pl, H5D_FILL_TIME_NEVER);
    H5Pset_shuffle(pl);
    H5Pset_deflate(pl, 9);
    return pl;
}

/* ---- one rank's arrays ----------------------------------------------------- */
/* field order of writeTaskCheckpoint (checkpoint.c:1169-1183) */
enum { A_TEMP, A_GR, A_FS, A_CE, A_OCE, A_CL, A_DIFF, A_MOLD, A_D, A_THR, A_DC, NARR };
static const char *const ARR_NAME[NARR] = { "temperature", "gr", "fs", "ce", "oce", "cl",
                                            "diff_id", "mold", "d", "nuc_threshold", "dc" };
static const size_t ARR_ESIZE[NARR] = { 8, 4, 8, 8, 8, 8, 1, 1, 8, 4, 24 };

typedef struct rank_arrays {
    void *a[NARR];
    mmca_nuc_point *nuc;
    int32_t nnuc;
} rank_arrays;

static void
free_arrays(rank_arrays *ra)
{
    for (int v = 0; v < NARR; v++)
        free(ra->a[v]);
    free(ra->nuc);
    memset(ra, 0, sizeof(*ra));
}

static int
fetch_rank(mmca_ctx *dom, const mmca_params *p, const double *oce, rank_arrays *ra)
{
    const size_t P = (size_t) (p->nx + 2) * (p->ny + 2) * (p->nz + 2);
    const size_t N = (size_t) p->nx * p->ny * p->nz;
    for (int v = 0; v < NARR; v++)
        if (!(ra->a[v] = calloc(P, ARR_ESIZE[v])))
            return fail("out of host memory");
    if (mmca_download_field(dom, MMCA_FIELD_TEMP, ra->a[A_TEMP]) || mmca_download_field(dom, MMCA_FIELD_GR, ra->a[A_GR]) ||
        mmca_download_field(dom, MMCA_FIELD_FS, ra->a[A_FS]) || mmca_download_field(dom, MMCA_FIELD_CE, ra->a[A_CE]) ||
        mmca_download_field(dom, MMCA_FIELD_CL, ra->a[A_CL]) || mmca_download_field(dom, MMCA_FIELD_MOLD, ra->a[A_MOLD]) ||
        mmca_download_field(dom, MMCA_FIELD_D, ra->a[A_D]) ||
        mmca_download_field(dom, MMCA_FIELD_THRESHOLD, ra->a[A_THR]) ||
        mmca_download_field(dom, MMCA_FIELD_DC, ra->a[A_DC]))
        return fail("field download failed");
    memcpy(ra->a[A_OCE], oce ? (const void *) oce : ra->a[A_CE], P * sizeof(double));
    /* diff_id is written as P chars from the int array (checkpoint.c:1176) */
    const size_t nint = (P + 3) / 4 < N + 2 ? (P + 3) / 4 : N + 2;
    int32_t *slots = calloc(nint, sizeof(int32_t));
    if (!slots)
        return fail("out of host memory");
    if (mmca_get_list_slots(dom, slots, (int64_t) nint))
    {
        free(slots);
        return fail("interface list download failed");
    }
    memcpy(ra->a[A_DIFF], slots, nint * 4 < P ? nint * 4 : P);
    free(slots);
    int32_t n = 0;
    if (mmca_get_nucleation_points(dom, NULL, 0, &n))
        return fail("nucleation points unavailable");
    ra->nnuc = n;
    if (n > 0)
    {
        if (!(ra->nuc = calloc((size_t) n, sizeof(mmca_nuc_point))) ||
            mmca_get_nucleation_points(dom, ra->nuc, n, &n))
            return fail("nucleation points unavailable");
    }
    return 0;
}

static int
write_chunk(hid_t loc, const char *name, hid_t type, int rank, const hsize_t *dims, hid_t link, const zjob *z)
{
    hid_t space = H5Screate_simple(rank, dims, NULL);
    hid_t pl = whole_chunk(rank, dims);
    hid_t ds = H5Dcreate(loc, name, type, space, link, pl, H5P_DEFAULT);
    hsize_t origin[3] = { 0, 0, 0 };
    int rc = (ds < 0 || H5Dwrite_chunk(ds, H5P_DEFAULT, z->mask, origin, z->out_n, z->out) < 0) ? -1 : 0;
    if (ds >= 0 && H5Dclose(ds) < 0)
        rc = -1;
    H5Pclose(pl);
    H5Sclose(space);
    return rc;
}

static int
write_u32(hid_t f, const char *name, uint32_t v)
{
    hsize_t one = 1;
    hid_t space = H5Screate_simple(1, &one, NULL);
    hid_t ds = H5Dcreate(f, name, H5T_NATIVE_UINT32, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = (ds < 0 || H5Dwrite(ds, H5T_NATIVE_UINT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) < 0) ? -1 : 0;
    if (ds >= 0)
        H5Dclose(ds);
    H5Sclose(space);
    return rc;
}

static int
write_bytes(hid_t g, const char *name, hid_t type, const void *data, hsize_t n)
{
    hid_t space = H5Screate_simple(1, &n, NULL);
    hid_t ds = H5Dcreate(g, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = (ds < 0 || (n > 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)) ? -1 : 0;
    if (ds >= 0 && H5Dclose(ds) < 0)
        rc = -1;
    H5Sclose(space);
    return rc;
}

/* file names of get_rank_cp_path / get_rank_cp_file (checkpoint.c:113-162) */
static void
cp_path(const mmca_run_cfg *cfg, int r, uint64_t ts, char *out, size_t n)
{
    if (r == 0)
        snprintf(out, n, "cp_main_%05lu.h5", (unsigned long) ts);
    else
        snprintf(out, n, "%s_r%d/cp_task_%05lu.h5", cfg->basefilename, r, (unsigned long) ts);
}

static void
cp_file(int r, char *out, size_t n)
{
    if (r == 0)
        snprintf(out, n, "cp_main.h5");
    else
        snprintf(out, n, "cp_task_%03d.h5", r);
}

/* one rank's task file (writeTaskCheckpoint, checkpoint.c:955-1238) */
static int
write_task_file(const mmca_run_cfg *cfg, const int nsub[3], int r, const mmca_params *p, const rank_arrays *ra,
                const zjob *jobs, const zjob *grains, int64_t ngr_slots, int32_t num_grains, uint64_t ts,
                const void *image, int64_t nimage, const uint64_t *renumber, int64_t nren)
{
    char fn[800];
    cp_path(cfg, r, ts, fn, sizeof(fn));
    if (r > 0)
    {
        char dir[700];
        snprintf(dir, sizeof(dir), "%s_r%d", cfg->basefilename, r);
        if (mkdir(dir, 0755) != 0 && errno != EEXIST)
            return fail("cannot create the rank directory");
    }
    hid_t f = H5Fcreate(fn, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0)
        return fail("cannot create a checkpoint file");
    int rc = 0;
    /* the low 32 bits of the 64-bit timestep (an unsigned int view of it) */
    rc |= write_u32(f, "/curTimeStep", (uint32_t) ts);
    rc |= write_u32(f, "/g_ts_idx", 0u);                   /* temp_type INTERNAL */
    hid_t gcpl = H5Pcreate(H5P_GROUP_CREATE);
    H5Pset_link_creation_order(gcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    hid_t sub = H5Gcreate(f, "SUBBLOCK", H5P_DEFAULT, gcpl, H5P_DEFAULT);
    int c[3];
    rank_xyz(nsub, r, c);
    char gname[64];
    snprintf(gname, sizeof(gname), "SubBlock_%dx%dx%d", c[0], c[1], c[2]);
    hid_t grp = H5Gcreate(sub, gname, H5P_DEFAULT, gcpl, H5P_DEFAULT);
    hid_t link = H5Pcreate(H5P_LINK_CREATE);
    H5Pset_create_intermediate_group(link, 1);
    if (sub < 0 || grp < 0)
        rc = -1;
    /* CoordTemp */
    coord_rec rec;
    memset(&rec, 0, sizeof(rec));
    const int pad[3] = { p->pad_bndy[0], p->pad_bndy[1], p->pad_bndy[2] };
    neighbours(nsub, pad, r, rec.neighbors);
    for (int a = 0; a < 3; a++)
        rec.coords[a] = (uint32_t) c[a];
    rec.subblockid = r;
    rec.nnuc = ra->nnuc;
    hid_t ct = coord_type();
    hsize_t one = 1;
    hid_t s1 = H5Screate_simple(1, &one, NULL);
    hid_t ds = H5Dcreate(grp, "CoordTemp", ct, s1, link, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0 || H5Dwrite(ds, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, &rec) < 0)
        rc = -1;
    if (ds >= 0)
        H5Dclose(ds);
    H5Sclose(s1);
    H5Tclose(ct);
    /* the padded fields */
    const hsize_t dims[3] = { (hsize_t) p->nz + 2, (hsize_t) p->ny + 2, (hsize_t) p->nx + 2 };
    hid_t dct = dc_type();
    const hid_t types[NARR] = { H5T_NATIVE_DOUBLE, H5T_NATIVE_INT, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE,
                                H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, H5T_NATIVE_CHAR, H5T_NATIVE_CHAR,
                                H5T_NATIVE_DOUBLE, H5T_NATIVE_FLOAT, dct };
    for (int v = 0; v < NARR && rc == 0; v++)
        rc = write_chunk(grp, ARR_NAME[v], types[v], 3, dims, link, &jobs[v]);
    H5Tclose(dct);
    if (rc == 0 && ra->nnuc > 0)
    {
        hid_t nt = nuc_type();
        hsize_t nn = (hsize_t) ra->nnuc;
        hid_t sn = H5Screate_simple(1, &nn, NULL);
        hid_t dn = H5Dcreate(grp, "nuc_pt", nt, sn, link, H5P_DEFAULT, H5P_DEFAULT);
        if (dn < 0 || H5Dwrite(dn, nt, H5S_ALL, H5S_ALL, H5P_DEFAULT, ra->nuc) < 0)
            rc = -1;
        if (dn >= 0)
            H5Dclose(dn);
        H5Sclose(sn);
        H5Tclose(nt);
    }
    H5Gclose(grp);
    /* grain table and count (storeGlobalGrainInfo, checkpoint.c:619-672) */
    if (rc == 0)
    {
        hid_t gt = grain_type();
        hsize_t gd = (hsize_t) ngr_slots;
        rc = write_chunk(f, "grainCache", gt, 1, &gd, H5P_DEFAULT, grains);
        H5Tclose(gt);
        hid_t sg = H5Screate_simple(1, &one, NULL);
        hid_t dg = H5Dcreate(f, "numGrain", H5T_NATIVE_INT, sg, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (dg < 0 || H5Dwrite(dg, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &num_grains) < 0)
            rc = -1;
        if (dg >= 0)
            H5Dclose(dg);
        H5Sclose(sg);
    }
    H5Gclose(sub);
    H5Pclose(gcpl);
    H5Pclose(link);
    /* what the reference's checkpoint lacks, in rank 0's file */
    if (rc == 0 && r == 0)
    {
        hid_t m = H5Gcreate(f, "MMCA", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (m < 0 || write_bytes(m, "state", H5T_NATIVE_UCHAR, image, (hsize_t) nimage) ||
            write_bytes(m, "renumber", H5T_NATIVE_UINT64, renumber, (hsize_t) nren))
            rc = -1;
        if (m >= 0)
            H5Gclose(m);
    }
    if (H5Fclose(f) < 0)
        rc = -1;
    if (rc)
        return fail("cannot write a checkpoint file");
    /* clean_checkpoint (checkpoint.c:165-177) */
    char fin[64];
    cp_file(r, fin, sizeof(fin));
    if (rename(fn, fin) != 0)
        return fail("cannot rename a checkpoint file");
    return 0;
}

/* The halo planes of cl, fs, dc, gr and d over all ranks from the history
 * (rebuild_halos); the ranks share one subblock shape. */
static int
rank_halos(const int nsub[3], const mmca_params *p, const mmca_ckpt_snap *const *hist, rank_arrays *ra, int nr)
{
    const int n[3] = { p->nx, p->ny, p->nz };
    const int pad[3] = { p->pad_bndy[0], p->pad_bndy[1], p->pad_bndy[2] };
    /* field, bytes per cell, sent at the start of a step, face cells already right */
    static const struct { int arr; size_t esize; int at_start, own_faces; } F[5] = {
        { A_CL, 8, 1, 0 }, { A_FS, 8, 1, 0 }, { A_DC, 24, 1, 0 }, { A_GR, 4, 1, 1 }, { A_D, 8, 0, 0 } };
    unsigned char **ptr = calloc((size_t) (MMCA_CKPT_HISTORY + 1) * nr, sizeof(*ptr));
    if (!ptr)
        return fail("out of host memory");
    int rc = 0;
    for (int v = 0; v < 5 && rc == 0; v++)
    {
        unsigned char **snap[MMCA_CKPT_HISTORY];
        unsigned char **cur = ptr + (size_t) MMCA_CKPT_HISTORY * nr;
        for (int k = 0; k < MMCA_CKPT_HISTORY; k++)
        {
            snap[k] = NULL;
            if (!hist[k])
                continue;
            unsigned char **row = ptr + (size_t) k * nr;
            int whole = 1;
            for (int r = 0; r < nr; r++)
            {
                const mmca_ckpt_snap *h = &hist[k][r];
                void *a = F[v].arr == A_CL ? (void *) h->cl : F[v].arr == A_FS ? (void *) h->fs
                        : F[v].arr == A_DC ? (void *) h->dc : F[v].arr == A_GR ? (void *) h->gr : (void *) h->d;
                row[r] = a;
                whole = whole && a;
            }
            if (whole)
                snap[k] = row;
        }
        for (int r = 0; r < nr; r++)
            cur[r] = ra[r].a[F[v].arr];
        rc = rebuild_halos(nsub, pad, n, F[v].esize, (unsigned char *const *const *) snap, MMCA_CKPT_HISTORY,
                           cur, F[v].at_start, F[v].own_faces);
    }
    free(ptr);
    return rc;
}

int
mmca_ckpt_write(const mmca_run_cfg *cfg, const int nsub[3], mmca_batch *bt, mmca_ctx *const *doms,
                const mmca_params *ps, const mmca_ckpt_snap *const *hist, uint64_t ts, const uint64_t *renumber,
                int64_t nren)
{
    const mmca_ckpt_snap *last = hist ? hist[MMCA_CKPT_HISTORY - 1] : NULL;
    const int nr = nsub[0] * nsub[1] * nsub[2];
    if (nr < 1 || (!bt && nr != 1))
        return fail("bad rank layout");
    /* the run's image without its fields */
    int64_t nimage = 0;
    if ((bt ? mmca_batch_save_state_ex(bt, MMCA_STATE_NO_FIELDS, NULL, 0, &nimage)
            : mmca_save_state_ex(doms[0], MMCA_STATE_NO_FIELDS, NULL, 0, &nimage)) != MMCA_OK)
        return fail("restart image unavailable");
    unsigned char *image = malloc((size_t) nimage);
    if (!image)
        return fail("out of host memory");
    if ((bt ? mmca_batch_save_state_ex(bt, MMCA_STATE_NO_FIELDS, image, nimage, &nimage)
            : mmca_save_state_ex(doms[0], MMCA_STATE_NO_FIELDS, image, nimage, &nimage)) != MMCA_OK)
    {
        free(image);
        return fail("restart image unavailable");
    }
    /* grain table: every slot up to MaxTotGrains (unused slots zero) */
    const int64_t nslots = ps[0].max_total_grains + 1;
    mmca_grain *gtab = calloc((size_t) nslots, sizeof(mmca_grain));
    int32_t ng = 0;
    int rc = 0;
    if (!gtab || mmca_get_grains(doms[0], gtab, (int32_t) (nslots < INT32_MAX ? nslots : INT32_MAX), &ng) != MMCA_OK)
        rc = fail("grain table unavailable");
    rank_arrays *ra = calloc((size_t) nr, sizeof(rank_arrays));
    zjob *jobs = calloc((size_t) nr * NARR + 1, sizeof(zjob));
    if (!ra || !jobs)
        rc = fail("out of host memory");
    for (int r = 0; r < nr && rc == 0; r++)
    {
        rc = fetch_rank(doms[r], &ps[r], last ? last[r].ce : NULL, &ra[r]);
        const size_t P = (size_t) (ps[r].nx + 2) * (ps[r].ny + 2) * (ps[r].nz + 2);
        for (int v = 0; v < NARR && rc == 0; v++)
        {
            zjob *z = &jobs[(size_t) r * NARR + v];
            z->data = ra[r].a[v];
            z->n = P;
            z->esize = ARR_ESIZE[v];
        }
    }
    if (rc == 0 && nr > 1 && hist)
        rc = rank_halos(nsub, &ps[0], hist, ra, nr);
    zjob *gj = jobs ? &jobs[(size_t) nr * NARR] : NULL;
    if (rc == 0)
    {
        gj->data = (const unsigned char *) gtab;
        gj->n = (size_t) nslots;
        gj->esize = sizeof(mmca_grain);
        if (compress_all(jobs, nr * NARR + 1))
            rc = fail("out of host memory");
    }
    for (int r = 0; r < nr && rc == 0; r++)
        rc = write_task_file(cfg, nsub, r, &ps[r], &ra[r], &jobs[(size_t) r * NARR], gj, nslots, ng, ts, image,
                             nimage, renumber, nren);
    if (jobs)
        for (int i = 0; i < nr * NARR + 1; i++)
            free(jobs[i].out);
    if (ra)
        for (int r = 0; r < nr; r++)
            free_arrays(&ra[r]);
    free(ra);
    free(jobs);
    free(gtab);
    free(image);
    return rc;
}

/* ---- restore ---------------------------------------------------------------- */
static int
read_all(hid_t loc, const char *name, hid_t mtype, void *out, size_t elems)
{
    hid_t ds = H5Dopen(loc, name, H5P_DEFAULT);
    if (ds < 0)
        return -1;
    hid_t space = H5Dget_space(ds);
    hssize_t n = H5Sget_simple_extent_npoints(space);
    int rc = (n < 0 || (size_t) n != elems || H5Dread(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) ? -1 : 0;
    H5Sclose(space);
    H5Dclose(ds);
    return rc;
}

static int64_t
dataset_size(hid_t loc, const char *name)
{
    hid_t ds = H5Dopen(loc, name, H5P_DEFAULT);
    if (ds < 0)
        return -1;
    hid_t space = H5Dget_space(ds);
    hssize_t n = H5Sget_simple_extent_npoints(space);
    H5Sclose(space);
    H5Dclose(ds);
    return (int64_t) n;
}

int
mmca_ckpt_read(const mmca_run_cfg *cfg, const int nsub[3], mmca_batch *bt, mmca_ctx *const *doms,
               const mmca_params *ps, uint64_t *renumber, int64_t max, int64_t *nren)
{
    (void) cfg;
    const int nr = nsub[0] * nsub[1] * nsub[2];
    if (nr < 1 || (!bt && nr != 1))
        return fail("bad rank layout");
    hid_t f0 = H5Fopen("cp_main.h5", H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f0 < 0)
        return fail("cp_main.h5 is missing or not an HDF5 file");
    int rc = 0;
    const int64_t nimage = dataset_size(f0, "/MMCA/state");
    const int64_t nr_ren = dataset_size(f0, "/MMCA/renumber");
    unsigned char *image = nimage > 0 ? malloc((size_t) nimage) : NULL;
    if (nimage <= 0 || nr_ren < 0 || !image || read_all(f0, "/MMCA/state", H5T_NATIVE_UCHAR, image, (size_t) nimage))
        rc = fail("cp_main.h5 holds no restart image of this driver");
    if (rc == 0 && nr_ren > max)
        rc = fail("renumbering longer than expected");
    if (rc == 0 && nr_ren > 0 && read_all(f0, "/MMCA/renumber", H5T_NATIVE_UINT64, renumber, (size_t) nr_ren))
        rc = fail("cannot read the renumbering");
    H5Fclose(f0);
    if (rc == 0 && (bt ? mmca_batch_load_state(bt, image, nimage) : mmca_load_state(doms[0], image, nimage)) != MMCA_OK)
        rc = fail("the restart image does not fit this run");
    free(image);
    if (nren)
        *nren = nr_ren;
    for (int r = 0; r < nr && rc == 0; r++)
    {
        if (ps[r].precision != MMCA_PREC_FP64)
            continue;                                /* FP32-deviation fields are in the image */
        char fn[64];
        cp_file(r, fn, sizeof(fn));
        hid_t f = H5Fopen(fn, H5F_ACC_RDONLY, H5P_DEFAULT);
        if (f < 0)
        {
            rc = fail("a rank's checkpoint file is missing");
            break;
        }
        int c[3];
        rank_xyz(nsub, r, c);
        char g[96];
        snprintf(g, sizeof(g), "/SUBBLOCK/SubBlock_%dx%dx%d", c[0], c[1], c[2]);
        hid_t grp = H5Gopen(f, g, H5P_DEFAULT);
        const size_t P = (size_t) (ps[r].nx + 2) * (ps[r].ny + 2) * (ps[r].nz + 2);
        double *buf = malloc(P * sizeof(dc_rec));
        hid_t dct = dc_type();
        if (grp < 0 || !buf)
            rc = fail("a rank's checkpoint file lacks its subblock");
        const struct { const char *name; int field; hid_t t; } io[] = {
            { "ce", MMCA_FIELD_CE, H5T_NATIVE_DOUBLE }, { "cl", MMCA_FIELD_CL, H5T_NATIVE_DOUBLE },
            { "fs", MMCA_FIELD_FS, H5T_NATIVE_DOUBLE }, { "gr", MMCA_FIELD_GR, H5T_NATIVE_INT },
            { "dc", MMCA_FIELD_DC, dct }, { "d", MMCA_FIELD_D, H5T_NATIVE_DOUBLE } };
        for (size_t i = 0; i < sizeof(io) / sizeof(io[0]) && rc == 0; i++)
            if (read_all(grp, io[i].name, io[i].t, buf, P) || mmca_upload_field(doms[r], io[i].field, buf) != MMCA_OK)
                rc = fail("cannot restore a field");
        /* the list array's slots past the list, which later checkpoints show
         * (diff_id holds the array's first P bytes) */
        const size_t N = (size_t) ps[r].nx * ps[r].ny * ps[r].nz;
        const size_t nint = P / 4 < N + 2 ? P / 4 : N + 2;
        if (rc == 0 && (read_all(grp, "diff_id", H5T_NATIVE_CHAR, buf, P) ||
                        mmca_set_list_slots(doms[r], (const int32_t *) buf, (int64_t) nint) != MMCA_OK))
            rc = fail("cannot restore the interface list array");
        H5Tclose(dct);
        free(buf);
        if (grp >= 0)
            H5Gclose(grp);
        H5Fclose(f);
    }
    return rc;
}
