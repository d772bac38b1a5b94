/*
 * VTK output of the C host driver in the reference's layout (io_vtk.c,
 * vtkCompress.c).  The fields come from the device in FP64 / int32
 * (mmca_download_field); each rank's .vtu is then built in memory and
 * written by its own host thread.
 */
#include <errno.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <zlib.h>

#include "mmca_vtk.h"

struct mmca_vtk {
    const mmca_run_cfg *cfg;
    int nsub[3], sbdim[3];
    int nranks;
};

/* ---- growing byte buffer ------------------------------------------------ */

typedef struct obuf {
    char *p;
    size_t n, cap;
    int failed;
} obuf;

static void
put_bytes(obuf *o, const void *d, size_t n)
{
    if (o->failed)
        return;
    if (o->n + n > o->cap)
    {
        size_t cap = o->cap ? o->cap : 4096;
        while (cap < o->n + n)
            cap *= 2;
        char *q = realloc(o->p, cap);
        if (!q)
        {
            o->failed = 1;
            return;
        }
        o->p = q;
        o->cap = cap;
    }
    memcpy(o->p + o->n, d, n);
    o->n += n;
}

static void
put_text(obuf *o, const char *s)
{
    put_bytes(o, s, strlen(s));
}

/* base64 (RFC 4648 alphabet, '=' padding), as writeBase64 (vtkCompress.c:55-69) */
static void
put_base64(obuf *o, const unsigned char *d, size_t n)
{
    static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    char q[4];
    for (size_t i = 0; i < n; i += 3)
    {
        const size_t left = n - i;
        const unsigned v = (unsigned) d[i] << 16 | (left > 1 ? (unsigned) d[i + 1] << 8 : 0u) |
                           (left > 2 ? (unsigned) d[i + 2] : 0u);
        q[0] = tab[(v >> 18) & 63];
        q[1] = tab[(v >> 12) & 63];
        q[2] = left > 1 ? tab[(v >> 6) & 63] : '=';
        q[3] = left > 2 ? tab[v & 63] : '=';
        put_bytes(o, q, 4);
    }
}

/* ---- zlib block stream (vtkCompress.c:73-129) ------------------------------
 * Equal-sized blocks, each compressed with zlib's default level; the header
 * is { blocks, block bytes, 0, compressed size of each block } in uint32,
 * written as its own base64 run before the one of the compressed data. */
typedef struct zstream {
    uint32_t *hdr;
    unsigned char *data;
    size_t n, cap;
    unsigned char *tmp;
    uLong tmp_cap;
    int failed;
} zstream;

static void
zs_block(zstream *z, const void *blk, uint32_t len)
{
    if (z->failed)
        return;
    const uint32_t nb = z->hdr ? z->hdr[0] : 0;
    uint32_t *h = realloc(z->hdr, sizeof(uint32_t) * (4 + (size_t) nb));
    if (!h)
    {
        z->failed = 1;
        return;
    }
    if (!z->hdr)
        h[0] = h[1] = h[2] = 0;
    z->hdr = h;
    if (h[1] == 0)
    {
        h[1] = len;
        z->tmp_cap = compressBound(len);
        z->tmp = malloc(z->tmp_cap);
        if (!z->tmp)
        {
            z->failed = 1;
            return;
        }
    }
    uLongf cl = z->tmp_cap;
    if (compress(z->tmp, &cl, blk, len) != Z_OK)
    {
        z->failed = 1;
        return;
    }
    h[3 + nb] = (uint32_t) cl;
    h[0] = nb + 1;
    if (z->n + cl > z->cap)
    {
        size_t cap = z->cap ? z->cap : 65536;
        while (cap < z->n + cl)
            cap *= 2;
        unsigned char *q = realloc(z->data, cap);
        if (!q)
        {
            z->failed = 1;
            return;
        }
        z->data = q;
        z->cap = cap;
    }
    memcpy(z->data + z->n, z->tmp, cl);
    z->n += cl;
}

static void
zs_emit(zstream *z, obuf *o)
{
    if (z->failed || !z->hdr)
        o->failed = 1;
    else
    {
        put_base64(o, (const unsigned char *) z->hdr, (3 + (size_t) z->hdr[0]) * sizeof(uint32_t));
        put_base64(o, z->data, z->n);
    }
    free(z->hdr);
    free(z->data);
    free(z->tmp);
    memset(z, 0, sizeof(*z));
}

/* ---- one rank's .vtu (writevtu, io_vtk.c:203-460) ------------------------- */

typedef struct vtu_job {
    const mmca_vtk *v;
    int rank, cx, cy, cz;
    int nx, ny, nz;
    const double *temp;            /* padded box */
    const int32_t *gr;             /* padded box */
    char path[700];
    int rc;
} vtu_job;

static void
build_vtu(vtu_job *j, obuf *o)
{
    const int nx = j->nx, ny = j->ny, nz = j->nz;
    const size_t sx = (size_t) nx + 2, sxy = sx * ((size_t) ny + 2);
    const double cs = j->v->cfg->cell_size;
    char line[256];
    zstream z;
    memset(&z, 0, sizeof(z));

    put_text(o, "<?xml version=\"1.0\"?>\n");
    put_text(o, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
                "compressor=\"vtkZLibDataCompressor\">\n");
    put_text(o, "  <UnstructuredGrid>\n");
    const int ncells = nx * ny * nz, npoints = (nx + 1) * (ny + 1) * (nz + 1);
    snprintf(line, sizeof(line), "      <Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\" >\n", npoints, ncells);
    put_text(o, line);
    put_text(o, "        <CellData>\n");

    /* cell arrays: one block per x slab, z fastest within it */
    const size_t slab = (size_t) ny * nz;
    double *tb = malloc(sizeof(double) * (slab ? slab : 1));
    int32_t *ib = malloc(sizeof(int32_t) * (slab ? slab : 1));
    if (!tb || !ib)
    {
        free(tb);
        free(ib);
        o->failed = 1;
        return;
    }
    put_text(o, "        <DataArray type=\"Float64\" Name=\"Temperature\" format=\"binary\">\n");
    for (int i = 1; i <= nx; i++)
    {
        size_t q = 0;
        for (int y = 1; y <= ny; y++)
            for (int k = 1; k <= nz; k++)
                tb[q++] = j->temp[k * sxy + y * sx + i];
        zs_block(&z, tb, (uint32_t) (slab * sizeof(double)));
    }
    zs_emit(&z, o);
    put_text(o, "\n        </DataArray>\n");
    put_text(o, "        <DataArray type=\"Int32\" Name=\"Grain\" format=\"binary\">\n");
    for (int i = 1; i <= nx; i++)
    {
        size_t q = 0;
        for (int y = 1; y <= ny; y++)
            for (int k = 1; k <= nz; k++)
                ib[q++] = j->gr[k * sxy + y * sx + i];
        zs_block(&z, ib, (uint32_t) (slab * sizeof(int32_t)));
    }
    zs_emit(&z, o);
    put_text(o, "\n        </DataArray>\n");
    put_text(o, "      </CellData>\n");
    free(tb);
    free(ib);

    /* points: cell-centre coordinates (scoord2realcoord, functions.c:122-140)
     * of corner indices 0..n, one block per x plane */
    put_text(o, "      <Points>\n");
    put_text(o, "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"binary\">\n");
    const size_t pslab = ((size_t) ny + 1) * ((size_t) nz + 1);
    double *pb = malloc(sizeof(double) * 3 * pslab);
    if (!pb)
    {
        o->failed = 1;
        return;
    }
    const int gx = j->cx * nx, gy = j->cy * ny, gz = j->cz * nz;
    for (int i = 0; i <= nx; i++)
    {
        size_t q = 0;
        for (int y = 0; y <= ny; y++)
            for (int k = 0; k <= nz; k++)
            {
                pb[q++] = 0.0 + (gx + i + 0.5) * cs;
                pb[q++] = 0.0 + (gy + y + 0.5) * cs;
                pb[q++] = 0.0 + (gz + k + 0.5) * cs;
            }
        zs_block(&z, pb, (uint32_t) (3 * pslab * sizeof(double)));
    }
    free(pb);
    zs_emit(&z, o);
    put_text(o, "        </DataArray>\n");
    put_text(o, "      </Points>\n");

    /* cells: hexahedra over the point grid (z fastest), offsets, types */
    put_text(o, "      <Cells>\n");
    put_text(o, "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"binary\">\n");
    int32_t *cb = malloc(sizeof(int32_t) * 8 * (slab ? slab : 1));
    if (!cb)
    {
        o->failed = 1;
        return;
    }
    const int row = nz + 1, plane = (ny + 1) * (nz + 1);
    for (int i = 0; i < nx; i++)
    {
        size_t q = 0;
        for (int y = 0; y < ny; y++)
            for (int k = 0; k < nz; k++)
            {
                const int c0 = i * plane + y * row + k;
                const int corner[8] = { c0, c0 + 1, c0 + row, c0 + row + 1, c0 + plane, c0 + plane + 1,
                                        c0 + plane + row, c0 + plane + row + 1 };
                for (int e = 0; e < 8; e++)
                    cb[q++] = corner[e];
            }
        zs_block(&z, cb, (uint32_t) (8 * slab * sizeof(int32_t)));
    }
    free(cb);
    zs_emit(&z, o);
    put_text(o, "        </DataArray>\n");
    put_text(o, "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"binary\">\n");
    int32_t *ob = malloc(sizeof(int32_t) * (size_t) (ncells ? ncells : 1));
    unsigned char *ty = malloc((size_t) (ncells ? ncells : 1));
    if (!ob || !ty)
    {
        free(ob);
        free(ty);
        o->failed = 1;
        return;
    }
    for (int n = 1; n <= ncells; n++)
        ob[n - 1] = n * 8;
    zs_block(&z, ob, (uint32_t) ((size_t) ncells * sizeof(int32_t)));
    zs_emit(&z, o);
    put_text(o, "        </DataArray>\n");
    put_text(o, "        <DataArray type=\"UInt8\" Name=\"types\" format=\"binary\">\n");
    memset(ty, 11, (size_t) ncells);             /* VTK_VOXEL, as the reference writes */
    zs_block(&z, ty, (uint32_t) ncells);
    zs_emit(&z, o);
    free(ob);
    free(ty);
    put_text(o, "        </DataArray>\n");
    put_text(o, "       </Cells>\n");
    put_text(o, "    </Piece>\n");
    put_text(o, "  </UnstructuredGrid>\n");
    put_text(o, "</VTKFile>\n");
}

static void *
vtu_thread(void *arg)
{
    vtu_job *j = arg;
    obuf o;
    memset(&o, 0, sizeof(o));
    build_vtu(j, &o);
    j->rc = -1;
    if (!o.failed)
    {
        FILE *f = fopen(j->path, "w");
        if (f)
        {
            const int ok = fwrite(o.p, 1, o.n, f) == o.n;
            if (fclose(f) == 0 && ok)
                j->rc = 0;
        }
    }
    if (j->rc)
        fprintf(stderr, "%d: Unable to write vtu file %s\n", j->rank, j->path);
    free(o.p);
    return NULL;
}

/* ---- entry points ---------------------------------------------------------- */

mmca_vtk *
mmca_vtk_open(const mmca_run_cfg *cfg, const int nsub[3], const int sbdim[3])
{
    mmca_vtk *v = calloc(1, sizeof(*v));
    if (!v)
        return NULL;
    v->cfg = cfg;
    for (int d = 0; d < 3; d++)
    {
        v->nsub[d] = nsub[d];
        v->sbdim[d] = sbdim[d];
    }
    v->nranks = nsub[0] * nsub[1] * nsub[2];
    for (int r = 0; r < v->nranks; r++)
    {
        char dir[600];
        snprintf(dir, sizeof(dir), "%s_%04d", cfg->basefilename, r);
        if (mkdir(dir, 0755) != 0 && errno != EEXIST)
        {
            fprintf(stderr, "Failed to create directory %s: %d\n", dir, errno);
            free(v);
            return NULL;
        }
    }
    return v;
}

/* writepvtu (io_vtk.c:39-88) */
static int
write_pvtu(const mmca_vtk *v, int ts)
{
    const char *b = v->cfg->basefilename;
    char fn[600];
    snprintf(fn, sizeof(fn), "%s_%d.pvtu", b, ts);
    FILE *fp = fopen(fn, "w");
    if (!fp)
    {
        fprintf(stderr, "0: Unable to open pvtu file %s\n", fn);
        return -1;
    }
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    fprintf(fp, "  <PUnstructuredGrid GhostLevel=\"0\">\n");
    fprintf(fp, "    <PCellData>\n");
    fprintf(fp, "      <PDataArray type=\"Float64\" Name=\"Temperature\"/>\n");
    fprintf(fp, "      <PDataArray type=\"Int32\" Name=\"Grain\"/>\n");
    fprintf(fp, "    </PCellData>\n");
    fprintf(fp, "    <PPoints>\n");
    fprintf(fp, "      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n");
    fprintf(fp, "    </PPoints>\n");
    for (int r = 0; r < v->nranks; r++)
        fprintf(fp, "    <Piece Source=\"%s_%04d/%d.vtu\"/>\n", b, r, ts);
    fprintf(fp, "  </PUnstructuredGrid>\n");
    fprintf(fp, "</VTKFile>\n");
    return fclose(fp) == 0 ? 0 : -1;
}

int
mmca_vtk_write(mmca_vtk *v, mmca_ctx *const *doms, const mmca_params *ps, uint64_t ts)
{
    const int nr = v->nranks, t = (int) ts;
    if (nr < 1 || write_pvtu(v, t))
        return -1;
    vtu_job *jobs = calloc((size_t) nr, sizeof(vtu_job));
    pthread_t *th = calloc((size_t) nr, sizeof(pthread_t));
    char *started = calloc((size_t) nr, 1);
    int rc = (jobs && th && started) ? 0 : -1;
    for (int r = 0; r < nr && rc == 0; r++)
    {
        vtu_job *j = &jobs[r];
        j->v = v;
        j->rank = r;
        j->cx = r % v->nsub[0];
        j->cy = (r / v->nsub[0]) % v->nsub[1];
        j->cz = r / (v->nsub[0] * v->nsub[1]);
        j->nx = ps[r].nx;
        j->ny = ps[r].ny;
        j->nz = ps[r].nz;
        const size_t P = (size_t) (j->nx + 2) * (j->ny + 2) * (j->nz + 2);
        double *temp = malloc(sizeof(double) * P);
        int32_t *gr = malloc(sizeof(int32_t) * P);
        j->temp = temp;
        j->gr = gr;
        snprintf(j->path, sizeof(j->path), "%s_%04d/%d.vtu", v->cfg->basefilename, r, t);
        if (!temp || !gr || mmca_download_field(doms[r], MMCA_FIELD_TEMP, temp) != MMCA_OK ||
            mmca_download_field(doms[r], MMCA_FIELD_GR, gr) != MMCA_OK)
            rc = -1;
    }
    for (int r = 0; r < nr && rc == 0; r++)
        if (pthread_create(&th[r], NULL, vtu_thread, &jobs[r]) == 0)
            started[r] = 1;
        else
            vtu_thread(&jobs[r]);
    for (int r = 0; r < nr && jobs; r++)
    {
        if (started && started[r])
            pthread_join(th[r], NULL);
        if (jobs[r].rc)
            rc = -1;
        free((void *) jobs[r].temp);
        free((void *) jobs[r].gr);
    }
    free(jobs);
    free(th);
    free(started);
    return rc;
}

void
mmca_vtk_close(mmca_vtk *v)
{
    free(v);
}
