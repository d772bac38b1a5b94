/*
 * Test infrastructure only (never linked into the product library).
 *
 * Link-time hook around the reference's per-step orchestrator
 * `doiteration()` (calculate.c:193).  The reference binary is linked with
 * `-Wl,--wrap=doiteration`, so its own `loop()` (loop.c:186) calls
 * `__wrap_doiteration` below, which runs the real step and then writes the
 * subblock state so the CPU restatement and the HIP path can be compared
 * with the reference bit for bit.
 *
 * Environment:
 *   MMCA_REF_DUMP_DIR    directory for outputs (required to dump anything)
 *   MMCA_REF_DUMP_STEPS  comma-separated list of timesteps for full dumps
 *                        (0 = state before the first step)
 *
 * Outputs:
 *   <dir>/hashes.txt        one line per step: FNV-1a 64 of each padded field
 *   <dir>/state_NNNNNN.bin  full padded fields at the requested steps
 *
 * Runs on several MPI ranks write, instead, one line per step and rank to
 * <dir>/ihash_rNN.txt: a word hash (ihash below) of each field's interior
 * cells, so that a coupled emulation whose edge and corner halo cells are
 * never written can be compared with each rank's subblock.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#include "globals.h"

extern SB_struct *lsp;
extern grain_t *grain_cache;

void __real_doiteration(void);
void __real_writeMainCheckpoint(void);

/*
 * The reference's main checkpoint writer reads gtasks.task, which nothing
 * allocates (checkpoint.c:1391), so every run that writes checkpoints dies
 * there on rank 0 before any rank's task file is complete.  With
 * MMCA_REF_TASK_CHECKPOINT set this wrapper skips it.  Rank 0 then writes
 * its task file to the same path (checkpoint.c:134-138 call the two writers
 * in turn, and get_rank_cp_path gives rank 0 the main file's name), which
 * replaces the main file in a run where the main writer works: what is left
 * on disk is what writeTaskCheckpoint (checkpoint.c:955-1238) writes on
 * every rank, renamed by clean_checkpoint (checkpoint.c:165-177).
 */
void
__wrap_writeMainCheckpoint(void)
{
    if (getenv("MMCA_REF_TASK_CHECKPOINT"))
        return;
    __real_writeMainCheckpoint();
}

static uint64_t
fnv1a(const void *p, size_t n)
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

static size_t
padded_cells(void)
{
    return (size_t) (bp->gsdimx + 2) * (bp->gsdimy + 2) * (bp->gsdimz + 2);
}

static int
want_full_dump(uint64_t ts)
{
    const char *s = getenv("MMCA_REF_DUMP_STEPS");
    if (!s)
        return 0;
    char buf[4096];
    strncpy(buf, s, sizeof(buf) - 1);
    buf[sizeof(buf) - 1] = 0;
    for (char *tok = strtok(buf, ","); tok; tok = strtok(NULL, ","))
        if ((uint64_t) strtoull(tok, NULL, 10) == ts)
            return 1;
    return 0;
}

static void
write_block(FILE *f, const void *p, size_t n)
{
    if (n && fwrite(p, 1, n, f) != n)
    {
        fprintf(stderr, "ref_hook: short write\n");
        exit(3);
    }
}

static void
full_dump(const char *dir, uint64_t ts)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s/state_%06lu.bin", dir, (unsigned long) ts);
    FILE *f = fopen(path, "wb");
    if (!f)
    {
        fprintf(stderr, "ref_hook: cannot open %s\n", path);
        exit(3);
    }
    size_t P = padded_cells();
    int32_t hdr[8] = { bp->gsdimx, bp->gsdimy, bp->gsdimz, bp->num_grains,
        lsp->gindex, 0, 0, 0 };
    uint64_t t = ts;
    write_block(f, "MMCAREF1", 8);
    write_block(f, hdr, sizeof(hdr));
    write_block(f, &t, sizeof(t));
    write_block(f, lsp->ce, P * sizeof(double));
    write_block(f, lsp->cl, P * sizeof(double));
    write_block(f, lsp->fs, P * sizeof(double));
    write_block(f, lsp->gr, P * sizeof(int));
    write_block(f, lsp->dc, P * sizeof(decentered_t));
    write_block(f, lsp->d, P * sizeof(double));
    write_block(f, lsp->temperature, P * sizeof(double));
    write_block(f, lsp->nuc_threshold, P * sizeof(float));
    write_block(f, lsp->mold, P * sizeof(int8_t));
    write_block(f, lsp->diff_id, (size_t) (lsp->gindex + 1) * sizeof(int));
    for (int g = 0; g < bp->num_grains; g++)
    {
        grain_t *gr = &grain_cache[g];
        write_block(f, gr->rotmat, 9 * sizeof(double));
        write_block(f, gr->rotang, 3 * sizeof(double));
        double nuc[3] = { gr->nuc_x, gr->nuc_y, gr->nuc_z };
        write_block(f, nuc, sizeof(nuc));
        write_block(f, &gr->nuc_timestep, sizeof(uint64_t));
        write_block(f, &gr->nuc_threshold, sizeof(double));
    }
    fclose(f);
}

static void
hash_line(const char *dir, uint64_t ts)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s/hashes.txt", dir);
    FILE *f = fopen(path, ts == 0 ? "w" : "a");
    if (!f)
        exit(3);
    size_t P = padded_cells();
    uint64_t hrot = 1469598103934665603ULL;
    if (bp->num_grains > 1)
        hrot = fnv1a(&grain_cache[1], (size_t) (bp->num_grains - 1) * sizeof(grain_t));
    fprintf(f, "%lu %d %d %016lx %016lx %016lx %016lx %016lx %016lx %016lx %016lx\n",
            (unsigned long) ts, bp->num_grains, lsp->gindex,
            (unsigned long) fnv1a(lsp->ce, P * sizeof(double)),
            (unsigned long) fnv1a(lsp->cl, P * sizeof(double)),
            (unsigned long) fnv1a(lsp->fs, P * sizeof(double)),
            (unsigned long) fnv1a(lsp->gr, P * sizeof(int)),
            (unsigned long) fnv1a(lsp->dc, P * sizeof(decentered_t)),
            (unsigned long) fnv1a(lsp->d, P * sizeof(double)),
            (unsigned long) fnv1a(lsp->diff_id + 1, (size_t) lsp->gindex * sizeof(int)),
            (unsigned long) hrot);
    fclose(f);
}

/* Hash of 64-bit words w_i: sum over i of mix(w_i + i * C1) modulo 2^64,
 * mix(x) = (x ^ (x >> 31)) * C2.  Order-dependent and easy to restate with
 * vector arithmetic. */
static uint64_t
ihash_add(uint64_t h, uint64_t w, uint64_t i)
{
    uint64_t x = w + i * 0x9E3779B97F4A7C15ULL;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ULL;
    return h + x;
}

/* interior cells of a padded field, raster order; `elem` bytes per cell
 * (4: zero-extended to a word; 8: one word; 24: three words) */
static uint64_t
ihash_field(const void *base, size_t elem)
{
    const int nx = bp->gsdimx, ny = bp->gsdimy, nz = bp->gsdimz;
    const unsigned char *p = (const unsigned char *) base;
    uint64_t h = 0, i = 0;
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++)
            {
                size_t idx = (size_t) x + (size_t) y * (nx + 2) + (size_t) z * (nx + 2) * (ny + 2);
                if (elem == 4)
                {
                    uint32_t v;
                    memcpy(&v, p + idx * 4, 4);
                    h = ihash_add(h, (uint64_t) v, i++);
                }
                else
                    for (size_t e = 0; e < elem / 8; e++)
                    {
                        uint64_t v;
                        memcpy(&v, p + idx * elem + 8 * e, 8);
                        h = ihash_add(h, v, i++);
                    }
            }
    return h;
}

static void
ihash_line(const char *dir, uint64_t ts)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s/ihash_r%02d.txt", dir, iproc);
    FILE *f = fopen(path, ts == 0 ? "w" : "a");
    if (!f)
        exit(3);
    uint64_t hl = 0;
    for (int i = 1; i <= lsp->gindex; i++)
        hl = ihash_add(hl, (uint64_t) (uint32_t) lsp->diff_id[i], (uint64_t) (i - 1));
    uint64_t hg = 0, i = 0;
    for (int g = 1; g < bp->num_grains; g++)
    {
        for (int q = 0; q < 9; q++)
        {
            uint64_t v;
            memcpy(&v, &grain_cache[g].rotmat[q / 3][q % 3], 8);
            hg = ihash_add(hg, v, i++);
        }
        hg = ihash_add(hg, (uint64_t) grain_cache[g].nuc_timestep, i++);
    }
    fprintf(f, "%lu %d %d %016lx %016lx %016lx %016lx %016lx %016lx %016lx\n",
            (unsigned long) ts, bp->num_grains, lsp->gindex,
            (unsigned long) ihash_field(lsp->ce, 8), (unsigned long) ihash_field(lsp->cl, 8),
            (unsigned long) ihash_field(lsp->fs, 8), (unsigned long) ihash_field(lsp->gr, 4),
            (unsigned long) ihash_field(lsp->dc, 24), (unsigned long) hl, (unsigned long) hg);
    fclose(f);
}

void
__wrap_doiteration(void)
{
    static int first = 1;
    const char *dir = getenv("MMCA_REF_DUMP_DIR");
    if (dir && nproc > 1)
    {
        if (first)
            ihash_line(dir, 0);
        first = 0;
        __real_doiteration();
        ihash_line(dir, bp->timestep);
        return;
    }
    if (first && dir)
    {
        /* state after initialisation, before step 1 */
        hash_line(dir, 0);
        if (want_full_dump(0))
            full_dump(dir, 0);
    }
    first = 0;
    __real_doiteration();
    if (dir)
    {
        hash_line(dir, bp->timestep);
        if (want_full_dump(bp->timestep))
            full_dump(dir, bp->timestep);
    }
}
