/*
 * mmca.h — C ABI of the MI355X-native cellular-automaton (CA) step of the
 * muMatScale solidification model.
 *
 * The reference has no plugin system: its operator API is the set of C
 * functions that `doiteration()` (calculate.c:193-352) and `loop()`
 * (loop.c:186-272) call on one subblock, passing state through globals
 * (`bp`, globals.h:422; `lsp`, calculate.c:40; `grain_cache`, grain.c:49).
 * This header replaces those globals by an opaque context that owns the
 * device buffers of one subblock, and exports one entry point per reference
 * operator plus a fused multi-step entry point.  Every function returns an
 * `int` status (MMCA_OK or an MMCA_ERR_* code) where the reference calls
 * `error()` (debug.h:54-58, which aborts the MPI job).
 *
 * Layout convention (same as the reference, globals.h:71-74, growth.c:70):
 * host arrays are padded boxes of (nx+2)*(ny+2)*(nz+2) cells, x fastest,
 * idx = i + j*(nx+2) + k*(nx+2)*(ny+2), interior i in [1,nx] etc.
 *
 * Threading: a context is used by one host thread at a time; all work is
 * ordered on the context's single HIP stream.
 */
#ifndef MMCA_H
#define MMCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMCA_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------- */
#define MMCA_OK                    0
#define MMCA_ERR_ARG               1  /* bad argument / size                  */
#define MMCA_ERR_HIP               2  /* HIP runtime error                    */
#define MMCA_ERR_COURANT           3  /* growth.c:43-50  rs|rl > CourantLimit */
#define MMCA_ERR_NEGATIVE_CE       4  /* growth.c:163-165 ce < 0 (CPU build)  */
#define MMCA_ERR_GRAIN_CAPACITY    5  /* grain.c:361-366 grain id > MaxTotGrains */
#define MMCA_ERR_NUC_THRESHOLD     6  /* grain.c:141-154 rejection loop does not end */
#define MMCA_ERR_NO_DEVICE         7  /* no usable GPU                        */
#define MMCA_ERR_STATE             8  /* call out of order                    */

/* ---- faces, order of the reference FACE_* constants (globals.h:46-51) ---- */
#define MMCA_FACE_TOP    0  /* +z */
#define MMCA_FACE_BOTTOM 1  /* -z */
#define MMCA_FACE_LEFT   2  /* -x */
#define MMCA_FACE_RIGHT  3  /* +x */
#define MMCA_FACE_FRONT  4  /* -y */
#define MMCA_FACE_BACK   5  /* +y */

/* ---- temperature models (temperature.c:24-126) --------------------------- */
#define MMCA_TEMP_UNIFORM   0  /* T = T0 - t*rate           (:36-39)  */
#define MMCA_TEMP_GRADIENT  1  /* moving linear gradient      (:41-61)  */
#define MMCA_TEMP_AM_BEAM   2  /* Gaussian-beam time integral (:62-123) */

/* ---- storage precision of the solute fields ------------------------------ */
#define MMCA_PREC_FP64      0  /* ce, cl, fs in FP64: bit-exact path */
#define MMCA_PREC_FP32DEV   1  /* ce, cl stored in FP32 as deviations from c0 */

/*
 * Parameters of one subblock: the part of BB_struct (globals.h:273-403) the
 * CA step reads, with the same meaning and units (SI).  Derived constants
 * are folded by the library exactly as the reference folds them
 * (growth.c:38-40, 181-189; read_ctrl.c:551).
 */
typedef struct mmca_params {
    int32_t nx, ny, nz;            /* NCellsPerSB                           */
    int32_t pad_bndy[3];           /* FaceCtrl per axis x,y,z: 1 = pad (mold wall), 0 = wrap */
    int32_t sb_coords[3];          /* subblock coordinates (0,0,0 for one block) */
    int32_t gdim[3];               /* global domain size in cells           */
    double  cell_size;             /* CellSize                              */
    double  origin[3];             /* origin_offset                         */
    double  dt;                    /* TimeStep                              */
    double  courant_limit;         /* CourantLimit (default 0.2)            */
    /* alloy */
    double  c_init;                /* Cinit                                 */
    double  liquidus_temp;         /* LiquidusTemp                          */
    double  solidus_temp;          /* SolidusTemp                           */
    double  m_solute[3];           /* m_solute0 a b c; slope used = c       */
    double  part_coef;             /* part_coef0 c                          */
    double  d_sol, d_liq;          /* Dsol, Dliq                            */
    double  fsgrow;                /* FSGrow (rev_fsgrow = 1.0/FSGrow)      */
    /* temperature */
    int32_t temp_model;            /* MMCA_TEMP_*                           */
    int32_t am_layers;             /* NoOfLayers                            */
    int32_t am_tracks;             /* NoOfTracks                            */
    int32_t gnsbx;                 /* subblocks along x (AM track length)   */
    double  t_init;                /* InitialTemperature                    */
    double  cool_rate;             /* CoolingRate                           */
    double  gradient, grad_coef;   /* Gradient, Grad_Coef                   */
    double  velocity, velo_coef;   /* Velocity, Velo_Coef                   */
    double  laser_power, laser_speed, laser_spot, laser_absorp;
    double  laser_hatch, sub_init_temp, th_conductivity, density, specific_heat;
    double  laser_start[3];        /* LaserStartLocation                    */
    /* grains */
    int64_t max_total_grains;      /* MaxTotGrains                          */
    int32_t check_negative_ce;     /* 1: CPU-build abort on ce<0 (growth.c:163) */
    int32_t precision;             /* MMCA_PREC_*                           */
    uint64_t timestep;             /* starting timestep (0 for a fresh run) */
} mmca_params;

/* FixedNuc entry (read_ctrl.c:298-310): position [m], angles [deg]. */
typedef struct mmca_fixed_nuc {
    double x, y, z;
    double ax, ay, az;
    double threshold;
} mmca_fixed_nuc;

/* NucVolume entry (read_ctrl.c:311-330). */
typedef struct mmca_nuc_volume {
    double density;
    double minx, miny, minz, maxx, maxy, maxz;
    double threshold;
} mmca_nuc_volume;

/*
 * Nucleation set-up (init_nucleation_sites grain.c:603-713 and
 * init_sb_nucleation grain.c:931-1010).  Lists are given in control-file
 * order; the library applies the reference's head-insertion order
 * (ll.c:59) itself.
 */
typedef struct mmca_nucleation_cfg {
    uint32_t random_seed;          /* RandSeedVal; stream = srand(rank+1+seed) */
    int32_t  rank;                 /* MPI rank whose stream this block uses */
    int32_t  oriented;             /* GNOriented                            */
    int32_t  do_surface;           /* MouldNuc (default 1)                  */
    double   bulk_density, bulk_mean, bulk_sigma;   /* MaxGrainDensity, GNGaussCentre/Sigma */
    double   surf_density, surf_mean, surf_sigma;   /* ...Surf                   */
    int32_t  n_fixed;
    int32_t  n_volumes;
    const mmca_fixed_nuc  *fixed;
    const mmca_nuc_volume *volumes;
} mmca_nucleation_cfg;

/* Grain record, same layout and meaning as grain_t (globals.h:244-254). */
typedef struct mmca_grain {
    double   nuc_x, nuc_y, nuc_z;
    uint64_t nuc_timestep;
    double   nuc_threshold;
    double   rotmat[3][3];
    double   rotang[3];
} mmca_grain;

/* Per-call report of mmca_step. */
typedef struct mmca_report {
    uint64_t timestep;        /* bp->timestep after the call               */
    int32_t  num_grains;      /* bp->num_grains (includes grain 0)         */
    int32_t  new_grains;      /* grains created during the call            */
    int32_t  gindex;          /* interface-list length of the last step    */
    int32_t  status;          /* MMCA_OK or error of the failing step      */
    uint64_t failed_step;     /* timestep at which status was raised       */
    double   solid_sum;       /* sum of min(1,fs) over the interior (growth.c:541-566), if requested */
    double   solute_sum;      /* sum of ce over the interior, if requested */
    double   solute_ledger;   /* FP32 storage: solute lost to rounding since creation (compensated
                                 FP64 sum); sum + ledger is the conserved total.  0 for FP64 */
} mmca_report;

/* Field identifiers for upload/download (padded host arrays). */
#define MMCA_FIELD_CE         0  /* double */
#define MMCA_FIELD_CL         1  /* double */
#define MMCA_FIELD_FS         2  /* double */
#define MMCA_FIELD_GR         3  /* int32  */
#define MMCA_FIELD_DC         4  /* double[3] per cell, packed x,y,z (decentered_t) */
#define MMCA_FIELD_D          5  /* double: half-diagonal of solid cells (upload: the array MMCA_TUNE_TRACK_D keeps) */
#define MMCA_FIELD_TEMP       6  /* double, download only: temperature at current timestep (the
                                    separable form on the interior with MMCA_TUNE_BEAM_SEP) */
#define MMCA_FIELD_THRESHOLD  7  /* float, download only: INFINITY where no site */
#define MMCA_FIELD_MOLD       8  /* int8, download only */
#define MMCA_FIELD_OCE        9  /* double: ce before the last diffusion */

typedef struct mmca_ctx mmca_ctx;

/* Library / device */
int  mmca_abi_version(void);
const char *mmca_status_string(int status);
int  mmca_device_count(int *count);

/*
 * y[i] = exp(x[i]) on `device` as the Gaussian-beam temperature evaluates it
 * (temperature.c:108-112 calls the C library's exp 400 times per cell): a
 * restatement of glibc's double exp, bit-identical with the host library.
 * Host arrays; no counterpart in the reference (a check of that exp).
 */
int  mmca_eval_exp(int device, const double *x, double *y, int64_t n);

/*
 * Achievable HBM bandwidth for the solute kernel's field mixes: a streaming
 * kernel with no arithmetic over `cells` cells (1 .. 2^40) of freshly
 * allocated buffers, one contiguous run of 1024 cells per workgroup, timed
 * with HIP events over `reps` launches after one warm-up launch.
 * mix 0: copy of one f64 field (16 B/cell); 1: the FP64 fields (read ce, cl,
 * fs f64 and gr i32, write ce, cl, fs: 52 B/cell); 2: the FP32-deviation
 * fields (read ce, cl f32, fs f64, gr i32, write ce, cl, fs: 36 B/cell).
 * Returns the average ms per launch and the bytes per cell.  No counterpart
 * in the reference (a measurement for the roofline report).
 */
int  mmca_bandwidth_probe(int device, int mix, int64_t cells, int reps, double *ms, double *bytes_per_cell);

/*
 * allocateFields + SetupInitialConditions (setup.c:113-236, 42-96) and
 * tempPrepareSB (temperature.c:156-215): ce = cl = Cinit, thresholds = INF,
 * other fields zero, mold on padded faces.
 */
int  mmca_create(const mmca_params *params, int device, mmca_ctx **out);
int  mmca_destroy(mmca_ctx *ctx);

/* init_nucleation_sites + init_sb_nucleation (grain.c:603-713, 931-1010). */
int  mmca_init_nucleation(mmca_ctx *ctx, const mmca_nucleation_cfg *cfg);

/*
 * `loop()` body without IO (loop.c:222-227): for each of nsteps,
 * timestep++ then doiteration().  The sums in the report are computed after
 * the last step when want_sums != 0 (computeFS, loop.c:173-182).
 */
int  mmca_step(mmca_ctx *ctx, int nsteps, int want_sums, mmca_report *report);

/* --- single operators of doiteration(), for stepping phase by phase ------ */
int  mmca_set_timestep(mmca_ctx *ctx, uint64_t timestep);          /* bp->timestep       */
int  mmca_cell_nucleation(mmca_ctx *ctx, int32_t *n_new);            /* grain.c:1222-1318  */
int  mmca_activate_new_grains(mmca_ctx *ctx);                        /* grain.c:841-916    */
int  mmca_sb_diffuse_alloy_decentered(mmca_ctx *ctx);               /* growth.c:30-170    */
int  mmca_fs_change_diffuse(mmca_ctx *ctx);                          /* growth.c:177-276   */
int  mmca_grow_octahedra(mmca_ctx *ctx);                             /* growth.c:339-407   */
int  mmca_grow_cell_reduction(mmca_ctx *ctx, int32_t *gindex);       /* growth.c:461-538   */
int  mmca_capture_octahedra_diffuse(mmca_ctx *ctx);                  /* growth.c:574-807   */
int  mmca_doiteration(mmca_ctx *ctx);                                /* calculate.c:193-352 */
int  mmca_solid_volume(mmca_ctx *ctx, double *volume);               /* growth.c:971-978   */

/* --- state transfer ------------------------------------------------------- */
int  mmca_upload_field(mmca_ctx *ctx, int field, const void *host);
int  mmca_download_field(mmca_ctx *ctx, int field, void *host);
int  mmca_get_grains(mmca_ctx *ctx, mmca_grain *out, int32_t max, int32_t *num_grains);
int  mmca_get_interface_list(mmca_ctx *ctx, int32_t *out, int32_t max, int32_t *gindex);
/*
 * Slots 0..n-1 of the interface-list array as they are (n <= cells + 2):
 * slots 1..gindex hold the list, slots past it what earlier, longer lists
 * left there, and slot 0 is never written -- as in the reference's diff_id
 * array (growth.c:461-538), which its checkpoint stores whole
 * (checkpoint.c:1169).
 */
int  mmca_get_list_slots(mmca_ctx *ctx, int32_t *out, int64_t n);
/* Restores slots gindex+1 .. n-1 (the stale part) from in[gindex+1 .. n-1];
 * slots 0..gindex are left as they are (n <= cells + 2). */
int  mmca_set_list_slots(mmca_ctx *ctx, const int32_t *in, int64_t n);
/*
 * The subblock's FixedNuc and NucVolume points (nucleation_t, globals.h:147-160:
 * position [m], global cell, growth direction, threshold) in the order
 * init_nucleation_sites gives them to the subblock (grain.c:603-713), as its
 * checkpoint stores them (nuc_pt, checkpoint.c:1188-1204).  Copies at most
 * max points to out and sets *n to their number.
 */
typedef struct mmca_nuc_point {
    double x, y, z;
    uint64_t cx, cy, cz;
    double gx, gy, gz;
    double threshold;
} mmca_nuc_point;
int  mmca_get_nucleation_points(mmca_ctx *ctx, mmca_nuc_point *out, int32_t max, int32_t *n);
/*
 * Replace the grain table by records 0..num_grains-1 and set num_grains, as
 * a restart restores grain_cache and bp->num_grains (addGrainCache,
 * checkpoint.c:676-745).  Record 0 stands for the liquid; rotation
 * matrices are used as given.  The next nucleated grain gets id num_grains.
 * Cells refer to grains through the gr field (mmca_upload_field).
 */
int  mmca_set_grains(mmca_ctx *ctx, const mmca_grain *grains, int32_t num_grains);
/*
 * Restart image of a block: fields, counters, interface list and dc_tmp
 * slots, grain table, and the random stream after the orientations
 * consumed so far.  The reference's restart (checkpoint.c:955-1238,
 * mainRestart/taskRestart) stores neither the random stream nor dc_tmp, so
 * a restarted run diverges from an uninterrupted one; a block restored with
 * mmca_load_state continues bit for bit.  mmca_save_state writes at most
 * `capacity` bytes to host memory `buf` and the size to *size; with
 * buf == NULL it only reports the size.  mmca_load_state needs a context
 * created and given its nucleation set-up with the same parameters.
 */
int  mmca_save_state(mmca_ctx *ctx, void *buf, int64_t capacity, int64_t *size);
int  mmca_load_state(mmca_ctx *ctx, const void *buf, int64_t size);
/*
 * As mmca_save_state, with flags.  MMCA_STATE_NO_FIELDS leaves ce, cl, fs,
 * gr and dc out of the image of an FP64 block (for a caller that stores them
 * itself, as the checkpoint files in the reference's layout do): after
 * mmca_load_state of such an image, mmca_upload_field must restore them.
 * FP32-deviation blocks keep their fields in the image.
 */
#define MMCA_STATE_NO_FIELDS 1
int  mmca_save_state_ex(mmca_ctx *ctx, int flags, void *buf, int64_t capacity, int64_t *size);
int  mmca_get_timestep(mmca_ctx *ctx, uint64_t *timestep);
/*
 * cache_io_data (loop.c:74-113) and the host copies the output writers read
 * (io_xdmf.c:605-655): padded FP32 copies of fs, ce and the temperature and
 * a copy of gr, converted on the device.  A NULL output is skipped.
 */
int  mmca_io_fields(mmca_ctx *ctx, float *fs_io, float *ce_io, float *temperature_io, int32_t *gr_io);
/*
 * The next n values of rand() from this block's random stream, for callers
 * that share the stream with the CA, as the XDMF writer's grain renumbering
 * does (getRenumberArray, io_xdmf.c:345-376: rand() % num_grains on rank 0).
 * The library draws grain orientations ahead of time; after this call the
 * orientations of grains created later come from the stream after these
 * values, as in the reference.
 */
int  mmca_rand(mmca_ctx *ctx, int32_t *out, int32_t n);
int  mmca_padded_cells(mmca_ctx *ctx, int64_t *cells);

/* --- host set-up without a device (for checking the set-up on a CPU) ----- */
/* Sites that mmca_init_nucleation would create: padded index, threshold and
 * matching FixedNuc/NucVolume entry (-1: random orientation), raster order. */
int  mmca_nucleation_sites(const mmca_params *params, const mmca_nucleation_cfg *cfg,
                           int32_t *idx_out, float *thr_out, int32_t *fixed_out, int64_t max,
                           int64_t *n_sites);
/* The first n random orientations drawn after that set-up (grain.c:118-129,
 * 3 draws each): Euler angles [rad] and row-major rotation matrices. */
int  mmca_orientation_draws(const mmca_params *params, const mmca_nucleation_cfg *cfg, int32_t n,
                            double *angles_out, double *rotmat_out);

/* --- tuning (no effect on results) --------------------------------------- */
#define MMCA_TUNE_ZM_PLANES  0   /* 1..1024: z-marching solute kernel, planes per workgroup; 0: cell-parallel
                                    (default 64 for FP64 blocks, 8 for FP32-deviation blocks and batches) */
#define MMCA_TUNE_ZM_ROWS    1   /* rows per z-marching workgroup: 4, 8 or 16 (default 16 for FP64 blocks,
                                    4 for FP32-deviation blocks and batches) */
#define MMCA_TUNE_XCD_REMAP  2   /* 1: workgroups of one XCD take contiguous tiles (default 1 for FP64 blocks,
                                    0 for FP32-deviation blocks and batches) */
#define MMCA_TUNE_GRAPH      3   /* batches: 1 = replay steps from a captured HIP graph */
#define MMCA_TUNE_ZM_LDS     4   /* 1 (default): z-marching kernel with its planes staged in LDS by direct loads; 0: in registers */
#define MMCA_TUNE_CAPTURE_CHUNK 5 /* >0: capture deals list slots in chunks of this many (rounded up to
                                      256), chunk q on XCD q % 8; 0: 256-slot groups over the whole grid */
#define MMCA_TUNE_TRACK_D    6   /* 1: keep d as the reference's array holds it (growth.c:392-403 writes
                                    solid cells only, so a cell that melts keeps its last value);
                                    MMCA_FIELD_D then gives that array */
#define MMCA_TUNE_BEAM_SEP   7   /* FP32-deviation Gaussian-beam blocks only (default 1 there): the
                                    beam temperature (temperature.c:104-112) as the sum over its 400
                                    nodes of the product of an x, a y and a z factor, computed per
                                    step for the whole block (within 1e-13 of the quadrature,
                                    not its bits); 0: the quadrature per cell.  FP64 storage
                                    always uses the quadrature (value 1 is refused there) */
#define MMCA_TUNE_CAPTURE_BAND 8 /* >0: capture works on units of this many rows of one plane, band q
                                     of every plane on XCD q % 8, units taken plane by plane; used
                                     when the block has at least 8 bands and the step builds the
                                     list itself (mmca_step, batches), before CAPTURE_CHUNK */
int  mmca_set_tuning(mmca_ctx *ctx, int key, int value);
int  mmca_get_tuning(mmca_ctx *ctx, int key, int *value);

/* Average duration [ms] of the dominant kernel over the last mmca_step call,
 * measured with HIP events on the context's stream (0 when not timed). */
int  mmca_set_kernel_timing(mmca_ctx *ctx, int enable);
int  mmca_kernel_time_ms(mmca_ctx *ctx, double *avg_ms, int64_t *launches);

/* --- batches: independent subblocks advanced together -------------------- */
/*
 * No reference counterpart: the reference advances one subblock per MPI
 * rank.  A batch holds `ndomains` subblocks of any sizes (same precision and
 * temperature-model class), each created as by mmca_create +
 * mmca_init_nucleation (nucleation may be NULL), and advances them with one
 * set of seven launches per step.  Each domain keeps its own timestep, grains,
 * RNG stream and errors; a domain that fails stops, the others go on.
 */
typedef struct mmca_batch mmca_batch;
int  mmca_batch_create(const mmca_params *params, const mmca_nucleation_cfg *nucleation, int ndomains,
                       int device, mmca_batch **out);
int  mmca_batch_destroy(mmca_batch *batch);
int  mmca_batch_size(mmca_batch *batch, int *ndomains);
/* the context of domain i (for downloads and single operators; do not destroy) */
int  mmca_batch_domain(mmca_batch *batch, int i, mmca_ctx **ctx);
/* n steps of every domain; reports[ndomains] may be NULL.  Returns the first
 * non-OK domain status (the others have stepped). */
int  mmca_batch_step(mmca_batch *batch, int nsteps, int want_sums, mmca_report *reports);
/*
 * Couples the domains of a fresh batch into one global domain split into
 * nsub[0] x nsub[1] x nsub[2] subblocks, as the reference runs one subblock
 * per MPI rank (ca_main.c:186-212, calculate.c:193-352).  Domain d is rank
 * d of the MPI Cart grid (x fastest); its params must give sb_coords =
 * (d % nsub[0], d / nsub[0] % nsub[1], d / (nsub[0] nsub[1])), the same
 * NCellsPerSB, gdim = nsub * NCellsPerSB, and its nucleation set-up must
 * name rank d (each rank samples its subblock from its own stream).  Each
 * step then exchanges the face halos of cl, fs and dc before diffusion and
 * of gr and fs after partitioning (ExchangeFacesForVar), and activates new
 * grains across the ranks in rank order into one shared grain table
 * (activateNewGrains, grain.c:841-916), while every rank keeps its own
 * interface list and dc_tmp slots.  A wrapped face (FaceCtrl 0) of the
 * global domain exchanges with the subblock at the other end of the axis
 * (sbID_from_sbcoords, functions.c:272-311); with one subblock along that
 * axis it keeps its halo, as one block does.  All domains must have the same
 * FaceCtrl.  Step the batch only through mmca_batch_step afterwards.
 */
int  mmca_batch_couple(mmca_batch *batch, const int32_t nsub[3]);
/*
 * Restart image of a whole batch, as mmca_save_state / mmca_load_state are
 * for one block (checkpoint.c:955-1238 replaced; the reference's checkpoint
 * stores neither the random streams nor dc_tmp).  For a coupled batch the
 * image holds every rank's fields, list, dc_tmp slots and stream, and the
 * shared grain table once.  Load into a batch created with the same
 * parameters and nucleation set-up and, for a coupled image, coupled with
 * the same nsub; the run then continues bit for bit.  A coupled batch that
 * stopped on an error (every rank stops at the step of the first one, as the
 * reference's error() ends every rank) is refused with MMCA_ERR_STATE.
 */
int  mmca_batch_save_state(mmca_batch *batch, void *buf, int64_t capacity, int64_t *size);
int  mmca_batch_load_state(mmca_batch *batch, const void *buf, int64_t size);
/* mmca_save_state_ex for a batch (every domain with the same flags). */
int  mmca_batch_save_state_ex(mmca_batch *batch, int flags, void *buf, int64_t capacity, int64_t *size);
int  mmca_batch_set_tuning(mmca_batch *batch, int key, int value);
int  mmca_batch_set_kernel_timing(mmca_batch *batch, int enable);
int  mmca_batch_kernel_time_ms(mmca_batch *batch, double *avg_ms, int64_t *launches);
/* Autotuner of the solute-kernel shape for a batch: builds a scratch batch
 * from the same inputs, runs `warmup` steps, then times `trial_steps` steps
 * of each candidate shape and returns the fastest (results do not depend on
 * the shape). */
int  mmca_batch_autotune(const mmca_params *params, const mmca_nucleation_cfg *nucleation, int ndomains,
                         int device, int warmup, int trial_steps, int *best_planes, int *best_rows,
                         double *best_ms);

/* --- adjoint of diffusion + partitioning on a recorded trajectory --------- */
/*
 * No reference counterpart: the reference has no derivatives.  A forward
 * run records, per step and cell, which branch of fs_change_diffuse
 * (growth.c:221-273) the cell took.  With those branches (and so every
 * nucleation and capture decision) frozen, one step maps the FP64 state
 * (ce, cl, fs) smoothly to the next one: diffusion (growth.c:85-113), then
 * the recorded branch.  A capture (growth.c:789-791) sets cl = ce and
 * fs = 0, which the LIQUID and REMELT branches of the captured cell have
 * already set, so it adds nothing to this map.
 */
#define MMCA_BRANCH_LIQUID  0   /* gr <= 0, or T < 1 or T > T_liq: cl = ce, fs = 0 (:228-234) */
#define MMCA_BRANCH_SOLID   1   /* fs == 1 and undercooled: cl = ce * (1/k)        (:236-240) */
#define MMCA_BRANCH_GROW    2   /* cl += dcl, fs += dfs                            (:254-263) */
#define MMCA_BRANCH_FULL    3   /* as GROW with fs clamped to 1                    (:264)     */
#define MMCA_BRANCH_REMELT  4   /* fs + dfs < 0: fs = 0, gr = 0, cl = ce           (:266-272) */

/*
 * Branch record of one step, 3 bits per cell: for each padded row (j, k),
 * j = 0..ny+1, k = 0..nz+1, and each 64-cell segment s = 0..ceil(nx/64)-1 of
 * its interior x, three uint64 words at word index
 * ((k * (ny+2) + j) * ceil(nx/64) + s) * 3 + p; bit (x-1) % 64 of word p is
 * bit p of the MMCA_BRANCH_* code of interior cell x.  Words of halo rows are
 * not written.  mmca_record_bytes gives the bytes of one step.
 */
int  mmca_record_bytes(const mmca_params *params, int64_t *bytes_per_step);

/*
 * Record the branch of every cell for the next n_steps steps of mmca_step or
 * mmca_doiteration into `codes` (device memory of n_steps * mmca_record_bytes
 * bytes, step s of the recording at codes + s * mmca_record_bytes).
 * Recording uses the cell-parallel solute kernel and FP64 storage
 * (MMCA_ERR_ARG otherwise); mmca_step refuses to run past the recorded
 * range.  codes == NULL or n_steps == 0 stops recording.
 */
int  mmca_set_record(mmca_ctx *ctx, uint8_t *codes, int64_t n_steps);

/*
 * One frozen-branch step of the FP64 state.  All pointers are device
 * memory of padded-cell size; `timestep` is the number of the recorded step
 * (its temperature time is timestep * dt), `codes` its branch record.  The
 * constants come from `params` (d_liq, d_sol and part_coef may differ from
 * the recorded run).  Only interior cells of the outputs are written, and
 * the outputs may not be the input buffers.  `codes` is the step's branch
 * record (mmca_record_bytes bytes).  `stream` is a hipStream_t (NULL: the
 * null stream).
 */
int  mmca_replay_step(const mmca_params *params, uint64_t timestep, const uint8_t *codes,
                      const double *ce, const double *cl, const double *fs,
                      double *ce_out, double *cl_out, double *fs_out, void *stream);

/* Device workspace, in bytes, that mmca_adjoint_step needs for `params`. */
int  mmca_adjoint_workspace_bytes(const mmca_params *params, int64_t *bytes);

/*
 * Adjoint (reverse-mode derivative) of mmca_replay_step.  (ce, cl, fs) is
 * the state at the start of the step; (g_ce, g_cl, g_fs) the derivative of
 * a scalar loss with respect to the state after it.  Writes the interior of
 * (g_ce_in, g_cl_in, g_fs_in), the derivative with respect to the state at
 * the start, and adds the derivative with respect to (d_liq, d_sol,
 * part_coef) to g_param[0..2] (device memory).  Each g_*_in may be the
 * same buffer as the matching g_* (in-place update).  The workspace is
 * device memory of
 * mmca_adjoint_workspace_bytes whose padded-cell part the library keeps at
 * zero on the halo; zero it once before the first call.
 */
int  mmca_adjoint_step(const mmca_params *params, uint64_t timestep, const uint8_t *codes,
                       const double *ce, const double *cl, const double *fs,
                       const double *g_ce, const double *g_cl, const double *g_fs,
                       double *g_ce_in, double *g_cl_in, double *g_fs_in, double *g_param,
                       void *workspace, void *stream);

/*
 * mmca_adjoint_step without a workspace argument and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a fixed 32 MiB device array of the library
 * (each GPU holds its own copy until the process ends), which covers blocks
 * of up to about 4.1 M padded cells; larger blocks are refused with
 * MMCA_ERR_ARG.  Every call on one GPU uses that GPU's array, so calls that
 * overlap on one GPU (other streams, other threads) race.  Results are those
 * of mmca_adjoint_step.
 */
int  mmca_adjoint_step_noalloc(const mmca_params *params, uint64_t timestep, const uint8_t *codes,
                               const double *ce, const double *cl, const double *fs,
                               const double *g_ce, const double *g_cl, const double *g_fs,
                               double *g_ce_in, double *g_cl_in, double *g_fs_in, double *g_param,
                               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MMCA_H */
