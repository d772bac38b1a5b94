/*
 * Checkpoint files of the C host driver in the layout of the reference's
 * writeTaskCheckpoint (checkpoint.c:955-1238): for every rank one HDF5 file,
 * rank 0's named cp_main.h5 and rank r's cp_task_<r %03d>.h5 after the
 * renaming of clean_checkpoint (checkpoint.c:165-177), each with
 *   /curTimeStep, /g_ts_idx, /numGrain, /grainCache (all grains slots),
 *   /SUBBLOCK/SubBlock_XxYxZ/{CoordTemp, temperature, gr, fs, ce, oce, cl,
 *   diff_id, mold, d, nuc_threshold, dc, nuc_pt}
 * over the padded subblock, chunked whole, shuffled and deflated at level 9.
 *
 * The reference's checkpoint holds neither its random streams nor the
 * dc_tmp slots, and its restart does not continue a run (its main-file
 * writer fails first).  Rank 0's file here also has a group /MMCA with what
 * a bit-for-bit continuation needs besides the fields: the run's restart
 * image without its fields (MMCA_STATE_NO_FIELDS: counters, interface lists,
 * dc_tmp slots, grain table, streams) and the XDMF writer's renumbering.
 * Restoring loads that image and then every rank's ce, cl, fs, gr, dc and d
 * from its task file's datasets.
 */
#ifndef MMCA_CKPT_H
#define MMCA_CKPT_H

#include <stdint.h>

#include "mmca.h"
#include "mmca_ctrl.h"

/* Steps before a checkpoint whose fields the writer uses: the halo planes
 * of the reference's arrays hold what its face exchanges of those steps
 * left there (an edge cell comes through a neighbour's halo, a corner
 * through two). */
#define MMCA_CKPT_HISTORY 3

/* One rank's padded fields after a step (as mmca_download_field gives them;
 * NULL where not kept). */
typedef struct mmca_ckpt_snap {
    double *ce, *cl, *fs, *dc, *d;
    int32_t *gr;
} mmca_ckpt_snap;

/*
 * Writes the checkpoint of timestep ts.  bt: the coupled batch of the ranks,
 * or NULL for one block (doms[0]).  doms[r], ps[r]: rank r's subblock and
 * parameters, rank r at Cart coordinates x = r % nsub[0],
 * y = r / nsub[0] % nsub[1], z = the rest.  hist[k][r]: rank r's fields after
 * step ts - MMCA_CKPT_HISTORY + k, or hist[k] NULL where that step was not
 * kept: ce after step ts - 1 is the file's oce (ce before the last step's
 * diffusion; ce itself without it), and with several ranks cl, fs, dc, gr
 * and d of those steps give the halo planes.  The blocks must keep the
 * reference's d array (MMCA_TUNE_TRACK_D).  renumber[0..nren): the XDMF
 * renumbering to store with it.  Returns 0, or -1 with a message on stderr.
 */
int mmca_ckpt_write(const mmca_run_cfg *cfg, const int nsub[3], mmca_batch *bt, mmca_ctx *const *doms,
                    const mmca_params *ps, const mmca_ckpt_snap *const *hist, uint64_t ts,
                    const uint64_t *renumber, int64_t nren);

/*
 * Restores the run written by mmca_ckpt_write into bt / doms, created with the
 * same parameters and nucleation set-up (and coupled with the same nsub).
 * renumber receives at most max entries of the stored renumbering and *nren
 * their number.  Returns 0, or -1 with a message on stderr.
 */
int mmca_ckpt_read(const mmca_run_cfg *cfg, const int nsub[3], mmca_batch *bt, mmca_ctx *const *doms,
                   const mmca_params *ps, uint64_t *renumber, int64_t max, int64_t *nren);

#endif
