/*
 * XDMF + HDF5 output of the C host driver, in the layout of the
 * reference's writer (io_xdmf.c): per rank and output step
 * <BaseFileName>_r<rank>/<step %05lu>.h5 with /SubBlock_XxYxZ/{Temperature,
 * CE, Grain, FracSolid}, the global file <BaseFileName>_gbl.h5 with the grain
 * renumbering and angles, and the time series <BaseFileName>_ts.xmf.
 */
#ifndef MMCA_XDMF_H
#define MMCA_XDMF_H

#include <stdint.h>

#include "mmca.h"
#include "mmca_ctrl.h"

typedef struct mmca_xdmf mmca_xdmf;

/* xdmf_prepareIO (io_xdmf.c:533-541): one output directory per rank, for
 * nsub[0]*nsub[1]*nsub[2] ranks of sbdim cells each (rank r at Cart
 * coordinates x = r % nsub[0], y = r / nsub[0] % nsub[1], z = the rest).
 * Returns NULL with a message on stderr on failure. */
mmca_xdmf *mmca_xdmf_open(const mmca_run_cfg *cfg, const int nsub[3], const int sbdim[3]);

/* writeData (loop.c:116-127) at timestep ts.  Before it returns: the grain
 * renumbering's draws from rank 0's random stream (mmca_rand), the grain
 * angles and the ranks' fields (mmca_io_fields).  A host thread then writes
 * rank 0's time-series entry and global file and every rank's HDF5 file
 * while the caller steps on; the next write, mmca_xdmf_flush and
 * mmca_xdmf_close wait for it.  doms[r] / ps[r]: the subblock of rank r (ps
 * must stay valid until then).  Returns 0, or -1 with a message on stderr
 * (also for a failure of the previous write's files). */
int mmca_xdmf_write(mmca_xdmf *x, mmca_ctx *const *doms, const mmca_params *ps, uint64_t ts);

/* Waits for the pending write; 0 if its files are complete. */
int mmca_xdmf_flush(mmca_xdmf *x);

void mmca_xdmf_close(mmca_xdmf *x);

/* The grain renumbering kept between writes (getRenumberArray's static
 * array), for restart files: get copies at most max entries into out and
 * returns the array's length; set replaces it. */
int64_t mmca_xdmf_get_renumber(const mmca_xdmf *x, uint64_t *out, int64_t max);
int mmca_xdmf_set_renumber(mmca_xdmf *x, const uint64_t *in, int64_t n);

#endif
