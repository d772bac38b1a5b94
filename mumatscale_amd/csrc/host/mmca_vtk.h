/*
 * VTK output of the C host driver (DataFormat VTK), in the layout of the
 * reference's writer (io_vtk.c, vtkCompress.c): per rank and output step
 * <BaseFileName>_<rank %04d>/<step>.vtu, an unstructured grid of the
 * subblock's cells with the FP64 temperature and the grain ids, and from rank
 * 0 the parallel index <BaseFileName>_<step>.pvtu.  Arrays are stored as
 * zlib blocks (one per x slab) behind a block-size header, base64 encoded.
 */
#ifndef MMCA_VTK_H
#define MMCA_VTK_H

#include <stdint.h>

#include "mmca.h"
#include "mmca_ctrl.h"

typedef struct mmca_vtk mmca_vtk;

/* vtk_prepareIOSubblock (io_vtk.c:463-474): one directory per rank, for
 * nsub[0]*nsub[1]*nsub[2] ranks of sbdim cells each (rank r at Cart
 * coordinates x = r % nsub[0], y = r / nsub[0] % nsub[1], z = the rest).
 * Returns NULL with a message on stderr on failure. */
mmca_vtk *mmca_vtk_open(const mmca_run_cfg *cfg, const int nsub[3], const int sbdim[3]);

/* writeData (loop.c:116-127) for DataFormat VTK at timestep ts: rank 0's
 * .pvtu, then every rank's .vtu from its temperature and grain fields.
 * doms[r] / ps[r]: the subblock of rank r.  Returns 0, or -1 with a message
 * on stderr. */
int mmca_vtk_write(mmca_vtk *v, mmca_ctx *const *doms, const mmca_params *ps, uint64_t ts);

void mmca_vtk_close(mmca_vtk *v);

#endif
