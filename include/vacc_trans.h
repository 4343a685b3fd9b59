/*
 * vacc_trans.h -- spatial translation for libvacc: the semi-Lagrangian PPM transport of
 * every velocity cell of every spatial cell along one spatial dimension, over "pencils"
 * (1-D lines of spatial cells), TRANS_SEMILAG_PPM.  Same return convention as vacc.h.
 *
 * The cells of a context are the spatial cells; pencils name them by their index in the
 * context.  A pencil of L positions holds VACC_TRANS_STENCIL stencil cells at each end
 * and the cells it propagates in between; per position it carries the cell's width along
 * the dimension (dz) and the target ratio (0: the position is not written).  A cell that
 * is written at any position starts from zero and receives the contributions of every
 * source position in pencil order, then position order; other cells keep their values.
 * Blocks are neither added nor removed: mass bound for a block its target cell lacks is
 * dropped, as in the reference (the block adjust with spatial neighbours provides them).
 *
 * Reference interfaces replaced (file:line in fmihpc/vlasiator):
 *   vacc_trans_map_1d         <- trans_map_1d_amr(mpiGrid, localPropagatedCells,
 *                                remoteTargetCells, nPencils, dimension, dt, popID)
 *                                vlasovsolver/cpu_trans_map_amr.cpp:214 (propagatePencil :36,
 *                                copy_trans_block_data_amr :162; GPU twin
 *                                gpu_trans_map_amr.cpp), PPM on non-uniform widths
 *                                cpu_1d_ppm_nonuniform.hpp:40
 *   vacc_trans_set_pencils    <- the DimensionPencils[dimension] set (setOfPencils,
 *                                cpu_trans_pencils.hpp:34-51: lengthOfPencils, idsStart, ids,
 *                                sourceDZ, targetRatios), pencils in bin order
 *   vacc_build_pencils_uniform <- prepareSeedIdsAndPencils (cpu_trans_pencils.cpp:1357) for a
 *                                grid without refinement: getSeedIds :920,
 *                                buildPencilsWithNeighbors :719,
 *                                computeSpatialSourceCellsForPencil :447
 */
#ifndef VACC_TRANS_H
#define VACC_TRANS_H

#include <stdint.h>

#include "vacc.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VACC_TRANS_STENCIL 2   /* VLASOV_STENCIL_WIDTH of TRANS_SEMILAG_PPM (definitions.h:78-84) */

/* Spatial cell kinds for the pencil builder (sysboundarytype / sysBoundaryLayer):
 * NORMAL = NOT_SYSBOUNDARY (propagated and written); BOUNDARY = a system-boundary cell of
 * layer 1 (propagated as a source, not written); INACTIVE = DO_NOT_COMPUTE or a deeper
 * boundary layer (not propagated; a stencil position on it takes the nearest good cell). */
enum { VACC_CELL_NORMAL = 0, VACC_CELL_BOUNDARY = 1, VACC_CELL_INACTIVE = 2 };

typedef struct vacc_pencils {
   uint32_t npencils;
   const uint32_t* length;   /* [npencils] positions of each pencil, stencil cells included */
   const uint64_t* start;    /* [npencils] first position of each pencil in cell/dz/ratio */
   const uint32_t* cell;     /* context cell index per position (VACC_INVALID_ID: none) */
   const float* dz;          /* cell width along the dimension per position */
   const float* ratio;       /* target ratio per position (0: not written) */
} vacc_pencils;

/* Pencils along dim of a grid of dims[0] x dims[1] x dims[2] cells without refinement,
 * cell index x + y*dims[0] + z*dims[0]*dims[1], kinds cell_kind[] (VACC_CELL_*), periodic
 * flags per axis and cell widths dx[3].  First call with length == NULL: *npencils and
 * *npositions are set.  Second call with arrays of those sizes: they are filled.  A
 * propagated cell starts a pencil where its -dim neighbour is across a periodic boundary,
 * inactive, or absent (a non-periodic edge). */
int vacc_build_pencils_uniform(const uint32_t dims[3], const int32_t periodic[3], const uint8_t* cell_kind,
                               const double dx[3], uint32_t dim, uint32_t* npencils, uint64_t* npositions,
                               uint32_t* length, uint64_t* start, uint32_t* cell, float* dz, float* ratio);

/* Installs the pencils of dimension dim for the context's cells (kept until replaced or the
 * cell set changes).  Errors if a pencil is shorter than 2*VACC_TRANS_STENCIL+1, a cell
 * index is outside the context, or a written cell whose contributions come from several
 * positions of a pencil, or from several pencils in a 16-bit storage form, has stencils that
 * span more than 12 distinct cells. */
int vacc_trans_set_pencils(vacc_ctx* ctx, uint32_t dim, const vacc_pencils* pencils);

/* Translates every cell of the context along dim by dt (converted to float, the reference's
 * Realf dt), over the installed pencils.  Threshold of the reconstruction: the species'
 * sparse_min_value.  With a 16-bit storage form the arithmetic is FP32 on the decoded
 * blocks and every written block is encoded again; blocks that are not written keep their
 * codes. */
int vacc_trans_map_1d(vacc_ctx* ctx, uint32_t dim, double dt);

#ifdef __cplusplus
}
#endif

#endif
