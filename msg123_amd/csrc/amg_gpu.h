// amg_gpu.h -- device smoothed-aggregation AMG used as an FP32 preconditioner
// inside the FP64 BiCGStab (MSG123_PRECON_AMG32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hip.h"
#include "msg123_gpu.h"

namespace msg123 {

struct AmgGpu;

// Hierarchy set-up for the 7-diagonal FP64 matrix A7 ([7][N] on the device,
// CSR neighbour order) of a box whose reference CSR pattern is (offind,
// offrow 1-based): the host aggregates level by level (the reference's
// strength filter and heap aggregation), the device computes P, A P and
// R A P.  The hierarchy keeps its aggregation; amg_gpu_numeric recomputes
// every value from a new matrix with the same pattern.
int amg_gpu_create(const Dims &dm, const int32_t *offind, const int32_t *offrow,
                   const double *A7, const msg123_amg_opts &opt, hipStream_t s,
                   AmgGpu **out);
void amg_gpu_destroy(AmgGpu *g);
int amg_gpu_numeric(AmgGpu *g, const double *A7, hipStream_t s);
// z = V-cycle(r), FP64 in and out, FP32 inside
int amg_gpu_apply(AmgGpu *g, const double *r, double *z, hipStream_t s);
int amg_gpu_nlevel(const AmgGpu *g);
// level l (0 = the box): unknowns, off-diagonal entries, entries of P from l-1
void amg_gpu_level_size(const AmgGpu *g, int l, int32_t *n, int32_t *nnz, int32_t *np);
// copy level l's matrix (l >= 1; reference layout: row ends, 1-based columns)
// and the prolongator from level l-1 to the host; any pointer may be NULL
int amg_gpu_level_get(AmgGpu *g, int l, int32_t *offind, int32_t *offrow, double *dmat,
                      double *lumat, int32_t *pindex, int32_t *poffrow, double *pval,
                      hipStream_t s);

}  // namespace msg123
