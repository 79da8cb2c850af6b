// amg_gpu.hip -- smoothed-aggregation AMG on the device (MSG123_PRECON_AMG32).
//
// Set-up (once per hierarchy): the host aggregates each level with the
// reference's strength filter and heap aggregation (amg_aggregate, the
// algorithm of Make_amg_matrix.f90:67-222) and derives every sparsity
// pattern: the Jacobi-smoothed prolongator P (:224-300), R = P^T (:373-425),
// A P, and the Galerkin operator R A P in the reference's entry order
// (:427-631).  The aggregation is kept for the life of the hierarchy.
//
// Numeric phase (every Newton iteration, device only, FP64):
//   P    = (I - w D~^-1 A_strong) T   per fine row, in the reference's order
//          (D~ = diagonal plus the weak entries of the row), so P is
//          bit-identical to the reference's for the same matrix;
//   A P  per fine row (fixed pattern);
//   R A P per coarse row (fixed pattern), diagonal kept apart.
// Then the FP32 copies the V-cycle reads: the box matrix as seven diagonals,
// coarse matrices in sliced ELL with 64 rows per slice (one wavefront per
// slice, coalesced), and w_s / diag for damped Jacobi.
//
// V-cycle (FP32 inside, FP64 in and out): npre damped-Jacobi sweeps from
// zero (the first two fused), residual, restriction R r, the next level,
// prolongation x += P x_c, npost sweeps; the coarsest level is solved with
// a dense inverse (Gauss-Jordan with partial pivoting in LDS, FP64) when it
// has at most kDenseMax unknowns, else with Jacobi sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "amg.h"
#include "amg_gpu.h"

#define AMGCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "msg123: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                               \
      return MSG123_EDEVICE;                                                          \
    }                                                                                 \
  } while (0)

namespace msg123 {
namespace {

constexpr int kDenseMax = 80;  // largest coarsest level solved with a dense inverse
constexpr int kAB = 256;       // threads per block of the AMG kernels
constexpr int16_t kWeak = -1, kSkip = -2, kNone = -3;

inline int grid(int n) { return (n + kAB - 1) / kAB; }

// Sliced ELL, 64 rows per slice: entry j of row r sits at
// soff[r/64] + j*64 + r%64; padding entries point at the row itself with 0.
struct Sell {
  const int *soff, *scol;
  const float *sval;
};

__device__ __forceinline__ float sell_dot(const Sell &A, int row, const float *__restrict__ x) {
  const int s = row >> 6, lane = row & 63;
  const int b = A.soff[s], w = (A.soff[s + 1] - b) >> 6;
  float t = 0.0f;
  for (int j = 0; j < w; j++) {
    const int k = b + j * 64 + lane;
    t += A.sval[k] * x[A.scol[k]];
  }
  return t;
}

struct Mat7f {
  const float *o[6];  // up, N, W, E, S, down
  const float *d;
};

__device__ __forceinline__ float off7(const Mat7f &A, const Dims &dm, const float *__restrict__ x,
                                      int i) {
  int j[6];
  bool ok[6];
  neighbours(dm, i, j, ok);
  float t = 0.0f;
#pragma unroll
  for (int d = 0; d < 6; d++)
    if (ok[d]) t += A.o[d][i] * x[j[d]];
  return t;
}

// ---------------------------------------------------------------------------
// Numeric set-up kernels (FP64)

// M-matrix projection of the box matrix: couplings max(a_ij, 0), diagonal
// -max(|a_ii|, sum of the clipped couplings)
__global__ void __launch_bounds__(kAB)
k_amg_project(int n, const double *__restrict__ A7, double *__restrict__ P7) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  const size_t N = (size_t)n;
  double s = 0.0;
#pragma unroll
  for (int d = 1; d < 7; d++) {
    const double v = fmax(A7[d * N + i], 0.0);
    P7[d * N + i] = v;
    s += v;
  }
  P7[i] = -fmax(fabs(A7[i]), s);
}

// FP32 copy of the box matrix and w_s / d
__global__ void __launch_bounds__(kAB)
k_amg_a0(int n, const double *__restrict__ A7, float *__restrict__ A32, float *__restrict__ dinv,
         double ws) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  const size_t N = (size_t)n;
#pragma unroll
  for (int d = 0; d < 7; d++) A32[d * N + i] = (float)A7[d * N + i];
  dinv[i] = (float)(ws / A7[i]);
}

// P of the box level: make_prolomat's arithmetic (Make_amg_matrix.f90:
// 300-371) on the seven diagonals: per strong entry with an aggregated
// column, P(i, slot) -= a_ij in CSR order; d~ = d + weak entries in CSR
// order; P *= w / d~; P(i, own) += 1 - w.
__global__ void __launch_bounds__(kAB)
k_amg_p0(Dims dm, const double *__restrict__ A7, const int16_t *__restrict__ smap,
         const int16_t *__restrict__ pown, const int *__restrict__ pptr, double *__restrict__ pval,
         float *__restrict__ pval32, double omega) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= dm.n) return;
  const size_t N = (size_t)dm.n;
  const int p0 = pptr[i], cnt = pptr[i + 1] - p0;
  if (cnt == 0) return;
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double tdm = A7[i];
#pragma unroll
  for (int d = 0; d < 6; d++) {
    const int code = smap[d * N + i];
    if (code == kNone || code == kSkip) continue;
    const double lu = A7[(d + 1) * N + i];
    if (code == kWeak) {
      tdm = tdm + lu;
    } else {
#pragma unroll
      for (int q = 0; q < 7; q++)
        if (q == code) acc[q] = acc[q] - lu;
    }
  }
  const int own = pown[i];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    if (q >= cnt) break;
    double v = acc[q] * omega / tdm;
    if (q == own) v = v + 1.0 - omega;
    pval[p0 + q] = v;
    pval32[p0 + q] = (float)v;
  }
}

// P of a coarse level (CSR matrix, diagonal apart), same arithmetic
__global__ void __launch_bounds__(kAB)
k_amg_pc(int n, const int *__restrict__ aptr, const double *__restrict__ aval,
         const double *__restrict__ adiag, const int16_t *__restrict__ smap,
         const int16_t *__restrict__ pown, const int *__restrict__ pptr, double *__restrict__ pval,
         float *__restrict__ pval32, double omega) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  const int p0 = pptr[i], p1 = pptr[i + 1];
  if (p0 == p1) return;
  for (int q = p0; q < p1; q++) pval[q] = 0.0;
  double tdm = adiag[i];
  for (int k = aptr[i]; k < aptr[i + 1]; k++) {
    const int code = smap[k];
    if (code == kWeak) tdm = tdm + aval[k];
    else if (code >= 0) pval[p0 + code] = pval[p0 + code] - aval[k];
  }
  const int own = pown[i];
  for (int q = p0; q < p1; q++) {
    double v = pval[q] * omega / tdm;
    if (q - p0 == own) v = v + 1.0 - omega;
    pval[q] = v;
    pval32[q] = (float)v;
  }
}

// out[e] = sum over e's pairs k of va[pa[k]] * vb[pb[k]], in list order.  The
// lists, built once per hierarchy on the host, name the operands of every
// entry of A P (va = the matrix, vb = P) and of R A P (va = P, vb = A P), so
// the numeric phase is pure gathers with a fixed summation order.
__global__ void __launch_bounds__(kAB)
k_gather(int ne, const int *__restrict__ ptr, const int *__restrict__ pa,
         const int *__restrict__ pb, const double *__restrict__ va, const double *__restrict__ vb,
         double *__restrict__ out) {
  const int e = blockIdx.x * kAB + threadIdx.x;
  if (e >= ne) return;
  double t = 0.0;
  for (int k = ptr[e]; k < ptr[e + 1]; k++) t += va[pa[k]] * vb[pb[k]];
  out[e] = t;
}

// After the Galerkin gather: 1 on diagonals the pattern lacks (as the
// reference initialises them), the FP32 sliced-ELL values, diag and w/diag.
__global__ void __launch_bounds__(kAB)
k_coarse_finish(int nc, const int *__restrict__ cptr, const char *__restrict__ hasdiag,
                double *__restrict__ flat, const int *__restrict__ soff, float *__restrict__ sval,
                float *__restrict__ cdiag32, float *__restrict__ cdinv, double ws) {
  const int I = blockIdx.x * kAB + threadIdx.x;
  if (I >= nc) return;
  double dg = flat[I];
  if (!hasdiag[I]) {
    dg = 1.0;
    flat[I] = dg;
  }
  cdiag32[I] = (float)dg;
  cdinv[I] = (float)(ws / dg);
  const double *cval = flat + nc;
  const int sb = soff[I >> 6], lane = I & 63;
  for (int q = cptr[I]; q < cptr[I + 1]; q++)
    sval[sb + (q - cptr[I]) * 64 + lane] = (float)cval[q];
}

// Dense inverse of the coarsest level (n <= kDenseMax) by Gauss-Jordan with
// partial pivoting in LDS, FP64; stored FP32 row-major.
__global__ void __launch_bounds__(kAB)
k_amg_dense_inv(int n, const int *__restrict__ cptr, const int *__restrict__ ccol,
                const double *__restrict__ cval, const double *__restrict__ cdiag,
                float *__restrict__ inv) {
  __shared__ double M[kDenseMax][2 * kDenseMax];
  __shared__ double fac[kDenseMax];
  __shared__ int piv;
  const int w = 2 * n;
  for (int q = threadIdx.x; q < n * w; q += kAB) {
    const int r = q / w, c = q % w;
    M[r][c] = (c >= n) ? ((c - n == r) ? 1.0 : 0.0) : 0.0;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += kAB) {
    M[r][r] = cdiag[r];
    for (int k = cptr[r]; k < cptr[r + 1]; k++) M[r][ccol[k]] = cval[k];
  }
  __syncthreads();
  for (int c = 0; c < n; c++) {
    if (threadIdx.x == 0) {
      int p = c;
      double best = fabs(M[c][c]);
      for (int r = c + 1; r < n; r++)
        if (fabs(M[r][c]) > best) {
          best = fabs(M[r][c]);
          p = r;
        }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != c)
      for (int q = threadIdx.x; q < w; q += kAB) {
        const double t = M[c][q];
        M[c][q] = M[p][q];
        M[p][q] = t;
      }
    __syncthreads();
    const double pinv = 1.0 / M[c][c];
    __syncthreads();
    for (int q = threadIdx.x; q < w; q += kAB) M[c][q] *= pinv;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += kAB) fac[r] = M[r][c];
    __syncthreads();
    for (int q = threadIdx.x; q < n * w; q += kAB) {
      const int r = q / w, cc = q % w;
      if (r != c) M[r][cc] -= fac[r] * M[c][cc];
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < n * n; q += kAB) inv[q] = (float)M[q / n][n + q % n];
}

// ---------------------------------------------------------------------------
// V-cycle kernels (FP32)

template <typename TB>
__device__ __forceinline__ float ld(const TB *p, int i) {
  return (float)p[i];
}

// two damped-Jacobi sweeps from zero on the box: x1 = D^-1 b (weighted),
// x = x1 + w D^-1 (b - A x1)
template <typename TB>
__global__ void __launch_bounds__(kAB)
k_v0_pre2(Dims dm, Mat7f A, const float *__restrict__ dinv, const TB *__restrict__ b,
          float *__restrict__ x) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= dm.n) return;
  int j[6];
  bool ok[6];
  neighbours(dm, i, j, ok);
  const float x1 = dinv[i] * ld(b, i);
  float t = A.d[i] * x1;
#pragma unroll
  for (int d = 0; d < 6; d++)
    if (ok[d]) t += A.o[d][i] * (dinv[j[d]] * ld(b, j[d]));
  x[i] = x1 + dinv[i] * (ld(b, i) - t);
}

// x = D^-1 b (weighted): one sweep from zero
template <typename TB>
__global__ void __launch_bounds__(kAB)
k_first(int n, const float *__restrict__ dinv, const TB *__restrict__ b, float *__restrict__ x) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i < n) x[i] = dinv[i] * ld(b, i);
}

// one sweep on the box: out = x + w D^-1 (b - A x); TO = float or double
template <typename TB, typename TO>
__global__ void __launch_bounds__(kAB)
k_v0_sweep(Dims dm, Mat7f A, const float *__restrict__ dinv, const TB *__restrict__ b,
           const float *__restrict__ x, TO *__restrict__ out) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= dm.n) return;
  const float t = A.d[i] * x[i] + off7(A, dm, x, i);
  out[i] = (TO)(x[i] + dinv[i] * (ld(b, i) - t));
}

template <typename TB>
__global__ void __launch_bounds__(kAB)
k_v0_resid(Dims dm, Mat7f A, const TB *__restrict__ b, const float *__restrict__ x,
           float *__restrict__ r) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= dm.n) return;
  r[i] = ld(b, i) - (A.d[i] * x[i] + off7(A, dm, x, i));
}

__global__ void __launch_bounds__(kAB)
k_vc_pre2(int n, Sell A, const float *__restrict__ diag, const float *__restrict__ dinv,
          const float *__restrict__ b, float *__restrict__ x1tmp, float *__restrict__ x) {
  // x1tmp holds dinv*b (written by k_first); x = x1 + dinv (b - A x1)
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  const float t = diag[i] * x1tmp[i] + sell_dot(A, i, x1tmp);
  x[i] = x1tmp[i] + dinv[i] * (b[i] - t);
}

__global__ void __launch_bounds__(kAB)
k_vc_sweep(int n, Sell A, const float *__restrict__ diag, const float *__restrict__ dinv,
           const float *__restrict__ b, const float *__restrict__ x, float *__restrict__ out) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  const float t = diag[i] * x[i] + sell_dot(A, i, x);
  out[i] = x[i] + dinv[i] * (b[i] - t);
}

__global__ void __launch_bounds__(kAB)
k_vc_resid(int n, Sell A, const float *__restrict__ diag, const float *__restrict__ b,
           const float *__restrict__ x, float *__restrict__ r) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  r[i] = b[i] - (diag[i] * x[i] + sell_dot(A, i, x));
}

// b_c(I) = sum_i P(i,I) r_i over R's row I
__global__ void __launch_bounds__(kAB)
k_restrict(int nc, const int *__restrict__ rptr, const int *__restrict__ rrow,
           const int *__restrict__ rpos, const float *__restrict__ pval32,
           const float *__restrict__ r, float *__restrict__ bc) {
  const int I = blockIdx.x * kAB + threadIdx.x;
  if (I >= nc) return;
  float t = 0.0f;
  for (int q = rptr[I]; q < rptr[I + 1]; q++) t += pval32[rpos[q]] * r[rrow[q]];
  bc[I] = t;
}

// x_i += sum_J P(i,J) xc_J
__global__ void __launch_bounds__(kAB)
k_prolong(int n, const int *__restrict__ pptr, const int *__restrict__ pcol,
          const float *__restrict__ pval32, const float *__restrict__ xc, float *__restrict__ x) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i >= n) return;
  float t = x[i];
  for (int q = pptr[i]; q < pptr[i + 1]; q++) t += pval32[q] * xc[pcol[q]];
  x[i] = t;
}

__global__ void __launch_bounds__(kAB)
k_dense_apply(int n, const float *__restrict__ inv, const float *__restrict__ b,
              float *__restrict__ x) {
  const int r = blockIdx.x * kAB + threadIdx.x;
  if (r >= n) return;
  float t = 0.0f;
  for (int c = 0; c < n; c++) t += inv[r * n + c] * b[c];
  x[r] = t;
}

template <typename T>
__global__ void k_to_double(int n, const float *__restrict__ x, T *__restrict__ z) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i < n) z[i] = (T)x[i];
}

__global__ void k_zero(int n, float *__restrict__ x) {
  const int i = blockIdx.x * kAB + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}

// ---------------------------------------------------------------------------

struct DevLevel {
  int n = 0, nnz = 0;  // unknowns, off-diagonal entries (levels >= 1)
  // matrix of this level (levels >= 1): CSR in the reference's entry order
  int *aptr = nullptr, *acol = nullptr;
  double *flat = nullptr;  // [n + nnz]: the diagonal, then the off-diagonal entries
  double *aval = nullptr, *adiag = nullptr;  // views into flat
  char *hasdiag = nullptr;
  int nsell = 0;
  int *soff = nullptr, *scol = nullptr;
  float *sval = nullptr, *diag32 = nullptr, *dinv = nullptr;
  // the box level: FP32 seven diagonals
  float *A32 = nullptr;
  // coarsening to the next level
  int np = 0, nap = 0;
  int *pptr = nullptr, *pcol = nullptr;
  double *pval = nullptr;
  float *pval32 = nullptr;
  int16_t *smap = nullptr, *pown = nullptr;
  int *rptr = nullptr, *rrow = nullptr, *rpos = nullptr;
  int *apptr = nullptr, *apcol = nullptr;
  double *apval = nullptr;
  // operand pairs of A P (per A P entry) and of R A P (per entry of the next
  // level's flat matrix)
  int *ap_ptr = nullptr, *ap_a = nullptr, *ap_b = nullptr;
  int *rp_ptr = nullptr, *rp_a = nullptr, *rp_b = nullptr;
  // host copies of the patterns kept for read-back
  std::vector<int> h_pptr, h_pcol;
  // V-cycle vectors
  float *b = nullptr, *x = nullptr, *y = nullptr, *r = nullptr;
  float *inv = nullptr;  // dense inverse of the coarsest level
};

}  // namespace

struct AmgGpu {
  Dims dm{};
  msg123_amg_opts opt{};
  std::vector<DevLevel> lv;
  std::vector<void *> allocs;
  bool dense = false;
  double *proj = nullptr;  // projected box matrix [7][N] when opt.project

  template <typename T>
  int alloc(T **p, size_t count) {
    if (hipMalloc((void **)p, sizeof(T) * (count ? count : 1)) != hipSuccess) return MSG123_ENOMEM;
    allocs.push_back((void *)*p);
    return MSG123_OK;
  }
  template <typename T>
  int upload(T **p, const std::vector<T> &v, hipStream_t s) {
    int rc = alloc(p, v.size());
    if (rc) return rc;
    if (!v.empty())
      AMGCHK(hipMemcpyAsync(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    return MSG123_OK;
  }
};

namespace {

// Host CSR of one level in the reference's layout (row ends, 1-based columns).
struct HostCsr {
  int n = 0;
  std::vector<int> offind, offrow;
  std::vector<double> dmat, lumat;
};

int direction(const Dims &dm, int row, int col) {
  const int o = col - row;
  if (o == -dm.nxy) return 0;
  if (o == -dm.nx) return 1;
  if (o == -1) return 2;
  if (o == 1) return 3;
  if (o == dm.nx) return 4;
  if (o == dm.nxy) return 5;
  return -1;
}

int numeric_level(AmgGpu *g, int l, const double *A7, hipStream_t s) {
  DevLevel &f = g->lv[l];
  DevLevel &c = g->lv[l + 1];
  const double omega = (double)g->opt.omega, ws = (double)g->opt.smooth_omega;
  const double *Aflat = (l == 0) ? A7 : f.flat;
  if (l == 0) {
    hipLaunchKernelGGL(k_amg_p0, dim3(grid(f.n)), dim3(kAB), 0, s, g->dm, A7, f.smap, f.pown,
                       f.pptr, f.pval, f.pval32, omega);
  } else {
    hipLaunchKernelGGL(k_amg_pc, dim3(grid(f.n)), dim3(kAB), 0, s, f.n, f.aptr, f.aval, f.adiag,
                       f.smap, f.pown, f.pptr, f.pval, f.pval32, omega);
  }
  hipLaunchKernelGGL(k_gather, dim3(grid(f.nap)), dim3(kAB), 0, s, f.nap, f.ap_ptr, f.ap_a,
                     f.ap_b, Aflat, f.pval, f.apval);
  const int ncf = c.n + c.nnz;
  hipLaunchKernelGGL(k_gather, dim3(grid(ncf)), dim3(kAB), 0, s, ncf, f.rp_ptr, f.rp_a, f.rp_b,
                     f.pval, f.apval, c.flat);
  hipLaunchKernelGGL(k_coarse_finish, dim3(grid(c.n)), dim3(kAB), 0, s, c.n, c.aptr, c.hasdiag,
                     c.flat, c.soff, c.sval, c.diag32, c.dinv, ws);
  AMGCHK(hipGetLastError());
  return MSG123_OK;
}

int finish_numeric(AmgGpu *g, hipStream_t s) {
  if (g->dense) {
    DevLevel &L = g->lv.back();
    hipLaunchKernelGGL(k_amg_dense_inv, dim3(1), dim3(kAB), 0, s, L.n, L.aptr, L.acol, L.aval,
                       L.adiag, L.inv);
    AMGCHK(hipGetLastError());
  }
  return MSG123_OK;
}

// Patterns of one coarsening step from host level f (index l) and the device
// arrays they need; returns 1 when the level cannot be coarsened.
int build_step(AmgGpu *g, int l, const HostCsr &f, hipStream_t s) {
  const Dims &dm = g->dm;
  const int nf = f.n;
  AmgAggregation ag;
  amg_aggregate(nf, f.offind.data(), f.offrow.data(), f.dmat.data(), f.lumat.data(),
                (double)g->opt.theta, ag);
  if (g->opt.complete_orphans) amg_complete_orphans(nf, f.offrow.data(), ag);
  const int nc = ag.nc;
  if (nc == 0 || nc >= nf) return 1;
  const int lev = l + 2;  // the reference's number of the coarse level
  int width = 30 + 80 * (lev - 2);
  if (width > nc) width = nc;
  std::vector<int> pptr, pcol1;
  if (!amg_prolong_pattern(nf, ag, f.offrow.data(), width, pptr, pcol1)) return 1;
  if (l == 0)
    for (int i = 0; i < nf; i++)
      if (pptr[i + 1] - pptr[i] > 7) return 1;  // the box kernel keeps at most 7 per row
  std::vector<int> pcol(pcol1.size());
  for (size_t q = 0; q < pcol1.size(); q++) pcol[q] = pcol1[q] - 1;
  // slot codes per matrix entry and the row's own slot
  const int nnz = f.offind[nf];
  std::vector<char> strong(nnz, 0);
  for (int i = 1; i <= nf; i++)
    for (int q = ag.sidx[i - 1]; q < ag.sidx[i]; q++) strong[ag.slu[q] - 1] = 1;
  auto slot_of = [&](int i0, int agg0) {  // i0, agg0 0-based
    for (int q = pptr[i0]; q < pptr[i0 + 1]; q++)
      if (pcol[q] == agg0) return q - pptr[i0];
    return -1;
  };
  std::vector<int16_t> smap(l == 0 ? (size_t)6 * nf : (size_t)nnz, kNone), pown(nf, -1);
  for (int i = 0; i < nf; i++) {
    for (int k = f.offind[i]; k < f.offind[i + 1]; k++) {
      const int col = f.offrow[k] - 1;
      int16_t code = kWeak;
      if (strong[k]) {
        const int a = ag.aggr[col + 1];
        code = a > 0 ? (int16_t)slot_of(i, a - 1) : kSkip;
      }
      if (l == 0) {
        const int d = direction(dm, i, col);
        if (d < 0) return MSG123_EINVAL;
        smap[(size_t)d * nf + i] = code;
      } else {
        smap[k] = code;
      }
    }
    if (ag.aggr[i + 1] > 0) pown[i] = (int16_t)slot_of(i, ag.aggr[i + 1] - 1);
  }
  // R = P^T
  std::vector<int> rindex, roffrow, rpos;
  amg_transpose(nf, nc, pptr, pcol1, rindex, roffrow, rpos);
  std::vector<int> rrow(roffrow.size());
  for (size_t q = 0; q < roffrow.size(); q++) rrow[q] = roffrow[q] - 1;
  // A P pattern: row i = union of P rows of i and its columns, first appearance order
  std::vector<int> apptr(nf + 1, 0), apcol;
  apcol.reserve((size_t)pptr[nf] * 4);
  {
    std::vector<int> mark(nc, -1);
    for (int i = 0; i < nf; i++) {
      auto addrow = [&](int j) {
        for (int q = pptr[j]; q < pptr[j + 1]; q++)
          if (mark[pcol[q]] != i) {
            mark[pcol[q]] = i;
            apcol.push_back(pcol[q]);
          }
      };
      addrow(i);
      for (int k = f.offind[i]; k < f.offind[i + 1]; k++) addrow(f.offrow[k] - 1);
      apptr[i + 1] = (int)apcol.size();
    }
  }
  // Galerkin pattern in the reference's order
  std::vector<int> coff, ccol1;
  std::vector<char> hasdiag;
  amg_galerkin_pattern(nf, f.offind.data(), f.offrow.data(), nc, rindex, roffrow, coff, ccol1,
                       hasdiag);
  std::vector<int> ccol(ccol1.size());
  for (size_t q = 0; q < ccol1.size(); q++) ccol[q] = ccol1[q] - 1;
  std::vector<char> hd(hasdiag.begin() + 1, hasdiag.end());
  // operand pairs of A P: entry (i, J) += a_ij P(j, J) over j = i, then the
  // row's columns in CSR order, each P row in order
  const int nap = apptr[nf];
  std::vector<int> ap_ptr(nap + 1, 0), ap_a, ap_b;
  std::vector<int> rp_ptr, rp_a, rp_b;
  bool bad = false;  // an operand outside the symbolic pattern (cannot happen)
  {
    std::vector<int> pos(nc, -1), stamp(nc, -1);
    for (int pass = 0; pass < 2; pass++) {
      std::vector<int> fill;
      std::fill(stamp.begin(), stamp.end(), -1);
      if (pass == 1) {
        for (int e = 0; e < nap; e++) ap_ptr[e + 1] += ap_ptr[e];
        ap_a.resize(ap_ptr[nap]);
        ap_b.resize(ap_ptr[nap]);
        fill.assign(ap_ptr.begin(), ap_ptr.end() - 1);
      }
      for (int i = 0; i < nf; i++) {
        for (int q = apptr[i]; q < apptr[i + 1]; q++) {
          pos[apcol[q]] = q;
          stamp[apcol[q]] = i;
        }
        auto add = [&](int aidx, int j) {
          for (int t = pptr[j]; t < pptr[j + 1]; t++) {
            if (stamp[pcol[t]] != i) {
              bad = true;
              continue;
            }
            const int e = pos[pcol[t]];
            if (pass == 0) {
              ap_ptr[e + 1]++;
            } else {
              ap_a[fill[e]] = aidx;
              ap_b[fill[e]++] = t;
            }
          }
        };
        add(i, i);  // the diagonal sits first in every flat matrix
        for (int k = f.offind[i]; k < f.offind[i + 1]; k++) {
          const int j = f.offrow[k] - 1;
          const int aidx = (l == 0) ? (direction(dm, i, j) + 1) * nf + i : nf + k;
          add(aidx, j);
        }
      }
    }
  }
  // operand pairs of R A P: flat coarse entry (I, J) += P(i, I) (A P)(i, J)
  // over R's row I (fine rows ascending), each A P row in order
  {
    const int ncf = nc + coff[nc];
    rp_ptr.assign(ncf + 1, 0);
    std::vector<int> pos(nc, -1), stamp(nc, -1);
    for (int pass = 0; pass < 2; pass++) {
      std::vector<int> fill;
      std::fill(stamp.begin(), stamp.end(), -1);
      if (pass == 1) {
        for (int e = 0; e < ncf; e++) rp_ptr[e + 1] += rp_ptr[e];
        rp_a.resize(rp_ptr[ncf]);
        rp_b.resize(rp_ptr[ncf]);
        fill.assign(rp_ptr.begin(), rp_ptr.end() - 1);
      }
      for (int I = 0; I < nc; I++) {
        pos[I] = I;
        stamp[I] = I;
        for (int k = coff[I]; k < coff[I + 1]; k++) {
          pos[ccol[k]] = nc + k;
          stamp[ccol[k]] = I;
        }
        for (int t = rindex[I]; t < rindex[I + 1]; t++) {
          const int i = rrow[t], pidx = rpos[t];
          for (int q = apptr[i]; q < apptr[i + 1]; q++) {
            if (stamp[apcol[q]] != I) {
              bad = true;
              continue;
            }
            const int e = pos[apcol[q]];
            if (pass == 0) {
              rp_ptr[e + 1]++;
            } else {
              rp_a[fill[e]] = pidx;
              rp_b[fill[e]++] = q;
            }
          }
        }
      }
    }
  }
  if (bad) return MSG123_EINVAL;
  // sliced ELL layout of the coarse matrix
  const int nsl = (nc + 63) / 64;
  std::vector<int> soff(nsl + 1, 0);
  for (int sl = 0; sl < nsl; sl++) {
    int w = 0;
    for (int r = sl * 64; r < std::min(nc, sl * 64 + 64); r++) w = std::max(w, coff[r + 1] - coff[r]);
    soff[sl + 1] = soff[sl] + w * 64;
  }
  std::vector<int> scol(soff[nsl]);
  for (int sl = 0; sl < nsl; sl++) {
    const int w = (soff[sl + 1] - soff[sl]) / 64;
    for (int lane = 0; lane < 64; lane++) {
      const int r = sl * 64 + lane;
      for (int j = 0; j < w; j++) {
        const int k = soff[sl] + j * 64 + lane;
        if (r < nc && j < coff[r + 1] - coff[r]) scol[k] = ccol[coff[r] + j];
        else scol[k] = r < nc ? r : 0;
      }
    }
  }
  // device arrays
  g->lv.emplace_back();
  DevLevel &C = g->lv.back();
  DevLevel &Fr = g->lv[l];  // taken after emplace_back, which may move the levels
  int rc;
  Fr.np = pptr[nf];
  Fr.nap = apptr[nf];
  Fr.h_pptr = pptr;
  Fr.h_pcol = pcol;
  if ((rc = g->upload(&Fr.pptr, pptr, s)) || (rc = g->upload(&Fr.pcol, pcol, s)) ||
      (rc = g->alloc(&Fr.pval, Fr.np)) || (rc = g->alloc(&Fr.pval32, Fr.np)) ||
      (rc = g->upload(&Fr.smap, smap, s)) || (rc = g->upload(&Fr.pown, pown, s)) ||
      (rc = g->upload(&Fr.rptr, rindex, s)) || (rc = g->upload(&Fr.rrow, rrow, s)) ||
      (rc = g->upload(&Fr.rpos, rpos, s)) || (rc = g->upload(&Fr.apptr, apptr, s)) ||
      (rc = g->upload(&Fr.apcol, apcol, s)) || (rc = g->alloc(&Fr.apval, Fr.nap)) ||
      (rc = g->upload(&Fr.ap_ptr, ap_ptr, s)) || (rc = g->upload(&Fr.ap_a, ap_a, s)) ||
      (rc = g->upload(&Fr.ap_b, ap_b, s)) || (rc = g->upload(&Fr.rp_ptr, rp_ptr, s)) ||
      (rc = g->upload(&Fr.rp_a, rp_a, s)) || (rc = g->upload(&Fr.rp_b, rp_b, s)))
    return rc;
  C.n = nc;
  C.nnz = coff[nc];
  C.nsell = soff[nsl];
  std::vector<char> hdv(hd);
  if ((rc = g->upload(&C.aptr, coff, s)) || (rc = g->upload(&C.acol, ccol, s)) ||
      (rc = g->alloc(&C.flat, (size_t)nc + C.nnz)) ||
      (rc = g->upload(&C.hasdiag, hdv, s)) || (rc = g->upload(&C.soff, soff, s)) ||
      (rc = g->upload(&C.scol, scol, s)) || (rc = g->alloc(&C.sval, C.nsell)) ||
      (rc = g->alloc(&C.diag32, nc)) || (rc = g->alloc(&C.dinv, nc)) ||
      (rc = g->alloc(&C.b, nc)) || (rc = g->alloc(&C.x, nc)) || (rc = g->alloc(&C.y, nc)) ||
      (rc = g->alloc(&C.r, nc)))
    return rc;
  C.adiag = C.flat;
  C.aval = C.flat + nc;
  AMGCHK(hipMemsetAsync(C.sval, 0, sizeof(float) * (size_t)C.nsell, s));
  return MSG123_OK;
}

Mat7f mat7f(const AmgGpu *g) {
  const size_t N = (size_t)g->dm.n;
  const float *A = g->lv[0].A32;
  return Mat7f{{A + N, A + 2 * N, A + 3 * N, A + 4 * N, A + 5 * N, A + 6 * N}, A};
}

Sell sell(const DevLevel &L) { return Sell{L.soff, L.scol, L.sval}; }

// Jacobi sweeps from zero on level l (l >= 1) with right-hand side L.b; the
// result is left in L.x
int coarse_sweeps(AmgGpu *g, int l, int nsweep, hipStream_t s) {
  DevLevel &L = g->lv[l];
  const int gb = grid(L.n);
  if (nsweep <= 0) {
    hipLaunchKernelGGL(k_zero, dim3(gb), dim3(kAB), 0, s, L.n, L.x);
    AMGCHK(hipGetLastError());
    return MSG123_OK;
  }
  hipLaunchKernelGGL(k_first<float>, dim3(gb), dim3(kAB), 0, s, L.n, L.dinv, L.b, L.x);
  for (int k = 1; k < nsweep; k++) {
    hipLaunchKernelGGL(k_vc_sweep, dim3(gb), dim3(kAB), 0, s, L.n, sell(L), L.diag32, L.dinv,
                       L.b, L.x, L.y);
    std::swap(L.x, L.y);
  }
  AMGCHK(hipGetLastError());
  return MSG123_OK;
}

// the matrix the hierarchy is built from: A7 or its projection
const double *source(AmgGpu *g, const double *A7, hipStream_t s) {
  if (!g->proj) return A7;
  hipLaunchKernelGGL(k_amg_project, dim3(grid(g->dm.n)), dim3(kAB), 0, s, g->dm.n, A7, g->proj);
  return g->proj;
}

}  // namespace

int amg_gpu_numeric(AmgGpu *g, const double *A7in, hipStream_t s) {
  const double *A7 = source(g, A7in, s);
  hipLaunchKernelGGL(k_amg_a0, dim3(grid(g->dm.n)), dim3(kAB), 0, s, g->dm.n, A7, g->lv[0].A32,
                     g->lv[0].dinv, (double)g->opt.smooth_omega);
  AMGCHK(hipGetLastError());
  for (int l = 0; l + 1 < (int)g->lv.size(); l++) {
    const int rc = numeric_level(g, l, A7, s);
    if (rc) return rc;
  }
  return finish_numeric(g, s);
}

int amg_gpu_create(const Dims &dm, const int32_t *offind, const int32_t *offrow,
                   const double *A7in, const msg123_amg_opts &opt, hipStream_t s, AmgGpu **out) {
  AmgGpu *g = new AmgGpu();
  g->dm = dm;
  g->opt = opt;
  const int n = dm.n;
  const size_t N = (size_t)n;
  int rc = MSG123_OK;
  auto fail = [&](int r) {
    amg_gpu_destroy(g);
    return r;
  };
  if (opt.project && (rc = g->alloc(&g->proj, 7 * N))) return fail(rc);
  const double *A7 = source(g, A7in, s);
  // the box level: host CSR values from the seven diagonals
  HostCsr f;
  f.n = n;
  f.offind.assign(offind, offind + n + 1);
  f.offrow.assign(offrow, offrow + offind[n]);
  {
    std::vector<double> a7(7 * N);
    AMGCHK(hipMemcpyAsync(a7.data(), A7, 8 * 7 * N, hipMemcpyDeviceToHost, s));
    AMGCHK(hipStreamSynchronize(s));
    f.dmat.assign(a7.begin(), a7.begin() + n);
    f.lumat.resize(f.offind[n]);
    for (int i = 0; i < n; i++)
      for (int k = f.offind[i]; k < f.offind[i + 1]; k++) {
        const int d = direction(dm, i, f.offrow[k] - 1);
        if (d < 0) return fail(MSG123_EINVAL);
        f.lumat[k] = a7[(size_t)(d + 1) * N + i];
      }
  }
  g->lv.emplace_back();
  {
    DevLevel &L0 = g->lv[0];
    L0.n = n;
    if ((rc = g->alloc(&L0.A32, 7 * N)) || (rc = g->alloc(&L0.dinv, N)) ||
        (rc = g->alloc(&L0.x, N)) || (rc = g->alloc(&L0.y, N)) || (rc = g->alloc(&L0.r, N)))
      return fail(rc);
  }
  hipLaunchKernelGGL(k_amg_a0, dim3(grid(n)), dim3(kAB), 0, s, n, A7, g->lv[0].A32, g->lv[0].dinv,
                     (double)opt.smooth_omega);
  const int maxlev = std::max(1, opt.max_levels);
  for (int l = 0; l + 1 < maxlev && f.n > opt.coarse_max; l++) {
    rc = build_step(g, l, f, s);
    if (rc == 1) break;
    if (rc) return fail(rc);
    if ((rc = numeric_level(g, l, A7, s))) return fail(rc);
    // the coarse matrix back on the host for the next aggregation
    DevLevel &C = g->lv[l + 1];
    HostCsr nx;
    nx.n = C.n;
    nx.offind.resize(C.n + 1);
    nx.offrow.resize(C.nnz);
    nx.dmat.resize(C.n);
    nx.lumat.resize(C.nnz);
    AMGCHK(hipMemcpyAsync(nx.offind.data(), C.aptr, 4 * (size_t)(C.n + 1), hipMemcpyDeviceToHost, s));
    AMGCHK(hipMemcpyAsync(nx.offrow.data(), C.acol, 4 * (size_t)C.nnz, hipMemcpyDeviceToHost, s));
    AMGCHK(hipMemcpyAsync(nx.dmat.data(), C.adiag, 8 * (size_t)C.n, hipMemcpyDeviceToHost, s));
    AMGCHK(hipMemcpyAsync(nx.lumat.data(), C.aval, 8 * (size_t)C.nnz, hipMemcpyDeviceToHost, s));
    AMGCHK(hipStreamSynchronize(s));
    for (auto &v : nx.offrow) v += 1;
    f = std::move(nx);
  }
  DevLevel &L = g->lv.back();
  if (g->lv.size() > 1 && L.n <= kDenseMax) {
    g->dense = true;
    if ((rc = g->alloc(&L.inv, (size_t)L.n * L.n))) return fail(rc);
  }
  if ((rc = finish_numeric(g, s))) return fail(rc);
  AMGCHK(hipStreamSynchronize(s));
  *out = g;
  return MSG123_OK;
}

void amg_gpu_destroy(AmgGpu *g) {
  if (!g) return;
  for (void *p : g->allocs) (void)hipFree(p);
  delete g;
}

int amg_gpu_nlevel(const AmgGpu *g) { return g ? (int)g->lv.size() : 0; }

void amg_gpu_level_size(const AmgGpu *g, int l, int32_t *n, int32_t *nnz, int32_t *np) {
  const DevLevel &L = g->lv[l];
  if (n) *n = L.n;
  if (nnz) *nnz = l == 0 ? -1 : L.nnz;
  if (np) *np = l == 0 ? 0 : g->lv[l - 1].np;
}

int amg_gpu_level_get(AmgGpu *g, int l, int32_t *offind, int32_t *offrow, double *dmat,
                      double *lumat, int32_t *pindex, int32_t *poffrow, double *pval,
                      hipStream_t s) {
  if (l < 1 || l >= (int)g->lv.size()) return MSG123_EINVAL;
  const DevLevel &C = g->lv[l], &F = g->lv[l - 1];
  if (offind)
    AMGCHK(hipMemcpyAsync(offind, C.aptr, 4 * (size_t)(C.n + 1), hipMemcpyDeviceToHost, s));
  if (offrow) AMGCHK(hipMemcpyAsync(offrow, C.acol, 4 * (size_t)C.nnz, hipMemcpyDeviceToHost, s));
  if (dmat) AMGCHK(hipMemcpyAsync(dmat, C.adiag, 8 * (size_t)C.n, hipMemcpyDeviceToHost, s));
  if (lumat) AMGCHK(hipMemcpyAsync(lumat, C.aval, 8 * (size_t)C.nnz, hipMemcpyDeviceToHost, s));
  if (pval) AMGCHK(hipMemcpyAsync(pval, F.pval, 8 * (size_t)F.np, hipMemcpyDeviceToHost, s));
  AMGCHK(hipStreamSynchronize(s));
  if (offrow)
    for (int q = 0; q < C.nnz; q++) offrow[q] += 1;
  if (pindex)
    for (size_t q = 0; q < F.h_pptr.size(); q++) pindex[q] = F.h_pptr[q];
  if (poffrow)
    for (size_t q = 0; q < F.h_pcol.size(); q++) poffrow[q] = F.h_pcol[q] + 1;
  return MSG123_OK;
}

int amg_gpu_apply(AmgGpu *g, const double *rin, double *z, hipStream_t s) {
  const int L = (int)g->lv.size();
  const Dims dm = g->dm;
  const int npre = g->opt.npre, npost = g->opt.npost;
  const Mat7f A0 = mat7f(g);
  DevLevel &B = g->lv[0];
  const int gb0 = grid(dm.n);
  if (L == 1) {
    // no coarse level: smoothing sweeps only
    const int ns = std::max(1, npre + npost);
    hipLaunchKernelGGL(k_first<double>, dim3(gb0), dim3(kAB), 0, s, dm.n, B.dinv, rin, B.x);
    for (int k = 1; k < ns; k++) {
      hipLaunchKernelGGL((k_v0_sweep<double, float>), dim3(gb0), dim3(kAB), 0, s, dm, A0, B.dinv,
                         rin, B.x, B.y);
      std::swap(B.x, B.y);
    }
    hipLaunchKernelGGL(k_to_double<double>, dim3(gb0), dim3(kAB), 0, s, dm.n, B.x, z);
    AMGCHK(hipGetLastError());
    return MSG123_OK;
  }
  // ---- down
  for (int l = 0; l < L - 1; l++) {
    DevLevel &F = g->lv[l];
    DevLevel &C = g->lv[l + 1];
    const int gb = grid(F.n);
    if (l == 0) {
      if (npre >= 2) {
        hipLaunchKernelGGL(k_v0_pre2<double>, dim3(gb), dim3(kAB), 0, s, dm, A0, F.dinv, rin, F.x);
      } else if (npre == 1) {
        hipLaunchKernelGGL(k_first<double>, dim3(gb), dim3(kAB), 0, s, dm.n, F.dinv, rin, F.x);
      } else {
        hipLaunchKernelGGL(k_zero, dim3(gb), dim3(kAB), 0, s, dm.n, F.x);
      }
      for (int k = 2; k < npre; k++) {
        hipLaunchKernelGGL((k_v0_sweep<double, float>), dim3(gb), dim3(kAB), 0, s, dm, A0, F.dinv,
                           rin, F.x, F.y);
        std::swap(F.x, F.y);
      }
      hipLaunchKernelGGL(k_v0_resid<double>, dim3(gb), dim3(kAB), 0, s, dm, A0, rin, F.x, F.r);
    } else {
      const Sell As = sell(F);
      if (npre >= 1) {
        hipLaunchKernelGGL(k_first<float>, dim3(gb), dim3(kAB), 0, s, F.n, F.dinv, F.b, F.y);
        if (npre >= 2) {
          hipLaunchKernelGGL(k_vc_pre2, dim3(gb), dim3(kAB), 0, s, F.n, As, F.diag32, F.dinv, F.b,
                             F.y, F.x);
        } else {
          std::swap(F.x, F.y);
        }
      } else {
        hipLaunchKernelGGL(k_zero, dim3(gb), dim3(kAB), 0, s, F.n, F.x);
      }
      for (int k = 2; k < npre; k++) {
        hipLaunchKernelGGL(k_vc_sweep, dim3(gb), dim3(kAB), 0, s, F.n, As, F.diag32, F.dinv, F.b,
                           F.x, F.y);
        std::swap(F.x, F.y);
      }
      hipLaunchKernelGGL(k_vc_resid, dim3(gb), dim3(kAB), 0, s, F.n, As, F.diag32, F.b, F.x, F.r);
    }
    hipLaunchKernelGGL(k_restrict, dim3(grid(C.n)), dim3(kAB), 0, s, C.n, F.rptr, F.rrow, F.rpos,
                       F.pval32, F.r, C.b);
    AMGCHK(hipGetLastError());
  }
  // ---- coarsest
  {
    DevLevel &C = g->lv[L - 1];
    if (g->dense) {
      hipLaunchKernelGGL(k_dense_apply, dim3(grid(C.n)), dim3(kAB), 0, s, C.n, C.inv, C.b, C.x);
      AMGCHK(hipGetLastError());
    } else {
      const int rc = coarse_sweeps(g, L - 1, g->opt.coarse_sweeps, s);
      if (rc) return rc;
    }
  }
  // ---- up
  for (int l = L - 2; l >= 0; l--) {
    DevLevel &F = g->lv[l];
    DevLevel &C = g->lv[l + 1];
    const int gb = grid(F.n);
    hipLaunchKernelGGL(k_prolong, dim3(gb), dim3(kAB), 0, s, F.n, F.pptr, F.pcol, F.pval32, C.x,
                       F.x);
    if (l == 0) {
      if (npost == 0) {
        hipLaunchKernelGGL(k_to_double<double>, dim3(gb), dim3(kAB), 0, s, dm.n, F.x, z);
      }
      for (int k = 0; k < npost; k++) {
        if (k == npost - 1) {
          hipLaunchKernelGGL((k_v0_sweep<double, double>), dim3(gb), dim3(kAB), 0, s, dm, A0,
                             F.dinv, rin, F.x, z);
        } else {
          hipLaunchKernelGGL((k_v0_sweep<double, float>), dim3(gb), dim3(kAB), 0, s, dm, A0,
                             F.dinv, rin, F.x, F.y);
          std::swap(F.x, F.y);
        }
      }
    } else {
      const Sell As = sell(F);
      for (int k = 0; k < npost; k++) {
        hipLaunchKernelGGL(k_vc_sweep, dim3(gb), dim3(kAB), 0, s, F.n, As, F.diag32, F.dinv, F.b,
                           F.x, F.y);
        std::swap(F.x, F.y);
      }
    }
    AMGCHK(hipGetLastError());
  }
  return MSG123_OK;
}

}  // namespace msg123
