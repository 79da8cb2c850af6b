// kernels.hip.h -- device kernels of the Richards Newton-Krylov step (gfx950).
//
// Data layout in HBM (N = nx*ny*nz cells, natural order: x fastest, then y
// north->south, then z top->bottom):
//   * per-cell state and soil parameters: structure of arrays, [N] each;
//   * face coefficients once per face: hydf/abyd of the east, south and
//     lower face of every cell ([N] each, 0 where the face is on the box
//     boundary); the west/north/upper faces of a cell are the east/south/
//     lower faces of its neighbours (the reference's values are bitwise
//     symmetric, checked when the context is created);
//   * the Jacobian as seven diagonals [7][N] in CSR neighbour order
//     (diag, up, N, W, E, S, down), zero where a neighbour is missing.
// Blocks are renumbered (xcd_block) so that each of the 8 XCDs works on one
// contiguous slab of the box.  Most kernels map one thread to one cell; the
// residual and Jacobian kernels march each thread down a z column.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msg123 {

constexpr int kBlock = 256;
constexpr int kXcd = 8;

struct Dims {
  int nx, ny, nz, n, nxy;
};

// Physical block id -> logical block id so that XCD x (= physical % 8) owns
// the contiguous logical range [x*per, (x+1)*per).
__device__ __forceinline__ int xcd_block(int nblocks) {
  const int b = blockIdx.x;
  const int per = (nblocks + kXcd - 1) / kXcd;
  const int xcd = b % kXcd, k = b / kXcd;
  const int full = nblocks - (per - 1) * kXcd;  // XCDs that get `per` blocks
  // blocks are dealt round robin; XCDs [0, full) get per, the rest per-1
  const int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + k;
}

// The Jacobian as seven diagonals in CSR neighbour order (FP64, or FP32 for
// the Newton systems of the line-preconditioned path).
template <typename T>
struct Mat7T {
  const T *d, *up, *n, *w, *e, *s, *dn;
};
using Mat7 = Mat7T<double>;
using Mat7F = Mat7T<float>;

// y_i = d_i x_i + sum over neighbours in CSR order (calc_matvec,
// Linear_solution.f90:917-951).  The diagonals are read once per product:
// non-temporal loads keep them from evicting x, which is read by up to seven
// rows, from the L2 (9% faster at 4.2M cells).  Every load is issued
// whatever the neighbour (a missing one reads the cell itself; its term is
// not added), so no branch stands around a load and each wait is for the
// loads it needs.
template <typename T>
__device__ __forceinline__ double apply7(const Mat7T<T> &A, const Dims &dm, const double *x,
                                         int i) {
  const int xx = i % dm.nx, y = (i / dm.nx) % dm.ny, z = i / dm.nxy;
  const bool bu = z > 0, bn = y > 0, bw = xx > 0, be = xx < dm.nx - 1, bs = y < dm.ny - 1,
             bd = z < dm.nz - 1;
  const double ad = (double)__builtin_nontemporal_load(&A.d[i]);
  const double au = (double)__builtin_nontemporal_load(&A.up[i]);
  const double an = (double)__builtin_nontemporal_load(&A.n[i]);
  const double aw = (double)__builtin_nontemporal_load(&A.w[i]);
  const double ae = (double)__builtin_nontemporal_load(&A.e[i]);
  const double as = (double)__builtin_nontemporal_load(&A.s[i]);
  const double adn = (double)__builtin_nontemporal_load(&A.dn[i]);
  const double x0 = x[i];
  const double xu = x[bu ? i - dm.nxy : i], xn = x[bn ? i - dm.nx : i];
  const double xw = x[bw ? i - 1 : i], xe = x[be ? i + 1 : i];
  const double xs = x[bs ? i + dm.nx : i], xd = x[bd ? i + dm.nxy : i];
  double t = ad * x0;
  if (bu) t = t + au * xu;
  if (bn) t = t + an * xn;
  if (bw) t = t + aw * xw;
  if (be) t = t + ae * xe;
  if (bs) t = t + as * xs;
  if (bd) t = t + adn * xd;
  return t;
}

// The six neighbours of cell i in CSR order (up, N, W, E, S, down); ok[d]
// is false where the neighbour lies outside the box.
__device__ __forceinline__ void neighbours(const Dims &dm, int i, int (&j)[6], bool (&ok)[6]) {
  const int xx = i % dm.nx, y = (i / dm.nx) % dm.ny, z = i / dm.nxy;
  j[0] = i - dm.nxy; ok[0] = z > 0;
  j[1] = i - dm.nx;  ok[1] = y > 0;
  j[2] = i - 1;      ok[2] = xx > 0;
  j[3] = i + 1;      ok[3] = xx < dm.nx - 1;
  j[4] = i + dm.nx;  ok[4] = y < dm.ny - 1;
  j[5] = i + dm.nxy; ok[5] = z < dm.nz - 1;
}

// ---------------------------------------------------------------------------
// wave64 / block reductions (deterministic for a fixed launch shape)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sums(double (&v)[NV], double *partials, int slot,
                                           int nslots) {
  constexpr int NW = (NT + 63) / 64;
  __shared__ double red[NV][NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) red[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = red[k][0];
#pragma unroll
      for (int q = 1; q < NW; q++) s += red[k][q];
      partials[(size_t)k * nslots + slot] = s;
    }
  }
}

}  // namespace msg123
