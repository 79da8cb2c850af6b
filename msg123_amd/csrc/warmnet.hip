// warmnet.hip -- the learned Newton warm start: a 3D convolution network
// (four 3x3x3 layers, 4 -> 16 -> 16 -> 16 -> 1 channels, ReLU between,
// zero padding) that predicts the correction to the linear extrapolation of
// the head from the previous two heads, the recharge and lnKs.  Training
// (forward, backward, Adam) and inference run here, on MFMA.
//
// Precision: weights and activations are stored in bf16 and multiplied on
// the matrix cores (v_mfma_f32_16x16x16_bf16) with fp32 accumulation;
// master weights, biases, the optimizer state, the network output and the
// loss are fp32.
//
// Layouts in HBM (N cells in the solver's natural order):
//   features  [N][4]  bf16;  activations and their gradients [N][16] bf16
//   (one 32-byte record per cell, so a lane's four channels are one 8-byte
//   load); output [N] fp32.
//   master parameters, flat fp32: layer l holds W[t][co][ci] (t = 3x3x3 tap,
//   (dz+1)*9 + (dy+1)*3 + (dx+1)) then b[co].
//   packed forward weights wf[chunk][co16][k16] bf16, one MFMA A operand per
//   chunk: chunk = tap for 16 input channels, four taps x 4 channels for the
//   first layer.  Packed backward weights wb[t][ci][co] = W[26-t][co][ci]:
//   the input gradient is the same convolution with mirrored taps.
//
// Convolution as an implicit GEMM on 16x16 tiles: D[co][cell] = sum_k
// A[co][k] B[k][cell], one wave per 16 consecutive cells; per MFMA lane l
// supplies A[l%16][4(l/16)..+3] (weights) and B[4(l/16)..+3][l%16] (four
// channels of one neighbour) and receives D[4(l/16)..+3][l%16] -- four
// consecutive output channels of one cell, stored as one 8-byte write.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "kernels.hip.h"
#include "msg123_gpu.h"
#include "warmnet.h"

namespace msg123 {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kH = 16;                      // hidden channels
constexpr int kTaps = 27;
constexpr int kCin[4] = {4, 16, 16, 16};
constexpr int kCout[4] = {16, 16, 16, 1};
constexpr int kDwBlocks = 512;              // partial-sum slots of the weight gradient

__host__ __device__ constexpr int wcount(int l) { return kTaps * kCout[l] * kCin[l]; }
__host__ __device__ constexpr int poff(int l) {
  return l == 0 ? 0 : poff(l - 1) + wcount(l - 1) + kCout[l - 1];
}
constexpr int kNParam = poff(3) + wcount(3) + kCout[3];
static_assert(kNParam == MSG123_NET_NPARAM, "parameter count");
__host__ __device__ constexpr int nchunk(int cin) { return cin == 16 ? kTaps : (kTaps + 3) / 4; }

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f4 mfma16(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// neighbour of cell (x,y,z) at tap t, or -1 outside the box
__device__ __forceinline__ int nbr(const Dims &d, int x, int y, int z, int t) {
  const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  const int xx = x + dx, yy = y + dy, zz = z + dz;
  if (xx < 0 || xx >= d.nx || yy < 0 || yy >= d.ny || zz < 0 || zz >= d.nz) return -1;
  return (zz * d.ny + yy) * d.nx + xx;
}

// ---------------------------------------------------------------------------
// Convolution (forward, and the input gradient with mirrored weights).
// EPI 0: bias + ReLU -> bf16 [N][16];  EPI 1: bias -> fp32 out[N] (channel 0);
// EPI 2: input gradient: times (act > 0) -> bf16 [N][16].
template <int CIN, int EPI>
__global__ void __launch_bounds__(256)
k_conv(Dims d, int nblocks, const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
       const float *__restrict__ bias, const uint16_t *__restrict__ act,
       uint16_t *__restrict__ y, float *__restrict__ yout) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = xcd_block(nblocks) * 4 + wave;
  const int j = lane & 15, g = lane >> 4;
  const int cell = tile * 16 + j;
  const bool live = cell < d.n;
  const int cx = cell % d.nx, cy = (cell / d.nx) % d.ny, cz = cell / d.nxy;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  constexpr int NCH = nchunk(CIN);
  // Every tap's input gathered first, each load issued whatever the tap
  // (outside the box: cell 0's values, replaced by zeros), so all of them
  // are in flight together; then the MFMA chain in tap order.
  s4 bs[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int t = (CIN == 16) ? ch : 4 * ch + g;
    const int c0 = (CIN == 16) ? 4 * g : 0;
    const int q = (live && t < kTaps) ? nbr(d, cx, cy, cz, t) : -1;
    const s4 v = *reinterpret_cast<const s4 *>(x + (size_t)(q >= 0 ? q : 0) * CIN + c0);
    const s4 zero4 = {0, 0, 0, 0};
    bs[ch] = q >= 0 ? v : zero4;
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const s4 a = *reinterpret_cast<const s4 *>(w + ((size_t)ch * 16 + j) * 16 + 4 * g);
    acc = mfma16(a, bs[ch], acc);
  }
  if (!live) return;
  if (EPI == 1) {
    if (g == 0) yout[cell] = acc[0] + bias[0];
    return;
  }
  s4 o;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int c = 4 * g + r;
    float v = acc[r];
    if (EPI == 0) {
      v += bias[c];
      v = v > 0.f ? v : 0.f;
    } else {
      v = bf2f(act[(size_t)cell * kH + c]) > 0.f ? v : 0.f;
    }
    o[r] = (short)f2bf(v);
  }
  *reinterpret_cast<s4 *>(y + (size_t)cell * kH + 4 * g) = o;
}

// The forward convolution (EPI 0 and 1 of k_conv) on boxes with nx a
// multiple of 16 and ny of 4: a block owns 16 x 4 cells of one level, each
// wave one row of 16, and first stages the 18 x 6 x 3 cells around them
// (zeros outside the box) in LDS, so every input is read from HBM about
// 1.7 times instead of once per tap through the caches.  The MFMA chain and
// its operands are k_conv's, so the outputs are the same bit for bit.
// A block takes ZL levels (2 where nz is even, else 1): ZL + 2 staged
// planes, and each wave runs the levels' MFMA chains side by side (four
// levels measured no faster than two).
constexpr int kTX = 16, kTY = 4, kHX = kTX + 2, kHY = kTY + 2;
template <int CIN, int EPI, int ZL>
__global__ void __launch_bounds__(256)
k_conv_tile(Dims d, const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
            const float *__restrict__ bias, uint16_t *__restrict__ y,
            float *__restrict__ yout) {
  constexpr int kHalo = kHX * kHY * (ZL + 2);
  __shared__ __attribute__((aligned(16))) uint16_t tile[kHalo * CIN];
  const int ntx = d.nx / kTX, nty = d.ny / kTY;
  const int b = blockIdx.x;
  const int bx = b % ntx, by = (b / ntx) % nty, bz = (b / (ntx * nty)) * ZL;
  const int x0 = bx * kTX, y0 = by * kTY;
  // stage: 4 bf16 (8 bytes) per load
  constexpr int kParts = CIN / 4;
  for (int q = threadIdx.x; q < kHalo * kParts; q += 256) {
    const int cl = q / kParts, part = q - cl * kParts;
    const int lz = cl / (kHX * kHY), ly = (cl / kHX) % kHY, lx = cl % kHX;
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly, gz = bz - 1 + lz;
    const bool in = gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz;
    const size_t gi = in ? ((size_t)gz * d.ny + gy) * d.nx + gx : 0;
    const s4 v = *reinterpret_cast<const s4 *>(x + gi * CIN + 4 * part);
    const s4 zero4 = {0, 0, 0, 0};
    *reinterpret_cast<s4 *>(tile + (size_t)cl * CIN + 4 * part) = in ? v : zero4;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  f4 acc[ZL];
#pragma unroll
  for (int l = 0; l < ZL; l++) acc[l] = f4{0.f, 0.f, 0.f, 0.f};
  constexpr int NCH = nchunk(CIN);
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int t = (CIN == 16) ? ch : 4 * ch + g;
    const int c0 = (CIN == 16) ? 4 * g : 0;
    const int tt = t < kTaps ? t : 0;
    const int dz = tt / 9, dy = (tt / 3) % 3, dx = tt % 3;  // offsets + 1
    const s4 a = *reinterpret_cast<const s4 *>(w + ((size_t)ch * 16 + j) * 16 + 4 * g);
    const s4 zero4 = {0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < ZL; l++) {
      const int cl = ((dz + l) * kHY + wave + dy) * kHX + j + dx;
      const s4 v = *reinterpret_cast<const s4 *>(tile + (size_t)cl * CIN + c0);
      acc[l] = mfma16(a, t < kTaps ? v : zero4, acc[l]);
    }
  }
#pragma unroll
  for (int l = 0; l < ZL; l++) {
    const int cell = ((bz + l) * d.ny + y0 + wave) * d.nx + x0 + j;
    if (EPI == 1) {
      if (g == 0) yout[cell] = acc[l][0] + bias[0];
      continue;
    }
    s4 o;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int c = 4 * g + r;
      float v = acc[l][r] + bias[c];
      v = v > 0.f ? v : 0.f;
      o[r] = (short)f2bf(v);
    }
    *reinterpret_cast<s4 *>(y + (size_t)cell * kH + 4 * g) = o;
  }
}

// Weight and bias gradient of one layer: dW[t][co][ci] = sum_p dY[p][co] X[p+off(t)][ci],
// db[co] = sum_p dY[p][co].  A = dY^T (16 x 16 cells), B = shifted X (16 cells x 16
// columns: ci, or four taps x 4 channels for CIN = 4).  Per block: partial sums in
// partials[block][NCH*256 + 16].
template <int CIN>
__global__ void __launch_bounds__(256)
k_conv_dw(Dims d, const uint16_t *__restrict__ dy, const uint16_t *__restrict__ x,
          float *__restrict__ partials) {
  constexpr int NCH = nchunk(CIN);
  constexpr int NW = NCH * 256 + 16;
  __shared__ float red[NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  f4 acc[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) acc[c] = f4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int nchunks = (d.n + 15) / 16;
  for (int k = blockIdx.x * 4 + wave; k < nchunks; k += gridDim.x * 4) {
    const int p0 = k * 16 + 4 * g;  // this lane's four cells p0..p0+3
    s4 a;
    int px[4], py[4], pz[4];
    bool pl[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int p = p0 + r;
      pl[r] = p < d.n;
      px[r] = p % d.nx;
      py[r] = (p / d.nx) % d.ny;
      pz[r] = p / d.nxy;
      a[r] = pl[r] ? (short)dy[(size_t)p * kH + j] : (short)0;
      bsum += pl[r] ? bf2f((uint16_t)a[r]) : 0.f;
    }
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
      const int t = (CIN == 16) ? ch : 4 * ch + j / 4;
      const int ci = (CIN == 16) ? j : j % 4;
      s4 b;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        short v = 0;
        if (pl[r] && t < kTaps) {
          const int q = nbr(d, px[r], py[r], pz[r], t);
          if (q >= 0) v = (short)x[(size_t)q * CIN + ci];
        }
        b[r] = v;
      }
      acc[ch] = mfma16(a, b, acc[ch]);
    }
  }
  // acc[ch][r] = dW[chunk ch][co = 4g + r][column j]; bias: lanes j, j+16,
  // j+32, j+48 hold channel j.  The four waves add into LDS in turn.
  bsum += __shfl_xor(bsum, 16, 64);
  bsum += __shfl_xor(bsum, 32, 64);
  for (int w = 0; w < 4; w++) {
    if (wave == w) {
#pragma unroll
      for (int ch = 0; ch < NCH; ch++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          float &dst = red[ch * 256 + (4 * g + r) * 16 + j];
          dst = (w == 0) ? acc[ch][r] : dst + acc[ch][r];
        }
      if (g == 0) {
        float &dst = red[NCH * 256 + j];
        dst = (w == 0) ? bsum : dst + bsum;
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < NW; i += blockDim.x)
    partials[(size_t)blockIdx.x * NW + i] = red[i];
}

// Sum the block partials (fixed order) and scatter them into the flat
// gradient layout of layer l.
__global__ void k_dw_final(const float *__restrict__ partials, int nblocks, int nw, int cin,
                           int cout, int woff, int boff, float *__restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; b++) s += partials[(size_t)b * nw + i];
  const int nch = nchunk(cin);
  if (i >= nch * 256) {
    const int co = i - nch * 256;
    if (co < cout) grad[boff + co] = s;
    return;
  }
  const int ch = i / 256, co = (i / 16) % 16, col = i % 16;
  const int t = (cin == 16) ? ch : 4 * ch + col / 4;
  const int ci = (cin == 16) ? col : col % 4;
  if (co < cout && t < kTaps) grad[woff + (t * cout + co) * cin + ci] = s;
}

// MSE loss of the fp32 output against the target; dY (bf16, channel 0) = 2 (y - t) / n
__global__ void __launch_bounds__(256)
k_loss(int n, int nblocks, const float *__restrict__ y, const float *__restrict__ tgt,
       uint16_t *__restrict__ dy, double *__restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    const float e = y[i] - tgt[i];
    v[0] = (double)e * (double)e;
    s4 z = {0, 0, 0, 0};
    s4 first = z;
    first[0] = (short)f2bf(2.0f * e / (float)n);
    uint16_t *row = dy + (size_t)i * kH;
    *reinterpret_cast<s4 *>(row) = first;
    *reinterpret_cast<s4 *>(row + 4) = z;
    *reinterpret_cast<s4 *>(row + 8) = z;
    *reinterpret_cast<s4 *>(row + 12) = z;
  }
  block_sums<1>(v, partials, blockIdx.x, nblocks);
}

__global__ void k_sum(const double *__restrict__ partials, int n, double *__restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) a += partials[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

__global__ void k_adam(int np, float lr, float b1, float b2, float eps, float c1, float c2,
                       float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m,
                       float *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Master fp32 parameters -> packed bf16 forward (wf) and backward (wb) weights of layer l.
__global__ void k_pack(const float *__restrict__ p, int cin, int cout, int woff,
                       uint16_t *__restrict__ wf, uint16_t *__restrict__ wb) {
  const int nf = nchunk(cin) * 256;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nf) {
    const int ch = i / 256, co = (i / 16) % 16, k = i % 16;
    const int t = (cin == 16) ? ch : 4 * ch + k / 4;
    const int ci = (cin == 16) ? k : k % 4;
    float v = 0.f;
    if (co < cout && t < kTaps) v = p[woff + (t * cout + co) * cin + ci];
    wf[i] = f2bf(v);
  }
  if (wb && i < kTaps * 256) {
    const int t = i / 256, ci = (i / 16) % 16, co = i % 16;
    float v = 0.f;
    if (co < cout && ci < cin) v = p[woff + ((kTaps - 1 - t) * cout + co) * cin + ci];
    wb[i] = f2bf(v);
  }
}

// Input features of one cell and the linear extrapolation h_lin (scaled units):
// f0 = r (h_n - h_nm1) L / s, f1 = (h_n - top) L / 1 m, f2 = recharge depth of the
// step / 1 mm on the top layer, f3 = (ln Ks + 11.5129) / 1.5; s = RMS of the
// extrapolation increment in metres (global, from the first pass).
__global__ void __launch_bounds__(256)
k_incr_rms(Dims d, int nblocks, const double *__restrict__ hn, const double *__restrict__ hm,
           double r, double L, double *__restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < d.n) {
    const double e = r * (hn[i] - hm[i]) * L;
    v[0] = e * e;
  }
  block_sums<1>(v, partials, blockIdx.x, nblocks);
}

__global__ void __launch_bounds__(256)
k_features(Dims d, const double *__restrict__ hn, const double *__restrict__ hm,
           const double *__restrict__ top, const float *__restrict__ lnks, double r, double L,
           double rech_depth, const double *__restrict__ sum2, uint16_t *__restrict__ xf,
           double *__restrict__ hlin, double *__restrict__ scale_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double s = sqrt(sum2[0] / (double)d.n) + 1.0e-12;
  if (i == 0 && scale_out) scale_out[0] = s;
  if (i >= d.n) return;
  const double inc = r * (hn[i] - hm[i]);
  if (hlin) hlin[i] = hn[i] + inc;
  s4 o;
  o[0] = (short)f2bf((float)(inc * L / s));
  o[1] = (short)f2bf((float)((hn[i] - top[i]) * L));
  o[2] = (short)f2bf(i < d.nxy ? (float)(rech_depth * 1.0e3) : 0.f);
  o[3] = (short)f2bf((lnks[i] + 11.512925f) / 1.5f);
  *reinterpret_cast<s4 *>(xf + (size_t)i * 4) = o;
}

// guess = h_lin + s y / L where the cell is unsaturated at h_n, h_lin elsewhere
__global__ void k_apply(int n, const double *__restrict__ hlin, const float *__restrict__ y,
                        const double *__restrict__ scale, double L, const double *__restrict__ hn,
                        const double *__restrict__ top, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool unsat = hn[i] < top[i];
  // unsaturated cells move at most half way to saturation (the extrapolation
  // obeys the same bound)
  out[i] = unsat ? fmin(hlin[i] + scale[0] * (double)y[i] / L, hn[i] + 0.5 * (top[i] - hn[i]))
                 : hlin[i];
}

// target of a training sample: (h_next - h_lin) L / s
__global__ void k_target(int n, const double *__restrict__ hnext, const double *__restrict__ hlin,
                         const double *__restrict__ scale, double L, float *__restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = (float)((hnext[i] - hlin[i]) * L / scale[0]);
}

}  // namespace msg123

using namespace msg123;

struct msg123_net {
  Dims d{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wf[4] = {}, *wb[4] = {};
  uint16_t *act[3] = {}, *dact[2] = {}, *dy = nullptr;
  float *yout = nullptr, *dwpart = nullptr;
  double *dpart = nullptr, *dsum = nullptr, *h_sum = nullptr;
  int adam_t = 0;
};

namespace {

#define NETCHK(x)                                   \
  do {                                              \
    if ((x) != hipSuccess) return MSG123_EDEVICE;   \
  } while (0)

int nblk(int n) { return (n + 255) / 256; }
int conv_blocks(int n) { return ((n + 15) / 16 + 3) / 4; }

template <typename T>
int nalloc(T **p, size_t count) {
  return hipMalloc((void **)p, sizeof(T) * (count ? count : 1)) == hipSuccess ? MSG123_OK
                                                                            : MSG123_ENOMEM;
}

int pack_all(msg123_net *net) {
  for (int l = 0; l < 4; l++) {
    hipLaunchKernelGGL(k_pack, dim3((kTaps * 256 + 255) / 256), dim3(256), 0, net->stream,
                       net->params, kCin[l], kCout[l], poff(l), net->wf[l], net->wb[l]);
  }
  NETCHK(hipGetLastError());
  return MSG123_OK;
}

// forward pass on features x; output in net->yout
int forward(msg123_net *net, const uint16_t *x) {
  const Dims d = net->d;
  const int nb = conv_blocks(d.n);
  hipStream_t s = net->stream;
  const float *p = net->params;
  if (d.nx % kTX == 0 && d.ny % kTY == 0) {
    auto layers = [&](auto zl) {
      constexpr int ZL = decltype(zl)::value;
      const int nt = (d.nx / kTX) * (d.ny / kTY) * (d.nz / ZL);
      hipLaunchKernelGGL((k_conv_tile<4, 0, ZL>), dim3(nt), dim3(256), 0, s, d, x, net->wf[0],
                         p + poff(0) + wcount(0), net->act[0], nullptr);
      hipLaunchKernelGGL((k_conv_tile<16, 0, ZL>), dim3(nt), dim3(256), 0, s, d, net->act[0],
                         net->wf[1], p + poff(1) + wcount(1), net->act[1], nullptr);
      hipLaunchKernelGGL((k_conv_tile<16, 0, ZL>), dim3(nt), dim3(256), 0, s, d, net->act[1],
                         net->wf[2], p + poff(2) + wcount(2), net->act[2], nullptr);
      hipLaunchKernelGGL((k_conv_tile<16, 1, ZL>), dim3(nt), dim3(256), 0, s, d, net->act[2],
                         net->wf[3], p + poff(3) + wcount(3), nullptr, net->yout);
    };
    if (d.nz % 2 == 0) layers(std::integral_constant<int, 2>{});
    else layers(std::integral_constant<int, 1>{});
    NETCHK(hipGetLastError());
    return MSG123_OK;
  }
  hipLaunchKernelGGL((k_conv<4, 0>), dim3(nb), dim3(256), 0, s, d, nb, x, net->wf[0],
                     p + poff(0) + wcount(0), nullptr, net->act[0], nullptr);
  hipLaunchKernelGGL((k_conv<16, 0>), dim3(nb), dim3(256), 0, s, d, nb, net->act[0], net->wf[1],
                     p + poff(1) + wcount(1), nullptr, net->act[1], nullptr);
  hipLaunchKernelGGL((k_conv<16, 0>), dim3(nb), dim3(256), 0, s, d, nb, net->act[1], net->wf[2],
                     p + poff(2) + wcount(2), nullptr, net->act[2], nullptr);
  hipLaunchKernelGGL((k_conv<16, 1>), dim3(nb), dim3(256), 0, s, d, nb, net->act[2], net->wf[3],
                     p + poff(3) + wcount(3), nullptr, nullptr, net->yout);
  NETCHK(hipGetLastError());
  return MSG123_OK;
}

int dw_layer(msg123_net *net, int l, const uint16_t *dyl, const uint16_t *xl) {
  const int nw = nchunk(kCin[l]) * 256 + 16;
  if (kCin[l] == 16)
    hipLaunchKernelGGL(k_conv_dw<16>, dim3(kDwBlocks), dim3(256), 0, net->stream, net->d, dyl, xl,
                       net->dwpart);
  else
    hipLaunchKernelGGL(k_conv_dw<4>, dim3(kDwBlocks), dim3(256), 0, net->stream, net->d, dyl, xl,
                       net->dwpart);
  hipLaunchKernelGGL(k_dw_final, dim3((nw + 255) / 256), dim3(256), 0, net->stream, net->dwpart,
                     kDwBlocks, nw, kCin[l], kCout[l], poff(l), poff(l) + wcount(l), net->grad);
  NETCHK(hipGetLastError());
  return MSG123_OK;
}

// loss and gradient of one sample into net->grad; *loss = mean squared error
int backward(msg123_net *net, const uint16_t *x, const float *tgt, double *loss) {
  const Dims d = net->d;
  const int nb = conv_blocks(d.n), nbl = nblk(d.n);
  hipStream_t s = net->stream;
  int rc;
  if ((rc = forward(net, x))) return rc;
  hipLaunchKernelGGL(k_loss, dim3(nbl), dim3(256), 0, s, d.n, nbl, net->yout, tgt, net->dy,
                     net->dpart);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, net->dpart, nbl, net->dsum);
  NETCHK(hipGetLastError());
  // layer 4 (no activation)
  if ((rc = dw_layer(net, 3, net->dy, net->act[2]))) return rc;
  // layer 3: d3 = conv^T(dy) * (a3 > 0)
  hipLaunchKernelGGL((k_conv<16, 2>), dim3(nb), dim3(256), 0, s, d, nb, net->dy, net->wb[3],
                     nullptr, net->act[2], net->dact[0], nullptr);
  if ((rc = dw_layer(net, 2, net->dact[0], net->act[1]))) return rc;
  hipLaunchKernelGGL((k_conv<16, 2>), dim3(nb), dim3(256), 0, s, d, nb, net->dact[0], net->wb[2],
                     nullptr, net->act[1], net->dact[1], nullptr);
  if ((rc = dw_layer(net, 1, net->dact[1], net->act[0]))) return rc;
  hipLaunchKernelGGL((k_conv<16, 2>), dim3(nb), dim3(256), 0, s, d, nb, net->dact[1], net->wb[1],
                     nullptr, net->act[0], net->dact[0], nullptr);
  if ((rc = dw_layer(net, 0, net->dact[0], x))) return rc;
  NETCHK(hipGetLastError());
  NETCHK(hipMemcpyAsync(net->h_sum, net->dsum, sizeof(double), hipMemcpyDeviceToHost, s));
  NETCHK(hipStreamSynchronize(s));
  if (loss) *loss = net->h_sum[0] / (double)d.n;
  return MSG123_OK;
}

}  // namespace

namespace msg123 {

int net_n(const msg123_net *net) { return net ? net->d.n : 0; }
hipStream_t net_stream(const msg123_net *net) { return net->stream; }

int net_features(msg123_net *net, const double *hn, const double *hm, const double *top,
                 const float *lnks, double r, double L, double rech_depth, uint16_t *x,
                 double *hlin, double *scale_dev) {
  const Dims d = net->d;
  const int nbl = nblk(d.n);
  hipLaunchKernelGGL(k_incr_rms, dim3(nbl), dim3(256), 0, net->stream, d, nbl, hn, hm, r, L,
                     net->dpart);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, net->stream, net->dpart, nbl, net->dsum);
  hipLaunchKernelGGL(k_features, dim3(nbl), dim3(256), 0, net->stream, d, hn, hm, top, lnks, r, L,
                     rech_depth, net->dsum, x, hlin, scale_dev);
  NETCHK(hipGetLastError());
  return MSG123_OK;
}

int net_guess(msg123_net *net, const uint16_t *x, const double *hlin, const double *scale_dev,
              double L, const double *hn, const double *top, double *out) {
  int rc;
  if ((rc = forward(net, x))) return rc;
  hipLaunchKernelGGL(k_apply, dim3(nblk(net->d.n)), dim3(256), 0, net->stream, net->d.n, hlin,
                     net->yout, scale_dev, L, hn, top, out);
  NETCHK(hipGetLastError());
  return MSG123_OK;
}

}  // namespace msg123

extern "C" int msg123_net_create(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, void *stream,
                                 msg123_net **out) {
  if (!out || nx < 1 || ny < 1 || nz < 1) return MSG123_EINVAL;
  *out = nullptr;
  const long long n64 = (long long)nx * ny * nz;
  if (n64 > (1LL << 30)) return MSG123_EINVAL;
  msg123_net *net = new msg123_net();
  net->d = Dims{nx, ny, nz, (int)n64, nx * ny};
  const size_t N = (size_t)n64;
  auto fail = [&](int e) {
    msg123_net_destroy(net);
    return e;
  };
  if (stream) {
    net->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess)
      return fail(MSG123_EDEVICE);
    net->own_stream = true;
  }
  int rc;
  if ((rc = nalloc(&net->params, kNParam)) || (rc = nalloc(&net->grad, kNParam)) ||
      (rc = nalloc(&net->m, kNParam)) || (rc = nalloc(&net->v, kNParam)) ||
      (rc = nalloc(&net->dy, N * kH)) || (rc = nalloc(&net->yout, N)) ||
      (rc = nalloc(&net->dwpart, (size_t)kDwBlocks * (kTaps * 256 + 16))) ||
      (rc = nalloc(&net->dpart, (size_t)nblk((int)N))) || (rc = nalloc(&net->dsum, 1)))
    return fail(rc);
  for (int k = 0; k < 3; k++)
    if ((rc = nalloc(&net->act[k], N * kH))) return fail(rc);
  for (int k = 0; k < 2; k++)
    if ((rc = nalloc(&net->dact[k], N * kH))) return fail(rc);
  for (int l = 0; l < 4; l++)
    if ((rc = nalloc(&net->wf[l], (size_t)nchunk(kCin[l]) * 256)) ||
        (rc = nalloc(&net->wb[l], (size_t)kTaps * 256)))
      return fail(rc);
  if (hipHostMalloc((void **)&net->h_sum, sizeof(double)) != hipSuccess) return fail(MSG123_ENOMEM);
  // He-uniform weights from a fixed generator (splitmix64), zero biases
  std::vector<float> p(kNParam, 0.0f);
  uint64_t st = seed;
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < 4; l++) {
    const double lim = std::sqrt(6.0 / (double)(kTaps * kCin[l])) * (l == 3 ? 0.1 : 1.0);
    for (int i = 0; i < wcount(l); i++) p[poff(l) + i] = (float)((2.0 * next() - 1.0) * lim);
  }
  if (hipMemcpy(net->params, p.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(net->m, 0, sizeof(float) * kNParam) != hipSuccess ||
      hipMemset(net->v, 0, sizeof(float) * kNParam) != hipSuccess)
    return fail(MSG123_EDEVICE);
  if ((rc = pack_all(net))) return fail(rc);
  if (hipStreamSynchronize(net->stream) != hipSuccess) return fail(MSG123_EDEVICE);
  *out = net;
  return MSG123_OK;
}

extern "C" void msg123_net_destroy(msg123_net *net) {
  if (!net) return;
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  void *bufs[] = {net->params, net->grad, net->m, net->v, net->dy, net->yout, net->dwpart,
                  net->dpart, net->dsum, net->act[0], net->act[1], net->act[2], net->dact[0],
                  net->dact[1], net->wf[0], net->wf[1], net->wf[2], net->wf[3], net->wb[0],
                  net->wb[1], net->wb[2], net->wb[3]};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (net->h_sum) (void)hipHostFree(net->h_sum);
  if (net->own_stream && net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

extern "C" int msg123_net_params(msg123_net *net, float *host, int32_t set) {
  if (!net || !host) return MSG123_EINVAL;
  if (set) {
    NETCHK(hipMemcpyAsync(net->params, host, sizeof(float) * kNParam, hipMemcpyHostToDevice,
                          net->stream));
    int rc = pack_all(net);
    if (rc) return rc;
  } else {
    NETCHK(hipMemcpyAsync(host, net->params, sizeof(float) * kNParam, hipMemcpyDeviceToHost,
                          net->stream));
  }
  NETCHK(hipStreamSynchronize(net->stream));
  return MSG123_OK;
}

extern "C" int msg123_net_forward(msg123_net *net, const uint16_t *x_dev, float *y_dev) {
  if (!net || !x_dev || !y_dev) return MSG123_EINVAL;
  int rc = forward(net, x_dev);
  if (rc) return rc;
  NETCHK(hipMemcpyAsync(y_dev, net->yout, sizeof(float) * net->d.n, hipMemcpyDeviceToDevice,
                        net->stream));
  NETCHK(hipStreamSynchronize(net->stream));
  return MSG123_OK;
}

extern "C" int msg123_net_grad(msg123_net *net, const uint16_t *x_dev, const float *t_dev,
                               float *grad_host, double *loss) {
  if (!net || !x_dev || !t_dev) return MSG123_EINVAL;
  int rc = backward(net, x_dev, t_dev, loss);
  if (rc) return rc;
  if (grad_host) {
    NETCHK(hipMemcpyAsync(grad_host, net->grad, sizeof(float) * kNParam, hipMemcpyDeviceToHost,
                          net->stream));
    NETCHK(hipStreamSynchronize(net->stream));
  }
  return MSG123_OK;
}

extern "C" int msg123_net_train_step(msg123_net *net, const uint16_t *x_dev, const float *t_dev,
                                     float lr, double *loss) {
  if (!net || !x_dev || !t_dev || !(lr > 0.f)) return MSG123_EINVAL;
  int rc = backward(net, x_dev, t_dev, loss);
  if (rc) return rc;
  net->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)net->adam_t), c2 = 1.f - std::pow(b2, (float)net->adam_t);
  hipLaunchKernelGGL(k_adam, dim3((kNParam + 255) / 256), dim3(256), 0, net->stream, kNParam, lr,
                     b1, b2, 1.0e-8f, c1, c2, net->params, net->grad, net->m, net->v);
  NETCHK(hipGetLastError());
  if ((rc = pack_all(net))) return rc;
  NETCHK(hipStreamSynchronize(net->stream));
  return MSG123_OK;
}

extern "C" int msg123_net_features(msg123_net *net, const double *hn_dev, const double *hm_dev,
                                   const double *top_dev, const float *lnks_dev, double r,
                                   double len_scal, double rech_depth, uint16_t *x_dev,
                                   double *hlin_dev, double *scale_dev) {
  if (!net || !hn_dev || !hm_dev || !top_dev || !lnks_dev || !x_dev || !scale_dev)
    return MSG123_EINVAL;
  int rc = net_features(net, hn_dev, hm_dev, top_dev, lnks_dev, r, len_scal, rech_depth, x_dev,
                        hlin_dev, scale_dev);
  if (rc) return rc;
  NETCHK(hipStreamSynchronize(net->stream));
  return MSG123_OK;
}

extern "C" int msg123_net_target(msg123_net *net, const double *hnext_dev, const double *hlin_dev,
                                 const double *scale_dev, double len_scal, float *t_dev) {
  if (!net || !hnext_dev || !hlin_dev || !scale_dev || !t_dev) return MSG123_EINVAL;
  hipLaunchKernelGGL(k_target, dim3(nblk(net->d.n)), dim3(256), 0, net->stream, net->d.n,
                     hnext_dev, hlin_dev, scale_dev, len_scal, t_dev);
  NETCHK(hipGetLastError());
  NETCHK(hipStreamSynchronize(net->stream));
  return MSG123_OK;
}
