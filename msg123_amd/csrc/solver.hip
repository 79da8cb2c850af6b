// solver.hip -- MI355X Newton-Krylov step for the 3D mixed-form Richards
// equation: device kernels, the host-side Newton/BiCGStab driver and the
// C ABI declared in include/msg123_gpu.h.
//
// FP64 throughout this file (residual, Jacobian, Krylov recurrences, D-ILU).
// Arithmetic inside each kernel follows the operation order of the
// reference routine it replaces (cited per kernel) and is compiled with
// -ffp-contract=off, so a kernel's result differs from the reference only
// where device libm (pow) and the order of a parallel reduction differ.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hip.h"
#include "amg_gpu.h"
#include "warmnet.h"
#include "msg123_gpu.h"

using namespace msg123;

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "msg123: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                               \
      return MSG123_EDEVICE;                                                          \
    }                                                                                 \
  } while (0)

namespace {

constexpr double MACHI_EPS = 1.1920928955078125e-07;  // epsilon(1._SP)
constexpr double VARMAX = 1.0e3, XMAX = 1.0e4, DINFI = 1.0e20, DSMAL = 1.0e-20;

// ===========================================================================
// Constitutive law: calc_srat_rperm (Calc_parameter.f90:15-77).
// se**0.5 is evaluated as sqrt(se): for se in (0,1] a correctly rounded pow
// and sqrt agree bit for bit.
struct Soil {
  const float *reta, *retn, *resi, *poro, *ss;
};

//
// FAST = true evaluates the same law with five transcendentals instead of
// four pow: with x = |psi alpha|, beta = x^n, Se = (1+beta)^-m,
// Se^(1/m) = 1/(1+beta), so 1 - (1 - Se^(1/m))^m = -expm1(-m log1p(1/beta)),
// and log1p(beta) = n log x + log1p(1/beta).  It also avoids the cancellation
// of 1 - (1 - Se^(1/m))^m in dry cells.  Used by the mixed path only; the
// plain path keeps the reference's form.
// Logarithms and quotients of the log/exp law, without the special-case
// handling and double-double steps of the general library routines (its
// arguments are finite and positive; 1-2 ulp is ample for a law that is
// held to 1e-14 of its closed form).
//   law_log1p_core(f), f in [sqrt(1/2) - 1, sqrt(2) - 1]: log(1 + f) by
//     s = f / (2 + f), log(1 + f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R)),
//     R = z sum_k 2 z^(k-1) / (2k + 1), z = s^2, nine terms (|s| <= 0.172:
//     the first term left out is below 3e-17 relative);
//   law_log(x) = e ln2 + log(1 + f) with x = 2^e (1 + f) reduced to
//     1 + f in [sqrt(1/2), sqrt(2)) (f exact), ln2 split into a 32-bit head
//     and a tail;
//   law_log1p(t), t in (0, 1]: the core directly, or ln2 + log(1 + f) with
//     1 + f = (1 + t) / 2, f = (t - 1) / 2 exact;
//   law_div(a, b): a / b from the hardware reciprocal, two Newton steps and
//     one residual correction.
__device__ __forceinline__ double law_div(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double law_log1p_core(double f) {
  const double s = law_div(f, 2.0 + f), z = s * s;
  double p = 0x1.af286bca1af28p-4;  // 2/19
  p = fma(z, p, 0x1.e1e1e1e1e1e1ep-4);  // 2/17
  p = fma(z, p, 0x1.1111111111111p-3);  // 2/15
  p = fma(z, p, 0x1.3b13b13b13b14p-3);  // 2/13
  p = fma(z, p, 0x1.745d1745d1746p-3);  // 2/11
  p = fma(z, p, 0x1.c71c71c71c71cp-3);  // 2/9
  p = fma(z, p, 0x1.2492492492492p-2);  // 2/7
  p = fma(z, p, 0x1.999999999999ap-2);  // 2/5
  p = fma(z, p, 0x1.5555555555555p-1);  // 2/3
  const double R = z * p, hfsq = 0.5 * f * f;
  return f - (hfsq - s * (hfsq + R));
}
constexpr double kLn2Hi = 0x1.62e42fee00000p-1, kLn2Lo = 0x1.a39ef35793c76p-33;
__device__ __forceinline__ double law_log(double x) {
  int e;
  double m = frexp(x, &e);  // m in [1/2, 1)
  if (m < 0x1.6a09e667f3bcdp-1) {  // sqrt(1/2)
    m = m + m;
    e = e - 1;
  }
  const double de = (double)e;
  return de * kLn2Hi + (law_log1p_core(m - 1.0) + de * kLn2Lo);
}
__device__ __forceinline__ double law_log1p(double t) {
  const bool big = t > 0x1.a827999fcef32p-2;  // sqrt(2) - 1
  const double r = law_log1p_core(big ? (t - 1.0) * 0.5 : t);
  return big ? kLn2Hi + (r + kLn2Lo) : r;
}

template <bool FAST>
__device__ __forceinline__ void vg_point(double per_phead, float reta, float retn, float resi,
                                         float poro, float ssv, double &srat, double &kr,
                                         double &sst) {
  if (FAST && per_phead < 0.0) {
    const double rn = (double)retn;
    const double retm = 1.0 - law_div(1.0, rn);
    const double nlx = rn * law_log(fabs(per_phead * (double)reta));  // log beta
    // log1p(beta) and log1p(1/beta) from one log1p: exp(-|log beta|) is
    // 1/beta where beta > 1 and beta otherwise
    const double lp = law_log1p(exp(-fabs(nlx)));
    const double l1 = nlx > 0.0 ? nlx + lp : lp;
    const double l2 = nlx > 0.0 ? lp : lp - nlx;
    const double se = exp(-retm * l1);
    const double theta = (double)resi + (double)(float)(poro - resi) * se;
    srat = law_div(theta, (double)poro);
    const double c = -expm1(-retm * l2);
    kr = sqrt(se) * (c * c);
    sst = 0.0;
  } else if (per_phead < 0.0) {
    const double rn = (double)retn;
    const double retm = 1.0 - 1.0 / rn;
    double theta, se;
    if (per_phead <= 0.0) {
      const double beta = pow(fabs(per_phead * (double)reta), rn);
      se = pow(1.0 + beta, -retm);
      theta = (double)resi + (double)(float)(poro - resi) * se;
    } else {
      se = 1.0;  // (1+0)**(-m)
      theta = (double)poro;
    }
    srat = theta / (double)poro;
    const double c = 1.0 - pow(1.0 - pow(se, 1.0 / retm), retm);
    kr = sqrt(se) * (c * c);
    sst = 0.0;
  } else {
    srat = 1.0;
    kr = 1.0;
    sst = (double)ssv;
  }
}

// The law at a pressure head and at that head + newper, as make_matrix's two
// calls give them (Make_linearsystem.f90:95-97), for the Jacobian's
// difference quotients (S(h+e) - S(h))/e and (kr(h+e) - kr(h))/e.
// FAST = false (the reference's path) evaluates the law twice.  FAST = true
// (the mixed path, whose Jacobian only steers Newton) takes the second
// point from the law's analytic derivative instead of a second evaluation:
// with g = beta/(1+beta), x = -m n / psi,
//   dSe/dpsi = x Se g,    dc/dpsi = x (1 - c)(1 - g)   (c = 1 - g^m),
//   dkr/dpsi = x sqrt(Se) c (c g / 2 + 2 (1 - c)(1 - g)),
// and S(h+e) = S + e dS, kr(h+e) = kr + e dkr, so the quotients give the
// derivative to the accuracy of a difference quotient with e ~ 1e-8 without
// its second set of five transcendentals.  Ss at h + newper is not needed.
// Within 64 e of saturation the second point is evaluated as before: for
// n < 2, dkr/dpsi grows without bound as psi -> 0 (kr ~ 1 - 2 (alpha|psi|)^(n-1)),
// and there the quotient over [psi, psi + e], bounded, is the better Newton
// slope (cells at the water table sit there).
constexpr double kPairCusp = 64.0;
template <bool FAST>
__device__ __forceinline__ void vg_pair(double per_phead, double newper, float reta, float retn,
                                        float resi, float poro, float ssv, double &srat,
                                        double &kr, double &sst, double &srat_p, double &kr_p) {
  if (!FAST) {
    double q2;
    vg_point<false>(per_phead, reta, retn, resi, poro, ssv, srat, kr, sst);
    vg_point<false>(per_phead + newper, reta, retn, resi, poro, ssv, srat_p, kr_p, q2);
    return;
  }
  if (per_phead < 0.0) {
    const double rn = (double)retn;
    const double retm = 1.0 - law_div(1.0, rn);
    const double nlx = rn * law_log(fabs(per_phead * (double)reta));  // log beta
    // t = exp(-|log beta|): 1/beta where beta > 1, beta otherwise
    const double t = exp(-fabs(nlx));
    const double lp = law_log1p(t);
    const double l1 = nlx > 0.0 ? nlx + lp : lp;  // log1p(beta)
    const double l2 = nlx > 0.0 ? lp : lp - nlx;  // log1p(1/beta)
    const double se = exp(-retm * l1);
    const double dsp = (double)(float)(poro - resi);
    const double theta = (double)resi + dsp * se;
    const double ipor = law_div(1.0, (double)poro);
    srat = law_div(theta, (double)poro);
    const double c = -expm1(-retm * l2);
    const double sq = sqrt(se);
    kr = sq * (c * c);
    sst = 0.0;
    const double inv = law_div(1.0, 1.0 + t);
    const double g = nlx > 0.0 ? inv : t * inv, omg = nlx > 0.0 ? t * inv : inv;
    const double x = law_div(-retm * rn, per_phead);
    const double dse = x * se * g;
    const double dkr = x * sq * c * (0.5 * c * g + 2.0 * (1.0 - c) * omg);
    srat_p = srat + newper * (dsp * dse * ipor);
    kr_p = kr + newper * dkr;
    if (per_phead + kPairCusp * newper >= 0.0) {
      double q2;
      vg_point<true>(per_phead + newper, reta, retn, resi, poro, ssv, srat_p, kr_p, q2);
    }
  } else {
    // saturated at h and at h + newper (newper > 0)
    srat = kr = srat_p = kr_p = 1.0;
    sst = (double)ssv;
  }
}

// h (optionally h_pre + lam*dh) -> S, kr, Ss; with `pair`, also S and kr at
// h + newper (make_matrix's two calls, Make_linearsystem.f90:95-97).
template <bool FAST>
__global__ void __launch_bounds__(kBlock)
k_vg(Dims dm, int nblocks, const double *__restrict__ h_in, const double *__restrict__ h_pre,
     const double *__restrict__ dh, double lam, double *__restrict__ h_out,
     const double *__restrict__ cell_top, Soil soil, double pert0, double newper, int pair,
     double *__restrict__ srat, double *__restrict__ kr, double *__restrict__ sst,
     double *__restrict__ srat_p, double *__restrict__ kr_p) {
  const int i = xcd_block(nblocks) * kBlock + threadIdx.x;
  if (i >= dm.n) return;
  double h;
  if (h_pre) {
    h = h_pre[i] + lam * dh[i];  // calc_funcl2norm (Nonlinear_solution.f90:775-777)
    h_out[i] = h;
  } else {
    h = h_in[i];
  }
  const float a = soil.reta[i], nn = soil.retn[i], rr = soil.resi[i], pp = soil.poro[i],
              sv = soil.ss[i];
  const double top = cell_top[i];
  double s, k, q;
  vg_point<FAST>(h - top + pert0, a, nn, rr, pp, sv, s, k, q);
  srat[i] = s;
  kr[i] = k;
  sst[i] = q;
  if (pair) {
    vg_point<FAST>(h - top + newper, a, nn, rr, pp, sv, s, k, q);
    srat_p[i] = s;
    kr_p[i] = k;
  }
}

// calc_hyd_upwind (Calc_parameter.f90:107-125)
__device__ __forceinline__ double upwind(double head_dif, double k1, double k2) {
  return head_dif >= 0.0 ? k1 : k2;
}

struct Faces {
  const double *hE, *aE, *hS, *aS, *hD, *aD;  // hydf / abyd of the E, S, lower face
  // free-drainage bottom: per (x, y) column the conductance of the bottom
  // face under unit gradient (Ks * area, scaled units), the outflow being
  // drain * kr of the bottom cell; null for the closed bottom
  const double *drain;
};

struct Surface {
  const double *hydf, *abyd, *top, *bott;  // [ns]
  const double *head;                      // surf_head [ns]
  double *rati;                            // surf_rati [ns] (state)
  const double *rech;                      // calc_rech [ns]
  // rivers and lakes per top cell, [6][ns]: tran = hydf_surf * abyd, stage,
  // bottom of the river, then of the lake (tran 0 where there is none); null
  // without either
  const double *rl;
};

// One river or lake entry of a top cell: func_riveterm / func_laketerm
// (Calc_function.f90:393-465) into f, and with JAC set_rivebound /
// set_lakebound (Make_linearsystem.f90:287-408, form_switch == 1) into d:
// -tran*kr on every entry, plus the FD derivative of kr where the exchange
// follows the cell head.  t3 = {tran, stage, bottom} columns of [3][ns].
template <bool JAC>
__device__ __forceinline__ void rl_term(const double *t3, int ns, int col, double h, double kr,
                                        double kp, double newper_inv, double &f, double &d) {
  const double tran = t3[col], head = t3[ns + col], bott = t3[2 * ns + col];
  double delh, over;
  if (h >= bott) {
    delh = head - h;
    over = 1.0;
  } else if (head > bott) {
    delh = head - bott;
    over = 0.0;
  } else {
    delh = 0.0;
    over = 1.0;
  }
  f = tran * delh * kr;
  if (JAC) d = -tran * kr + (kp - kr) * newper_inv * delh * tran * over;
}

// Face geometry, volume and cell top: per layer ([nz], index z) when
// `layered`, else per cell ([N], index i).
struct Geo {
  const double *aE, *aS, *aD;  // abyd of the E, S and lower face
  const double *vol, *top;
  int layered;
  const double *zero;  // 256 bytes of zeros: the address of a masked-off load
};

struct Storage {
  const double *h_old, *srat_old, *vol;
  const float *poro;
  // 1 / delt in device memory (set with delt), so that a captured graph of
  // kernels reading it holds for every step
  const double *delt_inv;
};

// func_surfterm (Calc_function.f90:304-391), no geography file (surf_reli = 0)
__device__ __forceinline__ double surf_term(const Surface &sf, int i, double h, double kr) {
  const double top = sf.top[i], bott = sf.bott[i], sh = sf.head[i];
  double delh, rati = sf.rati[i];
  if (h >= top) {
    delh = (sh > top) ? sh - h : top - h;
    rati = 1.0;
  } else if (h >= bott) {
    delh = (sh > bott) ? sh - h : bott - h;
    rati = 1.0;
  } else if (sh > bott) {
    delh = sh - bott;
    rati = 1.0;
  } else {
    delh = 0.0;
  }
  sf.rati[i] = rati;
  return sf.hydf[i] * sf.abyd[i] * delh * kr * rati;
}

// one connection of func_connflow (Calc_function.f90:212-254)
__device__ __forceinline__ void flow(double hj, double kj, double hy, double ab, double hi,
                                     double ki, double &conn) {
  const double dhd = hj - hi;
  const double rel = upwind(-dhd, ki, kj);
  conn = conn + hy * rel * ab * dhd;
}

// Mass balance of an accepted step: calc_mass (Calc_function.f90:103-174)
// summed over the domain as calc_out_mass does for the whole-domain output
// (Calc_output.f90:54-166), times dt, and beside it the Newton residual of
// calc_func at the accepted head.  Per block: storage change (Ss switch of
// the new and the old head, as calc_mass passes them), connection flow,
// recharge, surface exchange (with surf_old, as calc_mass passes it), the
// residual F (current-iterate Ss switch for both levels and surf_head, as
// calc_func), and |recharge|.  Reads only; no state is written.
// With ss_new / ss_old null, the storage coefficient of each head is
// evaluated here from cell_top and the Ss parameter: Ss where the pressure
// head h - top is not negative, 0 below (calc_srat_rperm's third output,
// Calc_parameter.f90:15-77), the same values the constitutive-law kernel
// writes.
__global__ void __launch_bounds__(kBlock)
k_mass(Dims dm, int nblocks, const double *__restrict__ h, const double *__restrict__ srat,
       const double *__restrict__ kr, const double *__restrict__ ss_new,
       const double *__restrict__ ss_old, Geo geo, const float *__restrict__ ss_param,
       Faces fc, Storage st, Surface sf, const double *__restrict__ surf_old, double dt,
       double *__restrict__ partials) {
  const int lb = xcd_block(nblocks);
  const int i = lb * kBlock + threadIdx.x;
  double red[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < dm.n) {
    const int x = i % dm.nx, y = (i / dm.nx) % dm.ny, z = i / dm.nxy;
    // face geometry, volume and top per layer on a layered grid (the same
    // values as the per-cell arrays, bit for bit), else per cell
    const bool lay = geo.layered != 0;
    const int gi = lay ? z : i;
    const double hi = h[i], ki = kr[i], si = __builtin_nontemporal_load(srat + i);
    const double ho = __builtin_nontemporal_load(st.h_old + i);
    const double so = __builtin_nontemporal_load(st.srat_old + i);
    double ssn, sso;
    if (ss_new) {
      ssn = ss_new[i];
      sso = ss_old[i];
    } else {
      const double top = geo.top[gi];
      const double ssv = (double)__builtin_nontemporal_load(ss_param + i);
      ssn = (hi - top + 0.0 < 0.0) ? 0.0 : ssv;
      sso = (ho - top + 0.0 < 0.0) ? 0.0 : ssv;
    }
    const double dsat = (double)__builtin_nontemporal_load(st.poro + i) * (si - so);
    const double delt_inv = *st.delt_inv;
    const double vol = geo.vol[gi];
    const double stom = -(((ssn * si * hi) - (sso * so * ho)) + dsat) * delt_inv * vol;
    const double stof = -(((ssn * si * hi) - (ssn * so * ho)) + dsat) * delt_inv * vol;
    double conn = 0.0;
#define MSG123_FLOW(J, HY, AB)                                    \
  {                                                               \
    const int j_ = (J);                                           \
    const double dhd = h[j_] - hi;                                \
    const double rel = upwind(-dhd, ki, kr[j_]);                  \
    conn = conn + (HY) * rel * (AB) * dhd;                        \
  }
    if (z > 0)
      MSG123_FLOW(i - dm.nxy, fc.hD[i - dm.nxy], geo.aD[lay ? z - 1 : i - dm.nxy]);
    if (y > 0) MSG123_FLOW(i - dm.nx, fc.hS[i - dm.nx], geo.aS[lay ? z : i - dm.nx]);
    if (x > 0) MSG123_FLOW(i - 1, fc.hE[i - 1], geo.aE[lay ? z : i - 1]);
    if (x < dm.nx - 1) MSG123_FLOW(i + 1, fc.hE[i], geo.aE[gi]);
    if (y < dm.ny - 1) MSG123_FLOW(i + dm.nx, fc.hS[i], geo.aS[gi]);
    if (z < dm.nz - 1) {
      MSG123_FLOW(i + dm.nxy, fc.hD[i], geo.aD[gi]);
    } else if (fc.drain) {
      conn = conn - fc.drain[i - z * dm.nxy] * ki;
    }
#undef MSG123_FLOW
    double rec = 0.0, surm = 0.0, surf = 0.0;
    if (z == 0) {
      rec = sf.rech[i];
      // func_surfterm without its surf_rati update, for two surface heads
      const double top = sf.top[i], bott = sf.bott[i], tran = sf.hydf[i] * sf.abyd[i];
      const double shs[2] = {surf_old[i], sf.head[i]};
      double out[2];
      for (int q = 0; q < 2; q++) {
        const double sh = shs[q];
        double delh, rati = sf.rati[i];
        if (hi >= top) {
          delh = (sh > top) ? sh - hi : top - hi;
          rati = 1.0;
        } else if (hi >= bott) {
          delh = (sh > bott) ? sh - hi : bott - hi;
          rati = 1.0;
        } else if (sh > bott) {
          delh = sh - bott;
          rati = 1.0;
        } else {
          delh = 0.0;
        }
        out[q] = tran * delh * ki * rati;
      }
      surm = out[0];
      surf = out[1];
    }
    const double fv = (z == 0) ? ((rec + surf) + stof + conn) : (stof + conn);
    red[0] = stom * dt;
    red[1] = conn * dt;
    red[2] = rec * dt;
    red[3] = surm * dt;
    red[4] = fv * dt;
    red[5] = fabs(rec) * dt;
  }
  block_sums<6>(red, partials, lb, nblocks);
}

// one connection of form_connflow (Make_linearsystem.f90:195-285) with its
// flow: adds to the flow sum and the diagonal, returns the off-diagonal
__device__ __forceinline__ double conn_jac(double hj, double kj, double kpj, double hy, double ab,
                                          double hi, double ki, double kpi, double newper,
                                          double newper_inv, double &conn, double &cond) {
  const double dhd = hj - hi;
  const double rel = upwind(-dhd, ki, kj);
  conn = conn + hy * rel * ab * dhd;
  const double rel_hyd = hy * rel;
  const double deri_dcon = -rel_hyd * ab;
  const double deri_lucon = rel_hyd * ab;
  const double dh1 = hy * upwind(-dhd + newper, kpi, kj);
  const double dh2 = hy * upwind(-dhd - newper, ki, kpj);
  const double con1 = (dh1 - rel_hyd) * newper_inv * dhd * ab;
  const double con2 = (dh2 - rel_hyd) * newper_inv * dhd * ab;
  cond = cond + deri_dcon + con1;
  return deri_lucon + con2;
}

// ---------------------------------------------------------------------------
// Fused stencil kernels: the constitutive law evaluated inside the residual
// and the Jacobian assembly.
//
// A block owns a TX x TY tile of (x, y) columns and a chunk of z levels
// [z0, z1) and marches down it.  At every level it evaluates the
// van Genuchten-Mualem law once per cell for the tile and the one-cell ring
// around it (the ring's x/y neighbours of the tile's edge cells), keeping
// h, kr (and kr at h + newper for the Jacobian) in LDS, and one level ahead:
// while level z is assembled, level z+1 (the lower neighbour) is already in
// LDS.  Three LDS buffers rotate, so one barrier per level separates the
// writes of level z+1 from the reads of levels z and z+1.  The upper
// neighbour's values come from the previous level in registers.  So S, kr
// and Ss never go through HBM: per cell the kernel reads h, h_old, S_old,
// the five float32 soil parameters and the three FP64 face conductances of
// the cell's E, S and lower faces, and writes its outputs.  Face geometry
// (A/d), cell volumes and cell tops are read per layer when they are uniform
// over each layer of the box (the usual layered grid; checked when the
// context is created), else per cell.
//
// Per-cell arithmetic is that of the reference routines, in their order:
// calc_srat_rperm (vg_point), func_stochn, func_connflow (flow), func_rechterm,
// func_surfterm (surf_term); form_stochn, form_connflow (conn_jac),
// set_surfbound for the Jacobian.
struct Tiling {
  int ntx, nty;   // tiles in x and y
  int zlen;       // z levels per chunk
  int nblocks;    // ntx * nty * chunks
};

constexpr int kFTX = 16, kFTY = 16;  // tile of the fused kernels
constexpr int kFusedBlocks = 1024;    // blocks the z chunking aims at (default)
constexpr int kFusedMaxBlocks = 8192; // largest target the tuner tries
// Launch variants of the fused kernels: the occupancy the compiler is held
// to (waves per SIMD; 1 = unconstrained register use) and the number of
// blocks the z chunking aims at (more blocks: shorter chunks, more law
// evaluations at chunk ends, more parallel slack).
constexpr int kMinWaves[] = {1, 4, 5};
constexpr int kNumMinWaves = 3;
struct FusedShape {
  int minw_idx = 0;
  int target = kFusedBlocks;
};

// z levels of one chunk at most (the layered geometry of a chunk is held
// in LDS: 40 * (zlen + 2) bytes)
constexpr int kFusedMaxZlen = 256;

template <int TX, int TY>
inline Tiling tiling_shape(const Dims &dm, int target_blocks) {
  Tiling t;
  t.ntx = (dm.nx + TX - 1) / TX;
  t.nty = (dm.ny + TY - 1) / TY;
  const int tiles = t.ntx * t.nty;
  int chunks = (target_blocks + tiles - 1) / tiles;
  const int min_chunks = (dm.nz + kFusedMaxZlen - 1) / kFusedMaxZlen;
  chunks = chunks < min_chunks ? min_chunks : chunks;
  chunks = chunks < 1 ? 1 : (chunks > dm.nz ? dm.nz : chunks);
  t.zlen = (dm.nz + chunks - 1) / chunks;
  t.nblocks = tiles * ((dm.nz + t.zlen - 1) / t.zlen);
  return t;
}

// dynamic LDS of a fused launch on a layered grid
inline size_t fused_geo_lds(const Tiling &t) { return sizeof(double) * 5 * (size_t)(t.zlen + 2); }

// The head of the step: h[i], or h_pre[i] + lam*dh[i] (calc_funcl2norm,
// Nonlinear_solution.f90:775-777) written to h_out for the block's own cells.
struct HeadIn {
  const double *h, *h_pre, *dh;
  double lam;
  double *h_out;
  __device__ __forceinline__ double at(int i) const { return h ? h[i] : h_pre[i] + lam * dh[i]; }
};

enum FusedMode { kResid = 0, kJv = 1, kAssem = 2 };

struct FusedOut {
  double *f;            // F (kResid), J.v (kJv) or rhs = -F (kAssem)
  double *A;            // kAssem: the seven diagonals
  double *srat;         // S at the head (srat_new)
  const double *f1;     // kJv: F at head_pre
  double eps_inv;       // kJv
  const double *rhs;    // kJv: rhs . jv is reduced
  double *partials;     // per-block reduction (|F|^2, rhs.jv or |rhs|^2)
  const int *run_if;    // device-driven Newton: run only where *run_if != 0
  const double *eps_inv_dev;  // kJv: eps_inv read from device memory
  float *Af;            // kAssem: the seven diagonals in FP32 instead of A
};

// Inputs of the law at one cell, loaded ahead of their use: the head (h,
// or h_pre in h and dh in d, combined when the law is evaluated, so the
// loads are not waited for when they are issued).
struct LawIn {
  double h, d, top;
  float a, n, r, p, s;
};

// The law's outputs at one cell: S, kr, Ss at h; S and kr at h + newper.
struct LawOut {
  double h, s, k, q, sp, kp;
};

// What a cell keeps in registers of its own law outputs from one level to
// the next: S at h and at h + newper, and Ss (0 or the float32 input).  Its
// h, kr and kr(h + newper), which the neighbours read too, stay in the
// LDS ring (levels z-1, z and z+1 are all there when level z is computed).
struct LawOwn {
  double s, sp;
  float q;
};

// Per-cell inputs of the residual / Jacobian besides the law's, loaded one
// level ahead of their use.
struct CellIn {
  double so, ho, hyd, hyE, hyW, hyS, hyN, f1, rhs;
  float poro;
};

template <int MODE, bool FAST, bool LAY, int TX, int TY, int MINW>
__global__ void __launch_bounds__(TX * TY + 2 * TX + 2 * TY, MINW)
k_fused(Dims dm, Tiling tg, HeadIn hin, Soil soil, Geo geo, Faces fc, Storage st, Surface sf,
        double newper, double newper_inv, int consistent, FusedOut out) {
  if (out.run_if && *out.run_if == 0) return;
  const double eps_inv = MODE == kJv ? (out.eps_inv_dev ? *out.eps_inv_dev : out.eps_inv) : 0.0;
  const double delt_inv = *st.delt_inv;
  // threads [0, TX*TY) own the tile's cells; threads [TX*TY, TX*TY + RING)
  // evaluate the law on the ring around the tile, so every wave evaluates it
  // once per level.  Per level: the law of level z+1 into LDS, one barrier,
  // the cells of level z; three LDS buffers (level mod 3) keep level z+1's
  // writes apart from the two levels the other waves may still read.
  constexpr int NOWN = TX * TY, PX = TX + 2, PY = TY + 2, RING = 2 * TX + 2 * TY;
  constexpr int NT = NOWN + RING, NBUF = 3;
  constexpr bool PAIR = MODE == kAssem;
  __shared__ double sh[NBUF][PY * PX], sk[NBUF][PY * PX];
  __shared__ double skp[PAIR ? NBUF : 1][PAIR ? PY * PX : 1];
  const int L = xcd_block(tg.nblocks);
  const int tiles = tg.ntx * tg.nty;
  const int chunk = L / tiles, trem = L - chunk * tiles;
  const int ty = trem / tg.ntx, tx = trem - ty * tg.ntx;
  const int z0 = chunk * tg.zlen, z1 = min(dm.nz, z0 + tg.zlen);
  const int nxy = dm.nxy;
  const int t = threadIdx.x;
  const bool own = t < NOWN;
  // the cell this thread evaluates the law on: its own, or a ring cell
  int vx, vy, vslot;
  if (own) {
    const int lx = t % TX, ly = t / TX;
    vx = tx * TX + lx;
    vy = ty * TY + ly;
    vslot = (ly + 1) * PX + (lx + 1);
  } else {
    const int r = t - NOWN;
    if (r < TX) { vx = tx * TX + r; vy = ty * TY - 1; vslot = r + 1; }
    else if (r < 2 * TX) { vx = tx * TX + r - TX; vy = ty * TY + TY; vslot = (PY - 1) * PX + (r - TX + 1); }
    else if (r < 2 * TX + TY) { vx = tx * TX - 1; vy = ty * TY + r - 2 * TX; vslot = (r - 2 * TX + 1) * PX; }
    else { vx = tx * TX + TX; vy = ty * TY + r - 2 * TX - TY; vslot = (r - 2 * TX - TY + 1) * PX + PX - 1; }
  }
  const bool valid = vx >= 0 && vx < dm.nx && vy >= 0 && vy < dm.ny;
  const bool active = own && valid;  // owns a cell of the box
  const int x = vx, y = vy, col = vy * dm.nx + vx, slot = vslot;
  // layered grids: the geometry of levels z0-1 .. z1 in LDS, as
  // [aE, aS, aD, vol, top][z - z0 + 1] (dynamic shared memory of 5 * nlev
  // doubles), so no global load of the z loop waits on another's issue order
  extern __shared__ double glv[];
  const int nlev = tg.zlen + 2;
  if constexpr (LAY) {
    for (int k = t; k < nlev; k += NT) {
      const int zz = z0 - 1 + k;
      if (zz >= 0 && zz < dm.nz) {
        glv[k] = geo.aE[zz];
        glv[nlev + k] = geo.aS[zz];
        glv[2 * nlev + k] = geo.aD[zz];
        glv[3 * nlev + k] = geo.vol[zz];
        glv[4 * nlev + k] = geo.top[zz];
      }
    }
    __syncthreads();
  }
  auto lv = [&](int q, int z) { return glv[q * nlev + (z - z0 + 1)]; };
  // Loads every lane issues: a[i] where ok, else a zero read from one shared
  // line.  With no branch around them the number of loads in flight is the
  // same on every path, so each wait is for exactly the load it needs.
  auto ld = [&](const double *a, int i, bool ok) { return *(ok ? a + i : geo.zero); };
  auto ldf = [&](const float *a, int i, bool ok) {
    return *(ok ? a + i : reinterpret_cast<const float *>(geo.zero));
  };
  // the same for inputs no other cell reads (streamed past the caches)
  auto ldn = [&](const double *a, int i, bool ok) {
    return __builtin_nontemporal_load(ok ? a + i : geo.zero);
  };
  auto ldnf = [&](const float *a, int i, bool ok) {
    return __builtin_nontemporal_load(ok ? a + i : reinterpret_cast<const float *>(geo.zero));
  };
  // law inputs of level z (z beyond the levels the chunk needs: zeros)
  auto load = [&](int z) {
    LawIn v{0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.f, 0.f};
    const bool ok = valid && z < dm.nz && z <= z1;
    const int i = z * nxy + col;
    if (hin.h) {
      v.h = ld(hin.h, i, ok);
    } else {
      v.h = ld(hin.h_pre, i, ok);
      v.d = ld(hin.dh, i, ok);
    }
    v.top = LAY ? lv(4, min(z, z1)) : ld(geo.top, i, ok);
    v.a = ldf(soil.reta, i, ok);
    v.n = ldf(soil.retn, i, ok);
    v.r = ldf(soil.resi, i, ok);
    v.p = ldf(soil.poro, i, ok);
    v.s = ldf(soil.ss, i, ok);
    return v;
  };
  // the law at level z into LDS (and the returned registers)
  auto level = [&](int z, const LawIn &v) {
    LawOwn o{0.0, 0.0, 0.f};
    if (valid && z < dm.nz && z <= z1) {
      // h, or calc_funcl2norm's h_pre + lam*dh (Nonlinear_solution.f90:775-777)
      const double h = hin.h ? v.h : v.h + hin.lam * v.d;
      double k, q, kp = 0.0;
      if (PAIR)
        vg_pair<FAST>(h - v.top + 0.0, newper, v.a, v.n, v.r, v.p, v.s, o.s, k, q, o.sp, kp);
      else
        vg_point<FAST>(h - v.top + 0.0, v.a, v.n, v.r, v.p, v.s, o.s, k, q);
      o.q = (float)q;  // 0 or the float32 Ss: exact
      const int b = z % NBUF;
      sh[b][slot] = h;
      sk[b][slot] = k;
      if (PAIR) skp[b][slot] = kp;
      if (own && hin.h_out && z < z1) __builtin_nontemporal_store(h, hin.h_out + z * nxy + col);
    }
    return o;
  };
  double red = 0.0;
  // upper neighbour (level z0 - 1, own column only) and the face above
  // upper neighbour of the chunk's first level (level z0 - 1, own column
  // only): its h, kr and kr(h + newper) into the LDS ring's slot of that
  // level, and the face above
  double hyu = 0.0, abu = 0.0;
  if (active && z0 > 0) {
    const int i = (z0 - 1) * nxy + col;
    const double h = hin.at(i);
    const double top = LAY ? lv(4, z0 - 1) : geo.top[i];
    const float a = soil.reta[i], nn = soil.retn[i], rr = soil.resi[i], pp = soil.poro[i],
                sv = soil.ss[i];
    double s0, k, q, kp = 0.0, s2;
    if (PAIR)
      vg_pair<FAST>(h - top + 0.0, newper, a, nn, rr, pp, sv, s0, k, q, s2, kp);
    else
      vg_point<FAST>(h - top + 0.0, a, nn, rr, pp, sv, s0, k, q);
    const int b = (z0 - 1) % NBUF;
    sh[b][slot] = h;
    sk[b][slot] = k;
    if (PAIR) skp[b][slot] = kp;
    hyu = fc.hD[i];
    abu = LAY ? lv(2, z0 - 1) : geo.aD[i];
  }
  auto load_cell = [&](int z) {
    CellIn v{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.f};
    const bool ok = active && z < z1;
    const int i = z * nxy + col;
    v.so = ldn(st.srat_old, i, ok);
    v.ho = ldn(st.h_old, i, ok);
    v.poro = ldnf(st.poro, i, ok);
    v.hyd = ldn(fc.hD, i, ok && z < dm.nz - 1);
    v.hyE = ld(fc.hE, i, ok && x < dm.nx - 1);
    v.hyW = ld(fc.hE, i - 1, ok && x > 0);
    v.hyS = ld(fc.hS, i, ok && y < dm.ny - 1);
    v.hyN = ld(fc.hS, i - dm.nx, ok && y > 0);
    if (MODE == kJv) {
      v.f1 = ldn(out.f1, i, ok);
      v.rhs = ldn(out.rhs, i, ok);
    }
    return v;
  };
  // one cell of level z: own values c, upper u (face hyu/abu), lower d
  auto cell = [&](int z, const LawOwn &c, const CellIn &ci, double &hyd_out, double &abd_out) {
    const int i = z * nxy + col;
    const int bc = z % NBUF, bu = (z + NBUF - 1) % NBUF, bd = (z + 1) % NBUF;
    // own column at levels z-1, z, z+1 from the LDS ring
    struct Nb { double h, k, kp; };
    const Nb u{sh[bu][slot], sk[bu][slot], PAIR ? skp[bu][slot] : 0.0};
    const Nb d{sh[bd][slot], sk[bd][slot], PAIR ? skp[bd][slot] : 0.0};
    double abd = 0.0;
    const double hyd = ci.hyd;
    if (z < dm.nz - 1) abd = LAY ? lv(2, z) : geo.aD[i];
    hyd_out = hyd;
    abd_out = abd;
    const double hyE = ci.hyE, hyW = ci.hyW, hyS = ci.hyS, hyN = ci.hyN;
    const double aE = LAY ? lv(0, z) : geo.aE[i];
    const double aS = LAY ? lv(1, z) : geo.aS[i];
    const double aW = LAY ? lv(0, z) : (x > 0 ? geo.aE[i - 1] : 0.0);
    const double aN = LAY ? lv(1, z) : (y > 0 ? geo.aS[i - dm.nx] : 0.0);
    const double vol = LAY ? lv(3, z) : geo.vol[i];
    const double *H = sh[bc], *K = sk[bc];
    const double hi = sh[bc][slot], ki = sk[bc][slot], kpi = PAIR ? skp[bc][slot] : 0.0;
    const double si = c.s, spi = c.sp, qi = (double)c.q;
    // storage change (mixed form, func_stochn with the current-iterate Ss
    // switch for both time levels, Calc_function.f90:53)
    const double so = ci.so, ho = ci.ho;
    const double poro = (double)ci.poro;
    const double ch1 = qi * so * ho;
    const double ch2 = qi * si * hi;
    const double ch = (ch2 - ch1) + poro * (si - so);
    const double stof = -ch * delt_inv * vol;
    double conn = 0.0, fv;
    if (MODE != kAssem) {
      // connection flows in CSR order: up, N, W, E, S, down
      if (z > 0) flow(u.h, u.k, hyu, abu, hi, ki, conn);
      if (y > 0) flow(H[slot - PX], K[slot - PX], hyN, aN, hi, ki, conn);
      if (x > 0) flow(H[slot - 1], K[slot - 1], hyW, aW, hi, ki, conn);
      if (x < dm.nx - 1) flow(H[slot + 1], K[slot + 1], hyE, aE, hi, ki, conn);
      if (y < dm.ny - 1) flow(H[slot + PX], K[slot + PX], hyS, aS, hi, ki, conn);
      if (z < dm.nz - 1) flow(d.h, d.k, hyd, abd, hi, ki, conn);
      else if (fc.drain) conn = conn - fc.drain[col] * ki;
      if (z == 0) {
        double funcvs = sf.rech[col] + surf_term(sf, col, hi, ki);
        if (sf.rl) {
          double fr, fl, dd;
          rl_term<false>(sf.rl, nxy, col, hi, ki, 0.0, 0.0, fr, dd);
          rl_term<false>(sf.rl + 3 * (size_t)nxy, nxy, col, hi, ki, 0.0, 0.0, fl, dd);
          funcvs = (funcvs + fr) + fl;  // recf + surf + rivf + lakf
        }
        fv = funcvs + stof + conn;
      } else {
        fv = stof + conn;
      }
      if (MODE == kJv) {
        const double jv = -(fv - ci.f1) * eps_inv;
        __builtin_nontemporal_store(jv, out.f + i);
        red += ci.rhs * jv;
      } else {
        __builtin_nontemporal_store(fv, out.f + i);
        red += fv * fv;
      }
    } else {
      const double *KP = skp[bc];
      // storage derivative (form_stochn, form_switch == 1)
      const double deri_srat = (spi - si) * newper_inv;
      const double deri_stor = qi * deri_srat * hi;
      const double stod = -(poro * deri_srat + qi * si + deri_stor) * delt_inv * vol;
      double cond = 0.0;
      double lu[6] = {0, 0, 0, 0, 0, 0};
      if (z > 0) lu[0] = conn_jac(u.h, u.k, u.kp, hyu, abu, hi, ki, kpi, newper, newper_inv, conn, cond);
      if (y > 0)
        lu[1] = conn_jac(H[slot - PX], K[slot - PX], KP[slot - PX], hyN, aN, hi, ki, kpi, newper,
                         newper_inv, conn, cond);
      if (x > 0)
        lu[2] = conn_jac(H[slot - 1], K[slot - 1], KP[slot - 1], hyW, aW, hi, ki, kpi, newper,
                         newper_inv, conn, cond);
      if (x < dm.nx - 1)
        lu[3] = conn_jac(H[slot + 1], K[slot + 1], KP[slot + 1], hyE, aE, hi, ki, kpi, newper,
                         newper_inv, conn, cond);
      if (y < dm.ny - 1)
        lu[4] = conn_jac(H[slot + PX], K[slot + PX], KP[slot + PX], hyS, aS, hi, ki, kpi, newper,
                         newper_inv, conn, cond);
      if (z < dm.nz - 1) {
        lu[5] = conn_jac(d.h, d.k, d.kp, hyd, abd, hi, ki, kpi, newper, newper_inv, conn, cond);
      } else if (fc.drain) {
        // outflow drain * kr and its FD derivative, as deri_ks on the top face
        const double dr = fc.drain[col];
        conn = conn - dr * ki;
        cond = cond - dr * (kpi - ki) * newper_inv;
      }
      double dmat;
      if (z == 0) {
        // func_surfterm for F, set_surfbound for the diagonal
        const double top = sf.top[col], bott = sf.bott[col], shd = sf.head[col];
        double delh, over, rati = sf.rati[col];
        if (hi >= top) {
          delh = (shd > top) ? shd - hi : top - hi;
          over = 1.0;
          rati = 1.0;
        } else if (hi >= bott) {
          delh = (shd > bott) ? shd - hi : bott - hi;
          over = 1.0;
          rati = 1.0;
        } else if (shd > bott) {
          delh = shd - bott;
          over = 0.0;
          rati = 1.0;
        } else {
          delh = 0.0;
          over = 1.0;
        }
        sf.rati[col] = rati;
        const double tran = sf.hydf[col] * sf.abyd[col];
        const double surf = tran * delh * ki * rati;
        const double deri_s = (consistent && hi < bott) ? 0.0 : -tran * ki;
        const double deri_ks = (kpi - ki) * newper_inv * delh * tran * over;
        double funcvs = sf.rech[col] + surf, dmats = deri_s + deri_ks;
        if (sf.rl) {
          double fr, dr, fl, dl;
          rl_term<true>(sf.rl, nxy, col, hi, ki, kpi, newper_inv, fr, dr);
          rl_term<true>(sf.rl + 3 * (size_t)nxy, nxy, col, hi, ki, kpi, newper_inv, fl, dl);
          funcvs = (funcvs + fr) + fl;  // recf + surf + rivf + lakf
          dmats = (dmats + dr) + dl;    // surd + rivd + lakd
        }
        fv = funcvs + stof + conn;
        dmat = dmats + stod + cond;
      } else {
        fv = stof + conn;
        dmat = stod + cond;
      }
      const double r = -fv;
      __builtin_nontemporal_store(r, out.f + i);
      red += r * r;
      const size_t N = (size_t)dm.n;
      if (out.Af) {
        __builtin_nontemporal_store((float)dmat, out.Af + i);
#pragma unroll
        for (int q = 0; q < 6; q++)
          __builtin_nontemporal_store((float)lu[q], out.Af + (size_t)(q + 1) * N + i);
      } else {
        __builtin_nontemporal_store(dmat, out.A + i);
#pragma unroll
        for (int q = 0; q < 6; q++)
          __builtin_nontemporal_store(lu[q], out.A + (size_t)(q + 1) * N + i);
      }
    }
    __builtin_nontemporal_store(si, out.srat + i);
  };
  // level z0's law now; the law inputs of level z0 + 1 and the cell inputs
  // of level z0 in flight.  Each load is issued into the registers its
  // previous contents just left (in after level(), ci after cell()), so
  // the loop carries no copy of a load still in flight: the law inputs load
  // while the cells of a level are computed, the cell inputs while the law
  // of the next level is.
  LawIn in = load(z0 + 1);
  CellIn ci = load_cell(z0);
  LawOwn cur = level(z0, load(z0));
  for (int z = z0; z < z1; z++) {
    const LawOwn n1 = level(z + 1, in);
    in = load(z + 2);
    __syncthreads();
    if (active) {
      double hyd, abd;
      cell(z, cur, ci, hyd, abd);
      hyu = hyd;
      abu = abd;
    }
    ci = load_cell(z + 1);
    cur = n1;
  }
  if (out.partials) {
    double v[1] = {red};
    block_sums<1, NT>(v, out.partials, L, tg.nblocks);
  }
}

// ---------------------------------------------------------------------------
// Well and sea cells (msg123_set_bterms): a pass over their list after the
// fused kernels.  calc_func adds them last, funcv = funcvs + stof + conf +
// welf + seaf (Calc_function.f90:86-91), and make_matrix adds sead last
// (Make_linearsystem.f90:136-141), so adding them to the fused kernels'
// outputs keeps the reference's order.  Per cell: welf = calc_well
// (func_wellterm, :277-302); per sea connection in order (func_sealterm,
// :467-513; set_seabound, Make_linearsystem.f90:499-547) the flow
// hydf kr abyd (sea level - h) and the diagonal -kr hydf abyd + FD term.
struct BCells {
  int n;                         // listed cells, ascending
  const int *cell;               // [n]
  const double *well;            // [n] calc_well (0 for a sea-only cell)
  const int *beg;                // [n + 1] sea connections of cell e: beg[e] .. beg[e+1]
  const double *hydf, *abyd, *lvl;  // per sea connection (lvl: read_seal as FP64)
};

constexpr int kBcT = 256;

// MODE kResid: out holds F; kAssem: out holds rhs = -F and A the diagonals.
// The norm the fused kernel left in partials[0] (|F|^2 or |rhs|^2) is
// corrected by the changed cells' squares, summed in a fixed order.
template <int MODE, bool FAST>
__global__ void __launch_bounds__(kBcT)
k_bcells(BCells bc, HeadIn hin, const double *__restrict__ cell_top, Soil soil, double newper,
         double newper_inv, double *__restrict__ out, double *__restrict__ Ad,
         float *__restrict__ Adf, double *__restrict__ partials, const int *__restrict__ run_if) {
  if (run_if && *run_if == 0) return;
  double corr = 0.0;
  for (int e = threadIdx.x; e < bc.n; e += kBcT) {
    const int c = bc.cell[e];
    const double h = hin.at(c);
    double s, kr, q, kp = 0.0, s2;
    if (MODE == kAssem)
      vg_pair<FAST>(h - cell_top[c] + 0.0, newper, soil.reta[c], soil.retn[c], soil.resi[c],
                    soil.poro[c], soil.ss[c], s, kr, q, s2, kp);
    else
      vg_point<FAST>(h - cell_top[c] + 0.0, soil.reta[c], soil.retn[c], soil.resi[c],
                     soil.poro[c], soil.ss[c], s, kr, q);
    double seaf = 0.0, sead = 0.0;
    for (int k = bc.beg[e]; k < bc.beg[e + 1]; k++) {
      const double delh = bc.lvl[k] - h;
      seaf = seaf + bc.hydf[k] * kr * bc.abyd[k] * delh;
      if (MODE == kAssem) {
        const double per = kp * bc.hydf[k], rel = kr * bc.hydf[k];
        const double deri = -rel * bc.abyd[k];
        const double deri_ks = (per - rel) * newper_inv * delh * bc.abyd[k];
        sead = sead + deri + deri_ks;
      }
    }
    const double fo = MODE == kAssem ? -out[c] : out[c];
    const double fn = (fo + bc.well[e]) + seaf;
    out[c] = MODE == kAssem ? -fn : fn;
    corr += fn * fn - fo * fo;
    if (MODE == kAssem) {
      if (Adf) Adf[c] = (float)((double)Adf[c] + sead);
      else Ad[c] = Ad[c] + sead;
    }
  }
  if (partials) {
    // one block: its sum lands in one[0], written and read by thread 0
    double v[1] = {corr};
    __shared__ double one[1];
    block_sums<1, kBcT>(v, one, 0, 1);
    if (threadIdx.x == 0) partials[0] = partials[0] + one[0];
  }
}

// J.v from F at the perturbed head (the fused residual's output with the
// well and sea cells added): jv = -(F - f1) eps_inv, rhs.jv per block into
// partials slot blockIdx.x of gridDim.x (the residual kernel's slot count)
__global__ void __launch_bounds__(kBlock)
k_jv_from_f(int n, const double *__restrict__ F, const double *__restrict__ f1, double eps_inv,
            const double *__restrict__ eps_inv_dev, const double *__restrict__ rhs,
            double *__restrict__ jv, double *__restrict__ partials,
            const int *__restrict__ run_if) {
  if (run_if && *run_if == 0) return;
  const double ei = eps_inv_dev ? *eps_inv_dev : eps_inv;
  double red = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double v = -(F[i] - f1[i]) * ei;
    jv[i] = v;
    red += rhs[i] * v;
  }
  if (partials) {
    double v[1] = {red};
    block_sums<1, kBlock>(v, partials, blockIdx.x, gridDim.x);
  }
}

// The boundary terms of calc_mass (Calc_function.f90:103-174) at head h with
// its kr: per cell (times dt) the well, river, lake and sea volumes, summed
// in a fixed order into res[0..3]; res[4] = their sum over all cells (the
// part of the residual's volume the fused-free k_mass does not hold).
// Rivers and lakes over the listed columns rl_cols.
__global__ void __launch_bounds__(kBcT)
k_bmass(int nxy, BCells bc, const double *__restrict__ rl, int ncols,
        const int *__restrict__ rl_cols, const double *__restrict__ h,
        const double *__restrict__ kr, double dt, double *__restrict__ res) {
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < bc.n; e += kBcT) {
    const int c = bc.cell[e];
    double seaf = 0.0;
    for (int k = bc.beg[e]; k < bc.beg[e + 1]; k++)
      seaf = seaf + bc.hydf[k] * kr[c] * bc.abyd[k] * (bc.lvl[k] - h[c]);
    v[0] += bc.well[e] * dt;
    v[3] += seaf * dt;
    v[4] += (bc.well[e] + seaf) * dt;
  }
  for (int e = threadIdx.x; e < ncols; e += kBcT) {
    const int col = rl_cols[e];
    double fr, fl, dd;
    rl_term<false>(rl, nxy, col, h[col], kr[col], 0.0, 0.0, fr, dd);
    rl_term<false>(rl + 3 * (size_t)nxy, nxy, col, h[col], kr[col], 0.0, 0.0, fl, dd);
    v[1] += fr * dt;
    v[2] += fl * dt;
    v[4] += (fr + fl) * dt;
  }
  __shared__ double sums[5];
  block_sums<5, kBcT>(v, sums, 0, 1);
  __syncthreads();
  if (threadIdx.x < 5) res[threadIdx.x] = sums[threadIdx.x];
}

// ===========================================================================
// Krylov building blocks (Linear_solution.f90:234-493, 917-951;
// Calc_simulation.f90:36-73).


// r = b - A x, rs = r, partial sk = r.r (calc_resi + the first dot)
template <typename T = double>
__global__ void __launch_bounds__(kBlock)
k_resid(Dims dm, int nblocks, Mat7T<T> A, const double *__restrict__ x, const double *__restrict__ b,
        double *__restrict__ r, double *__restrict__ rs, double *__restrict__ partials) {
  const int lb = xcd_block(nblocks);
  const int i = lb * kBlock + threadIdx.x;
  double red[1] = {0.0};
  if (i < dm.n) {
    const double ri = b[i] - apply7(A, dm, x, i);
    r[i] = ri;
    rs[i] = ri;
    red[0] = ri * ri;
  }
  block_sums<1>(red, partials, lb, nblocks);
}

// v = A z with partial v.rs (mode 0) or t = A z with partials t.r, t.t (mode 1)
// skip (device-driven solves): the solve has ended, nothing to do
template <int MODE, typename T = double>
__global__ void __launch_bounds__(kBlock)
k_spmv_dot(Dims dm, int nblocks, Mat7T<T> A, const double *__restrict__ z, double *__restrict__ v,
           const double *__restrict__ w1, double *__restrict__ partials,
           const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int lb = xcd_block(nblocks);
  const int i = lb * kBlock + threadIdx.x;
  double red[MODE == 0 ? 1 : 2] = {};
  if (i < dm.n) {
    const double vi = apply7(A, dm, z, i);
    v[i] = vi;
    red[0] = vi * __builtin_nontemporal_load(&w1[i]);
    if (MODE == 1) red[MODE] = vi * vi;
  }
  block_sums<MODE == 0 ? 1 : 2>(red, partials, lb, nblocks);
}

// plain y = A x
__global__ void __launch_bounds__(kBlock)
k_spmv(Dims dm, int nblocks, Mat7 A, const double *__restrict__ x, double *__restrict__ y) {
  const int i = xcd_block(nblocks) * kBlock + threadIdx.x;
  if (i < dm.n) y[i] = apply7(A, dm, x, i);
}

// p = r + beta*(p_in - omega*v); p_in is p itself, or the first search
// direction r0 where that was read in place
__global__ void __launch_bounds__(kBlock)
k_pupdate(int n, double beta, double omega, const double *__restrict__ r,
          const double *__restrict__ p_in, double *__restrict__ p, const double *__restrict__ v) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = r[i] + beta * (p_in[i] - omega * v[i]);
}

// x += a*z; r = r_in - a*v; partials r.r (and r.rs when rs != null).  r_in
// is r itself, or r0 where the solve's first residual is read in place.
__global__ void __launch_bounds__(kBlock)
k_xr_update(int n, int nblocks, const double *__restrict__ num, const double *__restrict__ den,
            int omega_step, double *__restrict__ x, const double *r_in, double *r,
            const double *__restrict__ z, const double *__restrict__ v,
            const double *__restrict__ rs, double *__restrict__ partials,
            const int *__restrict__ skip, int x_zero) {
  if (skip && *skip) {
    // the solve has ended (a breakdown just now): r holds r_in, as below
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (r_in != r && i < n) r[i] = r_in[i];
    return;
  }
  // the coefficient from the reductions just finished (device copies):
  // alpha = rho/(v.r-hat), omega = (t.s)/(t.t); where the host will see a
  // breakdown the update is skipped, as the host loop stops before it
  const double nu = num[0], de = den[0];
  const double a = nu / de;
  const bool ok = omega_step ? (de != 0.0 && a != 0.0 && isfinite(a))
                             : (de != 0.0 && isfinite(de) && isfinite(nu));
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double red[2] = {0.0, 0.0};
  if (i < n) {
    double ri = r_in[i];
    // x_zero: x is the solve's x0 = 0, not read (and written where the
    // update is skipped)
    if (ok) {
      x[i] = (x_zero ? 0.0 : x[i]) + a * z[i];
      ri = ri - a * v[i];
    } else if (x_zero) {
      x[i] = 0.0;
    }
    r[i] = ri;
    red[0] = ri * ri;
    if (rs) red[1] = ri * rs[i];
  }
  block_sums<2>(red, partials, blockIdx.x, nblocks);
}

// generic fused dots over (a, b): mode bits select a.b, |a|_1, a.a
__global__ void __launch_bounds__(kBlock)
k_dots3(int n, int nblocks, const double *__restrict__ a, const double *__restrict__ b,
        double *__restrict__ partials) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double red[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const double ai = a[i];
    red[0] = ai * ai;
    red[1] = b ? b[i] * ai : 0.0;
    red[2] = fabs(ai);
  }
  block_sums<3>(red, partials, blockIdx.x, nblocks);
}

// The sums run_backtr needs in one pass: dh.dh, h_pre.dh, |dh|_1 (k_dots3's
// three) and with ALG also rhs.rhs and kr_r.rhs (the algebraic slope); per
// block in `maxpart` the max of |dh_i|/|h_new_i| over h_new_i != 0
// (k_lam_length's values before the floor at 1, which k_max_final applies).
// Launched on at most kMaxrecBlocks blocks.
template <bool ALG>
__global__ void __launch_bounds__(kBlock)
k_backtr_sums(int n, int nblocks, const double *__restrict__ dh, const double *__restrict__ h_pre,
              const double *__restrict__ rhs, const double *__restrict__ kr_r,
              const double *__restrict__ h_new, double *__restrict__ partials,
              double *__restrict__ maxpart, const int *__restrict__ run_if) {
  if (run_if && *run_if == 0) return;
  __shared__ double sm[kBlock];
  double red[ALG ? 5 : 3] = {};
  double q = 0.0;
  // grid-stride: each thread folds its cells in index order
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += nblocks * kBlock) {
    const double ai = dh[i];
    red[0] += ai * ai;
    red[1] += h_pre[i] * ai;
    red[2] += fabs(ai);
    if (ALG) {
      const double ri = rhs[i];
      red[ALG ? 3 : 0] += ri * ri;
      red[ALG ? 4 : 0] += kr_r[i] * ri;
    }
    const double hn = h_new[i];
    if (hn != 0.0) q = fmax(q, fabs(ai) / fabs(hn));
  }
  block_sums<ALG ? 5 : 3>(red, partials, blockIdx.x, nblocks);
  sm[threadIdx.x] = q;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) maxpart[blockIdx.x] = sm[0];
}

// out = a + s*b (calc_vecjacf's jcvec, Calc_function.f90:602-606)
__global__ void __launch_bounds__(kBlock)
k_axpy(int n, const double *__restrict__ a, double s, const double *__restrict__ b,
       double *__restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = a[i] + s * b[i];
}

__global__ void __launch_bounds__(kBlock)
k_scale_copy(int n, double s, const double *__restrict__ a, double *__restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = s * a[i];
}

__global__ void __launch_bounds__(kBlock)
k_scale(int n, double s, double *__restrict__ a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = a[i] * s;
}

// Final fixed-order sum of NV partial arrays of length nslots -> out[NV].
constexpr int kFin = 1024;
constexpr int kMaxRed = 8;  // sums one reduction carries (partials are kMaxRed x slots)
constexpr int kScal = 19;   // result slots: 0-7 general, 8-15 the BiCGStab scalars, 16 1/delt,
                            // 17 |head_new|^2 at the start of calc_numsol, 18 the rho
                            // k_finalize_gate tested with
constexpr int kDeltSlot = 16;
constexpr int kHnormSlot = 17;
// Publishes `seq` in the host-mapped `flag` once the result stores of
// threads 0..nwriters-1 (made before the call) are visible system-wide: the
// host spins on the flag instead of a stream synchronisation.
__device__ __forceinline__ void publish(unsigned *flag, unsigned seq, int nwriters) {
  if ((int)threadIdx.x < nwriters) __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) *(volatile unsigned *)flag = seq;
}

// The fold of k_finalize (kFin threads; results in threads 0..nv-1, also
// stored into out2/out2_dev when those are given)
__device__ __forceinline__ void fold_sums(const double *__restrict__ partials, int nslots, int nv,
                                          double *__restrict__ out, double *__restrict__ out_dev,
                                          double *__restrict__ out2, double *__restrict__ out2_dev) {
  __shared__ double red[kMaxRed][kFin / 64];
  double s[kMaxRed] = {};
  for (int k = 0; k < nv; k++) {
    const double *p = partials + (size_t)k * nslots;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int q = threadIdx.x;
    for (; q + 3 * kFin < nslots; q += 4 * kFin) {
      a0 += p[q];
      a1 += p[q + kFin];
      a2 += p[q + 2 * kFin];
      a3 += p[q + 3 * kFin];
    }
    for (; q < nslots; q += kFin) a0 += p[q];
    s[k] = (a0 + a1) + (a2 + a3);
  }
  for (int k = 0; k < nv; k++) {
    const double w = wave_sum(s[k]);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double t = red[threadIdx.x][0];
    for (int q = 1; q < kFin / 64; q++) t += red[threadIdx.x][q];
    out[threadIdx.x] = t;
    out_dev[threadIdx.x] = t;
    if (out2) {
      out2[threadIdx.x] = t;
      out2_dev[threadIdx.x] = t;
    }
  }
}

__global__ void __launch_bounds__(kFin)
k_finalize(const double *__restrict__ partials, int nslots, int nv, double *__restrict__ out,
           double *__restrict__ out_dev, unsigned *flag, unsigned seq) {
  fold_sums(partials, nslots, nv, out, out_dev, nullptr, nullptr);
  publish(flag, seq, nv);
}

// The cumulative mass balance of msg123_run_box on the device: the fold of
// k_mass's 6 partial sums (k_finalize's order) and the boundary terms' sums
// bm (well, river, lake, sea, their residual part; null: none) added to
// acc = storage, connection, recharge, surface, well, river, lake, sea,
// total, residual, |recharge|.
constexpr int kMassAcc = 24;  // 11 accumulators, the last fold at 16..21
__global__ void __launch_bounds__(kFin)
k_mass_accum(const double *__restrict__ partials, int nslots, const double *__restrict__ bm,
             double *__restrict__ acc) {
  __shared__ double m[6];
  fold_sums(partials, nslots, 6, acc + 16, acc + 16, nullptr, nullptr);
  __syncthreads();
  if (threadIdx.x < 6) m[threadIdx.x] = acc[16 + threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    const double b[5] = {bm ? bm[0] : 0.0, bm ? bm[1] : 0.0, bm ? bm[2] : 0.0,
                         bm ? bm[3] : 0.0, bm ? bm[4] : 0.0};
    acc[0] += m[0];
    acc[1] += m[1];
    acc[2] += m[2];
    acc[3] += m[3];
    acc[4] += b[0];
    acc[5] += b[1];
    acc[6] += b[2];
    acc[7] += b[3];
    acc[8] += ((((((m[0] + m[1]) + m[2]) + m[3]) + b[0]) + b[1]) + b[2]) + b[3];
    acc[9] += m[4] + b[4];
    acc[10] += m[5];
  }
}

// k_finalize of BiCGStab's |s|^2 (slot 10) followed by the half-step's
// convergence test, made here so that the tail launched behind it and the
// host take the same decision: with rho = sk (NaN: slot 8) and the reference
// norm bnorm (NaN: rho), breakdown where v.r-hat (slot 9) is 0 or either is
// not finite, else converged where sqrt(|s|^2 / bnorm) <= errtol.  The
// decision goes to *gate and, as 0 or 1, to slot 15 of both result copies,
// and the rho it used to slot 18 (the tail launched behind may overwrite
// slot 8 before the host reads it).
__global__ void __launch_bounds__(kFin)
k_finalize_gate(const double *__restrict__ partials, int nslots, double *__restrict__ out,
                double *__restrict__ out_dev, unsigned *flag, unsigned seq, double sk_host,
                double bnorm_host, double errtol, int *__restrict__ gate) {
  fold_sums(partials, nslots, 1, out + 10, out_dev + 10, nullptr, nullptr);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double sk = isnan(sk_host) ? out_dev[8] : sk_host;
    const double bn = isnan(bnorm_host) ? sk : bnorm_host;
    const double sk2 = out_dev[9];
    const bool brk = sk2 == 0.0 || !isfinite(sk2) || !isfinite(sk);
    const bool conv = !brk && sqrt(out_dev[10] / bn) <= errtol;
    *gate = conv ? 1 : 0;
    out[15] = conv ? 1.0 : 0.0;
    out_dev[15] = conv ? 1.0 : 0.0;
    out[18] = sk;
  }
  publish(flag, seq, 1);
}

// k_finalize of one sum into two slot sets (one launch, one publication)
__global__ void __launch_bounds__(kFin)
k_finalize2(const double *__restrict__ partials, int nslots, double *__restrict__ out,
            double *__restrict__ out_dev, double *__restrict__ out2,
            double *__restrict__ out2_dev, unsigned *flag, unsigned seq,
            const int *__restrict__ run_if) {
  // with *run_if == 0 the sums were not made: only the publication
  if (!run_if || *run_if != 0) fold_sums(partials, nslots, 1, out, out_dev, out2, out2_dev);
  publish(flag, seq, 1);
}

// What the evaluation at head_new leaves besides F where a converged
// iteration skips it: S(head_new) into srat (the next step's srat_old) and
// the seepage flags surf_rati of the top layer (func_surfterm's update).
// Runs only where *run_if != 0.
// kr(head_new) goes to kr, where the step's mass balance reads it.
template <bool FAST>
__global__ void __launch_bounds__(kBlock)
k_conv_state(int n, int nxy, const int *__restrict__ run_if, const double *__restrict__ h,
             const double *__restrict__ cell_top, Soil soil, Surface sf,
             double *__restrict__ srat, double *__restrict__ kr) {
  if (*run_if == 0) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double hi = h[i];
    double si, ki, qi;
    vg_point<FAST>(hi - cell_top[i] + 0.0, soil.reta[i], soil.retn[i], soil.resi[i],
                   soil.poro[i], soil.ss[i], si, ki, qi);
    srat[i] = si;
    kr[i] = ki;
    if (i < nxy) (void)surf_term(sf, i, hi, ki);
  }
}

// ---------------------------------------------------------------------------
// check_abserrmax (Check_simulation.f90:37-91) fused with the Newton update
// head_new = head_pre + dh (Nonlinear_solution.f90:132-134): max |dh| and
// max |h| with the last index winning ties, and a NaN flag.  Per block:
// (ach, idx_ch, aval, idx_val, nan); finalised on the host in block order.
struct MaxRec {
  double ach, sch;    // |dh| and signed dh at ich
  double aval, sval;  // |h| and signed h at ival
  int ich;
  int ival;
  int nan;
  int pad;
};

__device__ __forceinline__ void maxrec_merge(MaxRec &a, const MaxRec &b) {
  // the larger value wins, on ties the larger index: order-independent
  if (b.ich >= 0 && (a.ich < 0 || b.ach > a.ach || (b.ach == a.ach && b.ich > a.ich))) {
    a.ach = b.ach;
    a.sch = b.sch;
    a.ich = b.ich;
  }
  if (b.ival >= 0 && (a.ival < 0 || b.aval > a.aval || (b.aval == a.aval && b.ival > a.ival))) {
    a.aval = b.aval;
    a.sval = b.sval;
    a.ival = b.ival;
  }
  a.nan |= b.nan;
}

// Grid-stride over the cells (kMaxrecBlocks blocks at most): each thread
// folds its cells into one record before the block's LDS tree.
constexpr int kMaxrecBlocks = 2048;
__global__ void __launch_bounds__(kBlock)
k_update_maxerr(int n, const double *__restrict__ h_pre, const double *__restrict__ dh,
                double *__restrict__ h_new, int do_update, MaxRec *__restrict__ out,
                const int *__restrict__ run_if) {
  if (run_if && *run_if == 0) return;
  __shared__ MaxRec sm[kBlock];
  MaxRec m{0.0, 0.0, 0.0, 0.0, -1, -1, 0, 0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double hn;
    if (do_update) {
      hn = h_pre[i] + dh[i];
      h_new[i] = hn;
    } else {
      hn = h_new[i];
    }
    const double d = hn - h_pre[i];
    MaxRec q;
    q.ach = fabs(d);
    q.sch = d;
    q.ich = i;
    q.aval = fabs(hn);
    q.sval = hn;
    q.ival = i;
    q.nan = (hn != hn) ? 1 : 0;
    if (q.ach != q.ach) q.ich = -1;
    if (q.aval != q.aval) q.ival = -1;
    maxrec_merge(m, q);
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      MaxRec a = sm[threadIdx.x];
      maxrec_merge(a, sm[threadIdx.x + s]);
      sm[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// The Newton update and everything the convergence test and the line search
// start from, in one pass (algebraic slope): head_new = head_pre + dh with
// check_abserrmax's records (as k_update_maxerr), and run_backtr's sums
// |dh|^2, head_pre.dh, |dh|_1, |rhs|^2, r.rhs and lam_length's max (as
// k_backtr_sums<true>, same grid and per-thread order, so the same sums).
__global__ void __launch_bounds__(kBlock)
k_newton_tail(int n, int nblocks, const double *__restrict__ h_pre, const double *__restrict__ dh,
              double *__restrict__ h_new, const double *__restrict__ rhs,
              const double *__restrict__ kr_r, MaxRec *__restrict__ rec,
              double *__restrict__ partials, double *__restrict__ maxpart,
              const int *__restrict__ run_if) {
  if (run_if && *run_if == 0) return;
  __shared__ MaxRec sm[kBlock];
  __shared__ double sq[kBlock];
  MaxRec m{0.0, 0.0, 0.0, 0.0, -1, -1, 0, 0};
  double red[5] = {};
  double q = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += nblocks * kBlock) {
    const double hp = h_pre[i], ai = dh[i], ri = rhs[i];
    const double hn = hp + ai;
    h_new[i] = hn;
    const double d = hn - hp;
    MaxRec t;
    t.ach = fabs(d);
    t.sch = d;
    t.ich = i;
    t.aval = fabs(hn);
    t.sval = hn;
    t.ival = i;
    t.nan = (hn != hn) ? 1 : 0;
    if (t.ach != t.ach) t.ich = -1;
    if (t.aval != t.aval) t.ival = -1;
    maxrec_merge(m, t);
    red[0] += ai * ai;
    red[1] += hp * ai;
    red[2] += fabs(ai);
    red[3] += ri * ri;
    red[4] += kr_r[i] * ri;
    if (hn != 0.0) q = fmax(q, fabs(ai) / fabs(hn));
  }
  block_sums<5>(red, partials, blockIdx.x, nblocks);
  sm[threadIdx.x] = m;
  sq[threadIdx.x] = q;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      MaxRec a = sm[threadIdx.x];
      maxrec_merge(a, sm[threadIdx.x + s]);
      sm[threadIdx.x] = a;
      sq[threadIdx.x] = fmax(sq[threadIdx.x], sq[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rec[blockIdx.x] = sm[0];
    maxpart[blockIdx.x] = sq[0];
  }
}

// One block folds the per-block records: the merge picks the largest value
// and, on ties, the largest index, so any order gives the result of the
// serial scan with the last index winning ties.
constexpr int kFinMax = 1024;
__global__ void __launch_bounds__(kFinMax)
k_maxrec_final(int nslots, const MaxRec *__restrict__ in, MaxRec *__restrict__ out,
               unsigned *flag, unsigned seq) {
  __shared__ MaxRec sm[kFinMax];
  MaxRec m{0.0, 0.0, 0.0, 0.0, -1, -1, 0, 0};
  for (int q = threadIdx.x; q < nslots; q += kFinMax) maxrec_merge(m, in[q]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kFinMax / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      MaxRec a = sm[threadIdx.x];
      maxrec_merge(a, sm[threadIdx.x + s]);
      sm[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
  publish(flag, seq, 1);
}

// max of nslots values (all >= 1) into out[0]
__global__ void __launch_bounds__(kFinMax)
k_max_final(int nslots, const double *__restrict__ in, double *__restrict__ out) {
  __shared__ double sm[kFinMax];
  double v = 1.0;
  for (int q = threadIdx.x; q < nslots; q += kFinMax) v = fmax(v, in[q]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = kFinMax / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sm[0];
    __threadfence_system();
  }
}

// The diagonal the convergence log reports: FP64, or the FP32 Newton system's
struct DiagLog {
  const double *d;
  const float *f;
  __device__ __forceinline__ double at(int i) const { return f ? (double)f[i] : d[i]; }
};

// out[0..2] = a[i], b[i], c[i] (one launch instead of three copies)
__global__ void k_pick3(int i, DiagLog a, const double *__restrict__ b,
                        const double *__restrict__ c, double *__restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = a.at(i);
    out[1] = b[i];
    out[2] = c[i];
  }
}

// Up to three device-to-device copies in one launch.  The runtime's blit
// copies (hipMemcpyAsync device to device) left about 30 us between two
// consecutive copies in a kernel trace; a kernel of the library's own
// starts behind the previous kernel like any other.
struct CopyJob {
  double *dst[3];
  const double *src[3];
  size_t n[3];
};
__global__ void __launch_bounds__(kBlock) k_copy3(CopyJob j) {
  const size_t step = (size_t)gridDim.x * kBlock;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    double *__restrict__ d = j.dst[q];
    const double *__restrict__ s = j.src[q];
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < j.n[q]; i += step) d[i] = s[i];
  }
}

__global__ void __launch_bounds__(kBlock) k_zero(size_t n, double *__restrict__ d) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    d[i] = 0.0;
}

__global__ void k_set1(double *p, double v) {
  if (threadIdx.x == 0) *p = v;
}

// k_pick3 at the cell of the largest change of a folded max record
__global__ void k_pick3_rec(const MaxRec *__restrict__ m, DiagLog a,
                            const double *__restrict__ b, const double *__restrict__ c,
                            double *__restrict__ out) {
  if (threadIdx.x == 0) {
    const int i = m->ich >= 0 ? m->ich : 0;
    out[0] = a.at(i);
    out[1] = b[i];
    out[2] = c[i];
  }
}

// The Newton tail's folds in one launch: k_maxrec_final (the record into
// rec_out), k_pick3_rec at its cell (a, b, c into pick_out), k_max_final
// (maxpart into max_out) and k_finalize of nv sums, each in its own
// kernel's operation order; one publication of seq covers them all.
__global__ void __launch_bounds__(kFin)
k_tail_final(int nslots, const MaxRec *__restrict__ recs, MaxRec *__restrict__ rec_out,
             DiagLog a, const double *__restrict__ b,
             const double *__restrict__ c, double *__restrict__ pick_out,
             const double *__restrict__ maxpart, double *__restrict__ max_out,
             const double *__restrict__ partials, int nv, double *__restrict__ out,
             double *__restrict__ out_dev, unsigned *flag, unsigned seq, int lconv, double L,
             double criteria, int *__restrict__ conv_run, const int *__restrict__ run_if) {
  static_assert(kFin == kFinMax, "one block shape for every fold");
  if (run_if && *run_if == 0) {
    // launched ahead of a Krylov solve that went on: nothing of the tail runs
    if (threadIdx.x == 0 && conv_run) {
      conv_run[0] = 0;
      conv_run[1] = 0;
    }
    return;
  }
  {
    __shared__ MaxRec sm[kFinMax];
    MaxRec m{0.0, 0.0, 0.0, 0.0, -1, -1, 0, 0};
    for (int q = threadIdx.x; q < nslots; q += kFinMax) maxrec_merge(m, recs[q]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = kFinMax / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        MaxRec t = sm[threadIdx.x];
        maxrec_merge(t, sm[threadIdx.x + s]);
        sm[threadIdx.x] = t;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const MaxRec r = sm[0];
      rec_out[0] = r;
      const int i = r.ich >= 0 ? r.ich : 0;
      pick_out[0] = a.at(i);
      pick_out[1] = b[i];
      pick_out[2] = c[i];
      if (conv_run) {
        // calc_numsol's convergence test at the full step, as the host makes
        // it from this record: converged, nothing reads F(head_new), and
        // conv_run[0] = 0 switches off what computes it; conv_run[1] = 1
        // runs k_conv_state instead
        const bool conv = lconv != 0 && !r.nan &&
                          fabs((r.ich >= 0) ? r.sch : 0.0) * L <= criteria &&
                          fabs((r.ival >= 0) ? r.sval : 0.0) * L < XMAX;
        conv_run[0] = conv ? 0 : 1;
        conv_run[1] = conv ? 1 : 0;
      }
    }
  }
  {
    __shared__ double sm[kFinMax];
    double v = 1.0;
    for (int q = threadIdx.x; q < nslots; q += kFinMax) v = fmax(v, maxpart[q]);
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = kFinMax / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) max_out[0] = sm[0];
  }
  fold_sums(partials, nslots, nv, out, out_dev, nullptr, nullptr);
  publish(flag, seq, nv > 0 ? nv : 1);
}

// ===========================================================================
// Device-driven Newton iterations (msg123_set_control): the scalars of the
// linear solve (solve_bicgs, Linear_solution.f90:234-493) and of the Newton
// iteration (calc_numsol, Nonlinear_solution.f90:73-290) live in device
// memory and are updated by one thread after each reduction fold, so the
// kernels of a Newton iteration are enqueued without the host waiting
// between them: the Krylov kernels skip once the solve has ended, the tail
// (update, line-search sums, residual at the new head, decision) runs only
// once it has, and the host reads one record per iteration.  Every update
// uses the host driver's formulas in its order.
enum DevStatus {
  kSolving = 0,    // the linear solve has not ended: enqueue more iterations
  kConverged = 1,  // Newton converged
  kContinue = 2,   // full step accepted by the line search, not converged
  kBackStop = 3,   // check_val >= VARMAX or |h| >= XMAX after the step
  kNeedHost = 4,   // the line search needs more than the full step: host driver
  kNaNStep = 5     // NaN in the update
};

struct DevNewton {
  // the linear solve
  double rho, rho0, alpha, omega, sk2, ts, tt, rnorm, bnorm, errtol, r0;
  int it, maxinn, done, conv, breakdown, pad0;
  // the Newton iteration
  double l2_new, l2_pre, l2_jnorm, eta, gradient, maxstep, eta_max, criteria, L;
  double sums[5], lam_length, slope0, eps, eps_inv, l2_h;
  double rec_dmat, rec_rhs, rec_head;
  MaxRec mr;
  int out_iter, status, back_flag, conv_flag, alg, searched;
  int search_go;  // the iteration's full step goes to the line search
  int pad1;
};

enum FoldStage { kStR0, kStS1, kStS2, kStS3, kStS4, kStSums, kStSlope, kStL2 };

// Fixed-order fold of nv partial arrays (k_finalize's order), then the
// stage's scalar update by thread 0.
template <int STAGE>
__global__ void __launch_bounds__(kFin)
k_fold_dev(const double *__restrict__ partials, int nslots, int nv, DevNewton *dn) {
  const bool tail = STAGE == kStSums || STAGE == kStL2;
  if (STAGE == kStSlope) {
    if (dn->search_go == 0) return;
  } else if (STAGE != kStR0 && (tail ? dn->done == 0 : dn->done != 0)) {
    return;
  }
  __shared__ double red[kMaxRed][kFin / 64];
  double s[kMaxRed] = {};
  for (int k = 0; k < nv; k++) {
    const double *p = partials + (size_t)k * nslots;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int q = threadIdx.x;
    for (; q + 3 * kFin < nslots; q += 4 * kFin) {
      a0 += p[q];
      a1 += p[q + kFin];
      a2 += p[q + 2 * kFin];
      a3 += p[q + 3 * kFin];
    }
    for (; q < nslots; q += kFin) a0 += p[q];
    s[k] = (a0 + a1) + (a2 + a3);
  }
  for (int k = 0; k < nv; k++) {
    const double w = wave_sum(s[k]);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[kMaxRed];
  for (int k = 0; k < nv; k++) {
    double t = red[k][0];
    for (int q = 1; q < kFin / 64; q++) t += red[k][q];
    v[k] = t;
  }
  DevNewton &d = *dn;
  if (STAGE == kStR0) {
    // a Newton iteration begins: |rhs|^2 of the assembly starts the solve
    // from x0 = 0 (r0 = rhs); the first iteration's norms are |rhs|^2
    d.out_iter += 1;
    if (d.out_iter > 1) d.l2_pre = d.l2_new;
    d.r0 = v[0];
    d.rho = v[0];
    if (d.out_iter == 1) {
      d.bnorm = v[0];
      d.l2_new = v[0];
      d.l2_pre = v[0];
    } else {
      d.bnorm = d.l2_pre;
    }
    d.errtol = d.eta;
    d.it = 0;
    d.done = 0;
    d.conv = 0;
    d.breakdown = 0;
    d.rnorm = 0.0;
    d.status = kSolving;
    d.back_flag = 0;
    d.conv_flag = 0;
    d.search_go = 0;
  } else if (STAGE == kStS1) {
    d.it += 1;
    d.sk2 = v[0];
    if (d.sk2 == 0.0 || !isfinite(d.sk2) || !isfinite(d.rho)) {
      d.breakdown = 1;
      d.done = 1;
    } else {
      d.alpha = d.rho / d.sk2;
    }
  } else if (STAGE == kStS2) {
    d.rnorm = v[0];
    d.conv = (sqrt(d.rnorm / d.bnorm) <= d.errtol) ? 1 : 0;
    if (d.conv) d.done = 1;
  } else if (STAGE == kStS3) {
    d.ts = v[0];
    d.tt = v[1];
    d.omega = d.ts / d.tt;
    if (d.tt == 0.0 || d.omega == 0.0 || !isfinite(d.omega)) {
      d.breakdown = 1;
      d.done = 1;
    }
  } else if (STAGE == kStS4) {
    d.rnorm = v[0];
    d.conv = (sqrt(d.rnorm / d.bnorm) <= d.errtol) ? 1 : 0;
    if (d.conv || d.it == d.maxinn) {
      d.done = 1;
    } else {
      d.rho0 = d.rho;
      d.rho = v[1];
    }
  } else if (STAGE == kStSums) {
    for (int k = 0; k < nv; k++) d.sums[k] = v[k];
  } else if (STAGE == kStSlope) {
    d.slope0 = v[0];
  } else if (STAGE == kStL2) {
    d.l2_h = v[0];
  }
}

// p = r + beta (p_in - omega v) with beta = (rho/rho0)(alpha/omega)
__global__ void __launch_bounds__(kBlock)
k_pupdate_dev(int n, const DevNewton *__restrict__ dn, const double *__restrict__ r,
              const double *__restrict__ p_in, double *__restrict__ p,
              const double *__restrict__ v) {
  if (dn->done) return;
  const double beta = (dn->rho / dn->rho0) * (dn->alpha / dn->omega), omega = dn->omega;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = r[i] + beta * (p_in[i] - omega * v[i]);
}

// dst = src unless the solve has ended (the AMG path's record of the last
// preconditioner input, Linear_solution.f90:525-527)
__global__ void __launch_bounds__(kBlock)
k_copy_skip(int n, const int *__restrict__ skip, const double *__restrict__ src,
            double *__restrict__ dst) {
  if (*skip) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock)
k_zero(int n, double *__restrict__ x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// the largest change record of the update, folded into the control record
__global__ void __launch_bounds__(kFinMax)
k_maxrec_dev(int nslots, const MaxRec *__restrict__ in, DevNewton *dn) {
  if (dn->done == 0) return;
  __shared__ MaxRec sm[kFinMax];
  MaxRec m{0.0, 0.0, 0.0, 0.0, -1, -1, 0, 0};
  for (int q = threadIdx.x; q < nslots; q += kFinMax) maxrec_merge(m, in[q]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = kFinMax / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      MaxRec a = sm[threadIdx.x];
      maxrec_merge(a, sm[threadIdx.x + s]);
      sm[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // does the full step go to the line search (calc_numsol :132-146)?
    const MaxRec m = sm[0];
    DevNewton &d = *dn;
    d.mr = m;
    const double check_val = fabs((m.ich >= 0) ? m.sch : 0.0) * d.L;
    const double maxunk = fabs((m.ival >= 0) ? m.sval : 0.0) * d.L;
    d.search_go = (m.nan || (d.conv && check_val <= d.criteria && maxunk < XMAX)) ? 0 : 1;
  }
}

// lam_length of the uncapped step (k_max_final into the control record)
__global__ void __launch_bounds__(kFinMax)
k_max_dev(int nslots, const double *__restrict__ in, DevNewton *dn) {
  if (dn->done == 0) return;
  __shared__ double sm[kFinMax];
  double v = 1.0;
  for (int q = threadIdx.x; q < nslots; q += kFinMax) v = fmax(v, in[q]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = kFinMax / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) dn->lam_length = sm[0];
}

// calc_vecjacf's step from the sums (v.v, x.v, |v|_1 with x = head_pre,
// v = dh), then f1 = F(head_pre) = -rhs and jc = head_pre + eps dh
__global__ void __launch_bounds__(kBlock)
k_fd_prep(int n, DevNewton *dn, const double *__restrict__ rhs, const double *__restrict__ h_pre,
          const double *__restrict__ dh, double *__restrict__ f1, double *__restrict__ jc) {
  if (dn->search_go == 0) return;
  const double l2_v = dn->sums[0], l2_x = dn->sums[1], l1_v = dn->sums[2];
  const double sgn = (l2_x < 0.0) ? -1.0 : 1.0;
  const double eps = sgn * sqrt(MACHI_EPS) * fmax(fabs(l2_x), l1_v) / l2_v;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    dn->eps = eps;
    dn->eps_inv = 1.0 / eps;
  }
  if (i < n) {
    f1[i] = -1.0 * rhs[i];
    jc[i] = h_pre[i] + eps * dh[i];
  }
}

// The end of a Newton iteration whose full step the line search accepts:
// calc_numsol's convergence tests (Nonlinear_solution.f90:132-178),
// run_backtr's sufficient-decrease test at lambda = 1 (:441-631; with the
// algebraic slope its guard) and set_eater (:414-439).  Anything else (a
// capped step, backtracking, the guard's difference-quotient slope) is left
// to the host driver (kNeedHost).  Then the convergence-log values at the
// cell of the largest change, and the record is published to the host.
__global__ void k_decide(DevNewton *dn, DiagLog dmat_log,
                         const double *__restrict__ rhs_log, const double *__restrict__ h_new,
                         DevNewton *dn_map, unsigned *flag, unsigned seq) {
  if (threadIdx.x == 0) {
    DevNewton &d = *dn;
    if (d.done == 0) {
      d.status = kSolving;
    } else {
      const MaxRec m = d.mr;
      d.l2_jnorm = d.rnorm;
      d.searched = 0;
      if (m.nan) {
        d.status = kNaNStep;
        d.conv_flag = 0;
      } else {
        const double maxvar = (m.ich >= 0) ? m.sch : 0.0;
        const double maxunk = fabs((m.ival >= 0) ? m.sval : 0.0) * d.L;
        const double check_val = fabs(maxvar) * d.L;
        int conv_flag = d.conv;
        if (conv_flag != 0) {
          conv_flag = 0;
          if (check_val <= d.criteria && maxunk < XMAX) conv_flag = 1;
        }
        if (conv_flag == 1) {
          d.conv_flag = 1;
          d.l2_new = d.l2_h;
          d.status = kConverged;
        } else {
          d.searched = 1;
          const double back_alpha = 1.0e-4;
          const double l2_pnorm = d.sums[0];
          double slope = d.slope0;
          bool host = sqrt(l2_pnorm) > d.maxstep;
          if (!host && d.alg) {
            slope = -d.sums[3] + d.sums[4];
            if (0.5 * d.l2_h > 0.5 * d.l2_pre + back_alpha * slope * 1.0 && d.l2_h > d.l2_pre)
              host = true;
          }
          if (!host && 0.5 * d.l2_h > 0.5 * d.l2_pre + back_alpha * 1.0 * slope) host = true;
          if (host) {
            d.status = kNeedHost;
          } else {
            d.l2_new = d.l2_h;
            d.gradient = slope;
            d.back_flag = (check_val >= VARMAX || maxunk >= XMAX) ? 1 : 0;
            if (!d.back_flag && check_val <= d.criteria && maxunk < XMAX) conv_flag = 1;
            d.conv_flag = conv_flag;
            if (conv_flag != 1 && d.back_flag != 1) {
              // set_eater with lambda = 1
              const double lam = 1.0, eta_min = 1.0e-4, eta_alpha = (1.0 + sqrt(5.0)) / 2.0;
              double eta_safe = pow(d.eta, eta_alpha);
              const double minus1 = 1.0 - lam;
              const double lin = sqrt(minus1 * minus1 * d.l2_pre + 2.0 * lam * minus1 * d.gradient +
                                      lam * lam * d.l2_jnorm);
              double eta = fabs(sqrt(d.l2_new) - lin) / sqrt(d.l2_pre);
              if (eta_safe < 0.1) eta_safe = 0.0;
              eta = fmax(eta, eta_safe);
              eta = fmax(eta, eta_min);
              d.eta = fmin(eta, d.eta_max);
            }
            d.status = conv_flag == 1 ? kConverged : (d.back_flag ? kBackStop : kContinue);
          }
        }
      }
      const int mi = m.ich >= 0 ? m.ich : 0;
      d.rec_dmat = dmat_log.at(mi);
      d.rec_rhs = rhs_log[mi];
      d.rec_head = h_new[mi];
    }
    *dn_map = d;
    __threadfence_system();
  }
  __syncthreads();
  if (threadIdx.x == 0) *(volatile unsigned *)flag = seq;
}

// set the host driver's values after an iteration it finished itself
__global__ void k_dn_set(DevNewton *dn, double eta, double l2_new, double l2_pre) {
  if (threadIdx.x == 0) {
    dn->eta = eta;
    dn->l2_new = l2_new;
    dn->l2_pre = l2_pre;
  }
}

// a calc_numsol call begins
__global__ void k_dn_init(DevNewton *dn, double eta_first, double eta_max, double criteria,
                          double L, double maxstep, int maxinn, int alg) {
  if (threadIdx.x == 0) {
    DevNewton d{};
    d.eta = eta_first;
    d.eta_max = eta_max;
    d.criteria = criteria;
    d.L = L;
    d.maxstep = maxstep;
    d.maxinn = maxinn;
    d.alg = alg;
    d.out_iter = 0;
    *dn = d;
  }
}

// lam_length = max(1, max |dh_i|/|h_i| over h_i != 0) (run_backtr :504-513)
__global__ void __launch_bounds__(kBlock)
k_lam_length(int n, const double *__restrict__ dh, const double *__restrict__ h,
             double *__restrict__ out) {
  __shared__ double sm[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v = 1.0;
  if (i < n && h[i] != 0.0) v = fmax(1.0, fabs(dh[i]) / fabs(h[i]));
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// ===========================================================================
// D-ILU(0) (precon_dilu / solve_ilu, Linear_solution.f90:742-855), bit-exact
// to the serial natural-order sweeps.  The box is cut into tiles of
// kTX x kTY x kTZ cells; tile (I,J,K) depends only on (I-1,J,K), (I,J-1,K),
// (I,J,K-1) (reverse for the backward sweep), so all tiles with I+J+K = D
// run in one launch, diagonals in order.  A tile's operands and the faces it
// needs from its predecessor tiles are first loaded into LDS (one latency
// for the whole tile), then one thread per x-line (jj,kk) handles cell ii at
// step t = ii + jj + kk: its x-, y- and z-predecessors were finished in
// step t-1, so every step reads and writes LDS only, with one barrier.
// The result is written back to HBM coalesced at the end.
constexpr int kTX = 16, kTY = 16, kTZ = 8;
constexpr int kTC = kTX * kTY * kTZ;     // cells per tile
constexpr int kTT = kTY * kTZ;           // threads per tile (one per x-line)

struct Tile {
  int I, J, K, pad;
};

__device__ __forceinline__ int lidx(int ii, int jj, int kk) { return (kk * kTY + jj) * kTX + ii; }

struct TileGeo {
  int x0, y0, z0;
  __device__ TileGeo(const Tile &t) : x0(t.I * kTX), y0(t.J * kTY), z0(t.K * kTZ) {}
  __device__ bool inside(const Dims &dm, int ii, int jj, int kk) const {
    return x0 + ii < dm.nx && y0 + jj < dm.ny && z0 + kk < dm.nz;
  }
  __device__ int cell(const Dims &dm, int ii, int jj, int kk) const {
    return ((z0 + kk) * dm.ny + (y0 + jj)) * dm.nx + (x0 + ii);
  }
};

// Load NA arrays of one tile into LDS.  All loads of a thread are issued
// before the first LDS store, so the tile costs one HBM latency.
template <int NA>
__device__ __forceinline__ void preload(const Dims &dm, const TileGeo &g,
                                        const double *const (&src)[NA], double *const (&dst)[NA]) {
  constexpr int R = kTC / kTT;
  double v[NA][R];
  int cc[R];
#pragma unroll
  for (int it = 0; it < R; it++) {
    const int q = threadIdx.x + it * kTT;
    const int ii = q % kTX, jj = (q / kTX) % kTY, kk = q / (kTX * kTY);
    cc[it] = g.inside(dm, ii, jj, kk) ? g.cell(dm, ii, jj, kk) : -1;
#pragma unroll
    for (int a = 0; a < NA; a++) v[a][it] = cc[it] >= 0 ? src[a][cc[it]] : 0.0;
  }
#pragma unroll
  for (int it = 0; it < R; it++) {
    const int q = threadIdx.x + it * kTT;
#pragma unroll
    for (int a = 0; a < NA; a++) dst[a][q] = v[a][it];
  }
}

// forward sweep y = (D+L)^-1 r scaled by D: y_i = (r_i - sum_{up,N,W} l y)/d_i
__device__ __forceinline__ void ilu_fwd_tile(const Dims &dm, const Tile &tile, const Mat7 &A,
                                             const double *__restrict__ dg,
                                             const double *__restrict__ r,
                                             double *__restrict__ out) {
  __shared__ double sY[kTC], sUp[kTC], sN[kTC], sW[kTC], sD[kTC];
  __shared__ double hW[kTZ][kTY], hN[kTZ][kTX], hU[kTY][kTX];
  const TileGeo g(tile);
  {
    const double *src[5] = {r, A.up, A.n, A.w, dg};
    double *dst[5] = {sY, sUp, sN, sW, sD};
    preload<5>(dm, g, src, dst);
  }
  for (int q = threadIdx.x; q < kTX * kTY + 2 * kTX * kTZ; q += kTT) {
    if (q < kTY * kTX) {  // upper face: (ii, jj, z0-1)
      const int ii = q % kTX, jj = q / kTX;
      if (g.z0 > 0 && g.inside(dm, ii, jj, 0)) hU[jj][ii] = out[g.cell(dm, ii, jj, 0) - dm.nxy];
    } else if (q < kTY * kTX + kTZ * kTX) {  // north face: (ii, y0-1, kk)
      const int e = q - kTY * kTX, ii = e % kTX, kk = e / kTX;
      if (g.y0 > 0 && g.inside(dm, ii, 0, kk)) hN[kk][ii] = out[g.cell(dm, ii, 0, kk) - dm.nx];
    } else {  // west face: (x0-1, jj, kk)
      const int e = q - kTY * kTX - kTZ * kTX;
      if (e < kTZ * kTY) {
        const int jj = e % kTY, kk = e / kTY;
        if (g.x0 > 0 && g.inside(dm, 0, jj, kk)) hW[kk][jj] = out[g.cell(dm, 0, jj, kk) - 1];
      }
    }
  }
  __syncthreads();
  const int jj = threadIdx.x % kTY, kk = threadIdx.x / kTY;
  for (int t = 0; t < kTX + kTY + kTZ - 2; t++) {
    const int ii = t - jj - kk;
    if (ii >= 0 && ii < kTX && g.inside(dm, ii, jj, kk)) {
      const int q = lidx(ii, jj, kk);
      double tmp = sY[q];
      if (g.z0 + kk > 0) tmp = tmp - sUp[q] * ((kk > 0) ? sY[q - kTX * kTY] : hU[jj][ii]);
      if (g.y0 + jj > 0) tmp = tmp - sN[q] * ((jj > 0) ? sY[q - kTX] : hN[kk][ii]);
      if (g.x0 + ii > 0) tmp = tmp - sW[q] * ((ii > 0) ? sY[q - 1] : hW[kk][jj]);
      sY[q] = tmp / sD[q];
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < kTC; q += kTT) {
    const int ii = q % kTX, jj2 = (q / kTX) % kTY, kk2 = q / (kTX * kTY);
    if (g.inside(dm, ii, jj2, kk2)) out[g.cell(dm, ii, jj2, kk2)] = sY[q];
  }
}

__global__ void __launch_bounds__(kTT)
k_ilu_fwd(Dims dm, const Tile *__restrict__ tiles, Mat7 A, const double *__restrict__ dg,
          const double *__restrict__ r, double *__restrict__ out, const int *__restrict__ skip) {
  if (skip && *skip) return;
  ilu_fwd_tile(dm, tiles[blockIdx.x], A, dg, r, out);
}

// backward sweep in place: z_i = y_i - (sum_{E,S,down} u z)/d_i
__device__ __forceinline__ void ilu_bwd_tile(const Dims &dm, const Tile &tile, const Mat7 &A,
                                             const double *__restrict__ dg,
                                             double *__restrict__ out) {
  __shared__ double sY[kTC], sE[kTC], sS[kTC], sDn[kTC], sD[kTC];
  __shared__ double hE[kTZ][kTY], hS[kTZ][kTX], hB[kTY][kTX];
  const TileGeo g(tile);
  {
    const double *src[5] = {out, A.e, A.s, A.dn, dg};
    double *dst[5] = {sY, sE, sS, sDn, sD};
    preload<5>(dm, g, src, dst);
  }
  for (int q = threadIdx.x; q < kTX * kTY + 2 * kTX * kTZ; q += kTT) {
    if (q < kTY * kTX) {  // lower face: (ii, jj, z0+kTZ)
      const int ii = q % kTX, jj = q / kTX;
      if (g.z0 + kTZ < dm.nz && g.inside(dm, ii, jj, 0))
        hB[jj][ii] = out[g.cell(dm, ii, jj, 0) + kTZ * dm.nxy];
    } else if (q < kTY * kTX + kTZ * kTX) {  // south face: (ii, y0+kTY, kk)
      const int e = q - kTY * kTX, ii = e % kTX, kk = e / kTX;
      if (g.y0 + kTY < dm.ny && g.inside(dm, ii, 0, kk))
        hS[kk][ii] = out[g.cell(dm, ii, 0, kk) + kTY * dm.nx];
    } else {  // east face: (x0+kTX, jj, kk)
      const int e = q - kTY * kTX - kTZ * kTX;
      if (e < kTZ * kTY) {
        const int jj = e % kTY, kk = e / kTY;
        if (g.x0 + kTX < dm.nx && g.inside(dm, 0, jj, kk))
          hE[kk][jj] = out[g.cell(dm, 0, jj, kk) + kTX];
      }
    }
  }
  __syncthreads();
  const int jj = threadIdx.x % kTY, kk = threadIdx.x / kTY;
  for (int t = 0; t < kTX + kTY + kTZ - 2; t++) {
    const int ii = (kTX - 1) - (t - (kTY - 1 - jj) - (kTZ - 1 - kk));
    if (ii >= 0 && ii < kTX && g.inside(dm, ii, jj, kk)) {
      const int q = lidx(ii, jj, kk);
      const int xx = g.x0 + ii, yy = g.y0 + jj, zz = g.z0 + kk;
      double tmp = 0.0;
      if (xx < dm.nx - 1) tmp = tmp + sE[q] * ((ii < kTX - 1) ? sY[q + 1] : hE[kk][jj]);
      if (yy < dm.ny - 1) tmp = tmp + sS[q] * ((jj < kTY - 1) ? sY[q + kTX] : hS[kk][ii]);
      if (zz < dm.nz - 1) tmp = tmp + sDn[q] * ((kk < kTZ - 1) ? sY[q + kTX * kTY] : hB[jj][ii]);
      sY[q] = sY[q] - tmp / sD[q];
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < kTC; q += kTT) {
    const int ii = q % kTX, jj2 = (q / kTX) % kTY, kk2 = q / (kTX * kTY);
    if (g.inside(dm, ii, jj2, kk2)) out[g.cell(dm, ii, jj2, kk2)] = sY[q];
  }
}

__global__ void __launch_bounds__(kTT)
k_ilu_bwd(Dims dm, const Tile *__restrict__ tiles, Mat7 A, const double *__restrict__ dg,
          double *__restrict__ out, const int *__restrict__ skip) {
  if (skip && *skip) return;
  ilu_bwd_tile(dm, tiles[blockIdx.x], A, dg, out);
}

// factor: d_i = a_ii - sum_{lower k} (l_ik * (1/d_k)) * u_ki, lower in CSR order up, N, W
__device__ __forceinline__ void ilu_factor_tile(const Dims &dm, const Tile &tile, const Mat7 &A,
                                                double *__restrict__ dg) {
  __shared__ double sD[kTC], sUp[kTC], sN[kTC], sW[kTC], sUE[kTC], sUS[kTC], sUD[kTC];
  __shared__ double hWd[kTZ][kTY], hWu[kTZ][kTY], hNd[kTZ][kTX], hNu[kTZ][kTX];
  __shared__ double hUd[kTY][kTX], hUu[kTY][kTX];
  const TileGeo g(tile);
  {
    const double *src[7] = {A.d, A.up, A.n, A.w, A.e, A.s, A.dn};
    double *dst[7] = {sD, sUp, sN, sW, sUE, sUS, sUD};
    preload<7>(dm, g, src, dst);
  }
  for (int q = threadIdx.x; q < kTX * kTY + 2 * kTX * kTZ; q += kTT) {
    if (q < kTY * kTX) {
      const int ii = q % kTX, jj = q / kTX;
      if (g.z0 > 0 && g.inside(dm, ii, jj, 0)) {
        const int c = g.cell(dm, ii, jj, 0) - dm.nxy;
        hUd[jj][ii] = dg[c];
        hUu[jj][ii] = A.dn[c];
      }
    } else if (q < kTY * kTX + kTZ * kTX) {
      const int e = q - kTY * kTX, ii = e % kTX, kk = e / kTX;
      if (g.y0 > 0 && g.inside(dm, ii, 0, kk)) {
        const int c = g.cell(dm, ii, 0, kk) - dm.nx;
        hNd[kk][ii] = dg[c];
        hNu[kk][ii] = A.s[c];
      }
    } else {
      const int e = q - kTY * kTX - kTZ * kTX;
      if (e < kTZ * kTY) {
        const int jj = e % kTY, kk = e / kTY;
        if (g.x0 > 0 && g.inside(dm, 0, jj, kk)) {
          const int c = g.cell(dm, 0, jj, kk) - 1;
          hWd[kk][jj] = dg[c];
          hWu[kk][jj] = A.e[c];
        }
      }
    }
  }
  __syncthreads();
  const int jj = threadIdx.x % kTY, kk = threadIdx.x / kTY;
  for (int t = 0; t < kTX + kTY + kTZ - 2; t++) {
    const int ii = t - jj - kk;
    if (ii >= 0 && ii < kTX && g.inside(dm, ii, jj, kk)) {
      const int q = lidx(ii, jj, kk);
      double d = sD[q];
      if (g.z0 + kk > 0) {
        const double du = (kk > 0) ? sD[q - kTX * kTY] : hUd[jj][ii];
        const double uu = (kk > 0) ? sUD[q - kTX * kTY] : hUu[jj][ii];
        const double dinv = 1.0 / du;
        d = d - sUp[q] * dinv * uu;
      }
      if (g.y0 + jj > 0) {
        const double dn = (jj > 0) ? sD[q - kTX] : hNd[kk][ii];
        const double un = (jj > 0) ? sUS[q - kTX] : hNu[kk][ii];
        const double dinv = 1.0 / dn;
        d = d - sN[q] * dinv * un;
      }
      if (g.x0 + ii > 0) {
        const double dw = (ii > 0) ? sD[q - 1] : hWd[kk][jj];
        const double uw = (ii > 0) ? sUE[q - 1] : hWu[kk][jj];
        const double dinv = 1.0 / dw;
        d = d - sW[q] * dinv * uw;
      }
      sD[q] = d;
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < kTC; q += kTT) {
    const int ii = q % kTX, jj2 = (q / kTX) % kTY, kk2 = q / (kTX * kTY);
    if (g.inside(dm, ii, jj2, kk2)) dg[g.cell(dm, ii, jj2, kk2)] = sD[q];
  }
}

__global__ void __launch_bounds__(kTT)
k_ilu_factor(Dims dm, const Tile *__restrict__ tiles, Mat7 A, double *__restrict__ dg) {
  ilu_factor_tile(dm, tiles[blockIdx.x], A, dg);
}

// The same sweeps in one launch.  Blocks take tiles from a ticket counter in
// the order of the tile list (diagonal by diagonal; reversed for the backward
// sweep), so every tile a block waits for was taken earlier by a block that
// is running: the waits end whatever the order in which blocks are
// dispatched.  A finished tile is published by writing the sweep's epoch to
// its flag (every thread's stores fenced at agent scope first, then one
// release store); a tile's predecessors are awaited with acquire loads.  A
// wait is bounded (spin_max polls): past it the block raises ctl[1] and
// every block stops at its next wait, so the launch always drains; the host
// reports the flag as a device error.
enum IluKind { kIluFactor = 0, kIluFwd = 1, kIluBwd = 2 };

__device__ __forceinline__ bool tile_wait(const unsigned *flag, unsigned epoch,
                                          const unsigned *err, int spin_max) {
  for (int k = 0; k < spin_max; k++) {
    if (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == epoch) return true;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return false;
}

template <int KIND>
__global__ void __launch_bounds__(kTT)
k_ilu_wave(Dims dm, const Tile *__restrict__ tiles, int ntiles, int ni, int nj, int nk, Mat7 A,
           double *__restrict__ dg, const double *__restrict__ r, double *__restrict__ out,
           const int *__restrict__ skip, unsigned *ctl, unsigned *flags, unsigned epoch,
           int spin_max) {
  if (skip && *skip) return;
  __shared__ int s_t, s_ok;
  // The ticket and the outcome of the waits are made wave-uniform
  // (readfirstlane): the loop's exits must be uniform branches, or the
  // compiler structures the loop as divergent and the lanes beside thread 0
  // would come round to the barrier without thread 0's new ticket.  A block
  // takes at most ntiles + 1 tickets.
  for (int turn = 0; turn <= ntiles; turn++) {
    if (threadIdx.x == 0) s_t = (int)atomicAdd(&ctl[0], 1u);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_t);
    if (t >= ntiles) return;
    const Tile tl = tiles[KIND == kIluBwd ? ntiles - 1 - t : t];
    if (threadIdx.x == 0) {
      // predecessors: the tiles before this one along x, y and z (after it
      // for the backward sweep)
      const int d = KIND == kIluBwd ? 1 : -1;
      const int I = tl.I + d, J = tl.J + d, K = tl.K + d;
      bool ok = true;
      if (I >= 0 && I < ni)
        ok = ok && tile_wait(&flags[(tl.K * nj + tl.J) * ni + I], epoch, &ctl[1], spin_max);
      if (ok && J >= 0 && J < nj)
        ok = tile_wait(&flags[(tl.K * nj + J) * ni + tl.I], epoch, &ctl[1], spin_max);
      if (ok && K >= 0 && K < nk)
        ok = tile_wait(&flags[(K * nj + tl.J) * ni + tl.I], epoch, &ctl[1], spin_max);
      if (!ok) __hip_atomic_store(&ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_ok = ok;
    }
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_ok)) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if constexpr (KIND == kIluFactor) ilu_factor_tile(dm, tl, A, dg);
    else if constexpr (KIND == kIluFwd) ilu_fwd_tile(dm, tl, A, dg, r, out);
    else ilu_bwd_tile(dm, tl, A, dg, out);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(&flags[(tl.K * nj + tl.J) * ni + tl.I], epoch, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_ilu_wave_reset(unsigned *ctl) {
  if (threadIdx.x == 0) ctl[0] = 0u;
}

// ===========================================================================
// FP32 vertical-line preconditioner (the mixed-precision path).  The
// Jacobian of a layered box is dominated by its vertical couplings (face
// area / distance is (dx dy / dz) vertically vs (dy dz / dx) horizontally:
// 400:1 at 5 m x 5 m x 0.25 m cells), so M = blockdiag(T_col), the
// tridiagonal of every vertical column, captures most of A.  Factor: Thomas
// pivots computed in FP64 and stored in FP32 (1/pivot, c'); apply: one
// thread per column, forward elimination into LDS, back substitution, FP32
// arithmetic, FP64 in and out.  Right-preconditioned BiCGStab uses the
// computed z explicitly (x += alpha z, v = A z), so the FP32 rounding of M
// does not enter the FP64 residual recurrence.
constexpr int kCol = 64;  // columns per block (one wave)

// za: the up coupling in FP32 (null where A is FP32: the apply reads A.up)
template <typename T>
__global__ void __launch_bounds__(kCol)
k_zline_factor(Dims dm, Mat7T<T> A, float *__restrict__ zinv, float *__restrict__ zcp,
               float *__restrict__ za) {
  const int t = blockIdx.x * kCol + threadIdx.x;
  if (t >= dm.nxy) return;
  // One thread per column is one wave per SIMD at 256x256 columns: the
  // coefficients of eight levels are loaded before their eliminations, so a
  // column pays one load latency per eight levels.  The up coupling of the
  // top level and the down coupling of the bottom level are stored zeros.
  double cp = 0.0;
  int k = 0;
  for (; k + 8 <= dm.nz; k += 8) {
    double dd[8], uu[8], nn[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int c = (k + u) * dm.nxy + t;
      dd[u] = A.d[c];
      uu[u] = A.up[c];
      nn[u] = A.dn[c];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int c = (k + u) * dm.nxy + t;
      const double piv = dd[u] - uu[u] * cp;
      const double inv = 1.0 / piv;
      cp = nn[u] * inv;
      zinv[c] = (float)inv;
      zcp[c] = (float)cp;
      if (za) za[c] = (float)uu[u];
    }
  }
  for (; k < dm.nz; k++) {
    const int c = k * dm.nxy + t;
    const double a = A.up[c];
    const double piv = A.d[c] - a * cp;
    const double inv = 1.0 / piv;
    cp = A.dn[c] * inv;
    zinv[c] = (float)inv;
    zcp[c] = (float)cp;
    if (za) za[c] = (float)a;
  }
}

__global__ void __launch_bounds__(kCol)
k_zline_apply(Dims dm, const float *__restrict__ zinv, const float *__restrict__ zcp,
              const float *__restrict__ za, const double *__restrict__ r,
              double *__restrict__ z, const int *__restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ float sdp[];  // [nz][kCol]
  const int t = blockIdx.x * kCol + threadIdx.x;
  if (t >= dm.nxy) return;
  float dp = 0.0f;
  int k = 0;
  for (; k + 4 <= dm.nz; k += 4) {
    float rr[4], aa[4], ii[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int c = (k + u) * dm.nxy + t;
      rr[u] = (float)r[c];
      aa[u] = za[c];
      ii[u] = zinv[c];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      dp = (rr[u] - aa[u] * dp) * ii[u];
      sdp[(k + u) * kCol + threadIdx.x] = dp;
    }
  }
  for (; k < dm.nz; k++) {
    const int c = k * dm.nxy + t;
    dp = ((float)r[c] - za[c] * dp) * zinv[c];
    sdp[k * kCol + threadIdx.x] = dp;
  }
  float x = 0.0f;
  k = dm.nz - 1;
  for (; k >= 3; k -= 4) {
    float cc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) cc[u] = zcp[(k - u) * dm.nxy + t];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      x = sdp[(k - u) * kCol + threadIdx.x] - cc[u] * x;
      z[(k - u) * dm.nxy + t] = (double)x;
    }
  }
  for (; k >= 0; k--) {
    x = sdp[k * kCol + threadIdx.x] - zcp[k * dm.nxy + t] * x;
    z[k * dm.nxy + t] = (double)x;
  }
}

// ===========================================================================
// FP32 multicolour ILU(0) (MSG123_PRECON_MCILU32).  Cells are red (x+y+z
// even) or black (odd).  In the red-first ordering the 7-point matrix is
// [D_r C; E D_b] with diagonal D_r and D_b, and ILU(0) on that pattern is
// M = [D_r 0; E S] [I D_r^-1 C; 0 I] with S = D_b - diag(E D_r^-1 C):
//   red pivot    d_r = a_rr,
//   black pivot  d_b = a_bb - sum_k a_bk a_kb / a_kk   (k: its red neighbours,
//                in the order up, N, W, E, S, down).
// z = M^-1 r takes two passes, each over one colour with no dependency
// inside it:
//   black  z_b = (r_b - sum_k a_bk (r_k / a_kk)) / d_b,
//   red    z_r = (r_r - sum_k a_rk z_k) / a_rr.
// Pivots are computed in FP64; the factors (1/pivot and the six couplings)
// are stored and applied in FP32.  Cells 2p and 2p+1 always differ in
// colour (with an even row length a row starts at an even index; with an odd
// one the colour alternates with the index inside each level, and a level of
// odd size makes it alternate across levels too), so pair p holds one cell
// of each colour and the factors are packed per colour by p:
// mc[(colour * 7 + slot) * np + p], slot 0 = 1/pivot, slots 1..6 = couplings.

__device__ __forceinline__ int mc_colour(const Dims &dm, int c) {
  return (c % dm.nx + (c / dm.nx) % dm.ny + c / dm.nxy) & 1;
}

// The six neighbours of cell c in CSR order (up, N, W, E, S, down); a
// missing one is reported as c itself with ok = false.
__device__ __forceinline__ void mc_nbrs(const Dims &dm, int c, int *k, bool *ok) {
  const int x = c % dm.nx, y = (c / dm.nx) % dm.ny, z = c / dm.nxy;
  ok[0] = z > 0;
  ok[1] = y > 0;
  ok[2] = x > 0;
  ok[3] = x < dm.nx - 1;
  ok[4] = y < dm.ny - 1;
  ok[5] = z < dm.nz - 1;
  k[0] = ok[0] ? c - dm.nxy : c;
  k[1] = ok[1] ? c - dm.nx : c;
  k[2] = ok[2] ? c - 1 : c;
  k[3] = ok[3] ? c + 1 : c;
  k[4] = ok[4] ? c + dm.nx : c;
  k[5] = ok[5] ? c + dm.nxy : c;
}

__global__ void __launch_bounds__(kBlock)
k_mc_factor(Dims dm, Mat7 A, int np, float *__restrict__ mc, double *__restrict__ piv_out) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const size_t P = (size_t)np;
  const int col0 = mc_colour(dm, 2 * p);
  for (int h = 0; h < 2; h++) {
    const int c = 2 * p + h;
    const int col = col0 ^ h;
    float *out = mc + (size_t)col * 7 * P + p;
    if (c >= dm.n) {  // the last pair of an odd cell count has one cell
      for (int s = 0; s < 7; s++) out[s * P] = 0.0f;
      continue;
    }
    const double *off[6] = {A.up, A.n, A.w, A.e, A.s, A.dn};
    double piv = A.d[c];
    if (col) {
      // a_kb: the neighbour's coupling back to c is its opposite direction
      const double *back[6] = {A.dn, A.s, A.e, A.w, A.n, A.up};
      int k[6];
      bool ok[6];
      mc_nbrs(dm, c, k, ok);
      for (int s = 0; s < 6; s++) {
        const double t = off[s][c] * back[s][k[s]] / A.d[k[s]];
        piv = piv - (ok[s] ? t : 0.0);
      }
    }
    piv_out[c] = piv;
    out[0] = (float)(1.0 / piv);
    for (int s = 0; s < 6; s++) out[(s + 1) * P] = (float)off[s][c];
  }
}

// black pass: z_b = (r_b - sum_k a_bk (r_k / a_kk)) / d_b
__global__ void __launch_bounds__(kBlock)
k_mc_black(Dims dm, int np, const float *__restrict__ mc, const double *__restrict__ r,
           double *__restrict__ z, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const int b = 2 * p + (mc_colour(dm, 2 * p) ? 0 : 1);
  if (b >= dm.n) return;
  const size_t P = (size_t)np;
  const float *B = mc + 7 * P, *R = mc;
  int k[6];
  bool ok[6];
  mc_nbrs(dm, b, k, ok);
  float rk[6], ik[6], a[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    rk[s] = (float)r[k[s]];
    ik[s] = R[k[s] >> 1];
    a[s] = B[(s + 1) * P + p];
  }
  float t = (float)r[b];
#pragma unroll
  for (int s = 0; s < 6; s++) t = t - (ok[s] ? a[s] * (rk[s] * ik[s]) : 0.0f);
  z[b] = (double)(t * B[p]);
}

// red pass: z_r = (r_r - sum_k a_rk z_k) / a_rr, z_k from the black pass
__global__ void __launch_bounds__(kBlock)
k_mc_red(Dims dm, int np, const float *__restrict__ mc, const double *__restrict__ r,
         double *__restrict__ z, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const int rc = 2 * p + (mc_colour(dm, 2 * p) ? 1 : 0);
  if (rc >= dm.n) return;
  const size_t P = (size_t)np;
  const float *R = mc;
  int k[6];
  bool ok[6];
  mc_nbrs(dm, rc, k, ok);
  float zk[6], a[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    zk[s] = ok[s] ? (float)z[k[s]] : 0.0f;
    a[s] = R[(s + 1) * P + p];
  }
  float t = (float)r[rc];
#pragma unroll
  for (int s = 0; s < 6; s++) t = t - (ok[s] ? a[s] * zk[s] : 0.0f);
  z[rc] = (double)(t * R[p]);
}

// ===========================================================================
// Host side

inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }

// delt, and 1 / delt in device memory for the kernels (a launch on the
// context's stream, so kernels enqueued before it keep the previous value)
int set_delt(msg123_ctx *c, double delt);

// dst_k = src_k (n_k doubles each; n_k = 0: unused) on the context's stream
int dcopy(msg123_ctx *c, double *d0, const double *s0, size_t n0, double *d1 = nullptr,
          const double *s1 = nullptr, size_t n1 = 0, double *d2 = nullptr,
          const double *s2 = nullptr, size_t n2 = 0);

struct Timer {
  double total_ms = 0.0;
  int64_t launches = 0;
};

struct PendingEvent {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

struct msg123_ctx {
  Dims dm{};
  int device = 0;  // the HIP device current when the context was created
  int ns = 0, nconn = 0, nb = 0;
  FusedShape fshape[2];  // launch variant of the fused residual [0] and assembly [1] kernels
  Tiling ftg[2]{};       // ... and their tilings
  Geo geo{};     // face geometry, volumes and cell tops, per layer or per cell
  double *geo_tab = nullptr;  // per-layer tables [5][nz] when the geometry is layered
  bool aux_stale = false;     // rel_perm / sst not of head_new (refresh_aux)
  int32_t *probe_idx = nullptr;  // msg123_run_box probe cells and their gathered heads
  double *probe_val = nullptr;
  int n_probe_cap = 0;
  // device-driven Newton iterations: the control record in device memory,
  // its mapped host mirror, and the skip flag the Krylov kernels read (null
  // in host-driven solves)
  struct DevNewton *dn = nullptr, *dn_host = nullptr, *dn_map = nullptr;
  const int *skip = nullptr;
  int control = MSG123_CONTROL_HOST;
  int dev_chunk = 1;  // BiCGStab iterations enqueued before the first wait
  double len_scal = 1.0, newper = 0.0, newper_inv = 0.0;
  msg123_params prm{};
  hipStream_t stream = nullptr;
  // CSR pattern kept on the host for layout conversions
  std::vector<int32_t> offind, offrow, dir;
  // device arrays
  double *cell_top = nullptr, *cell_vol = nullptr;
  float *soil = nullptr;  // 5 x N: reta, retn, resi, poro, ss
  double *faces = nullptr;  // 6 x N: hE aE hS aS hD aD
  double *drain = nullptr;  // nx*ny free-drainage bottom conductances (when set)
  bool drain_on = false;
  // river, lake, well and sea terms (msg123_set_bterms)
  double *rl = nullptr;        // [6][ns]: river tran, stage, bottom; lake likewise
  bool rl_on = false;
  int *rl_cols = nullptr;      // columns with a river or a lake
  int n_rl_cols = 0;
  int *bc_cell = nullptr, *bc_beg = nullptr;  // well and sea cells, sea ranges
  double *bc_well = nullptr;   // calc_well of each listed cell (0: sea only)
  double *bc_sea = nullptr;    // [3][bc_nsea]: hydf_seal, abyd_seal, sea level
  int bc_n = 0, bc_nsea = 0;
  double *f_bnd = nullptr;     // F at the perturbed head (J.v with well/sea cells)
  double *bm = nullptr;        // [5] k_bmass results
  double bm_last[5] = {0, 0, 0, 0, 0};  // of the last msg123_calc_mass
  std::vector<double> hydf_surf_h;  // host copy of hydf_surf (river/lake tran)
  double *surf = nullptr;   // 4 x ns: hydf abyd top bott
  double *surf_head = nullptr, *surf_rati = nullptr, *calc_rech = nullptr, *surf_old = nullptr;  // ns
  double *h_new = nullptr, *h_old = nullptr, *h_pre = nullptr, *dh = nullptr;
  double *srat_new = nullptr, *srat_old = nullptr, *rel_perm = nullptr, *sst = nullptr;
  double *srat_p = nullptr, *kr_p = nullptr;
  double *fbuf = nullptr, *f1 = nullptr, *jvec = nullptr, *rhs = nullptr;
  double *A = nullptr;  // 7 x N
  double *kr_r = nullptr, *kr_rs = nullptr, *kr_p_ = nullptr, *kr_v = nullptr, *kr_z = nullptr,
         *kr_t = nullptr, *dg = nullptr;
  float *zinv = nullptr, *zcp = nullptr, *za = nullptr;  // FP32 line factors
  float *mc = nullptr;         // FP32 multicolour ILU(0) factors [2][7][mc_np]
  int mc_np = 0;               // cell pairs, (n + 1) / 2
  double *amg_last = nullptr;  // AMG path: last preconditioner input (level-1 rhs)
  AmgGpu *amg32 = nullptr;     // device AMG hierarchy (MSG123_PRECON_AMG32)
  int cur_precon = -1;         // preconditioner of the running solve (-1: prm.precon)
  bool backtr_pending = false; // run_backtr's sums were enqueued and waited for already
  std::map<std::string, hipGraphExec_t> graphs;  // MSG123_CONTROL_GRAPH: captured chunks
  // The next Newton iteration's system assembled ahead, at the full step's
  // head, in place of the line search's F(head_new) (rhs2 = -F, A2): used
  // when the full step is kept, exchanged with rhs and A.
  double *A2 = nullptr, *rhs2 = nullptr;
  bool spec_ready = false;
  unsigned spec_r0_seq = 0;
  int *conv_run = nullptr;  // [3] the full step converged: 0 1, else 1 0 (k_tail_final);
                            // [2] the Krylov solve ended at the half-step (k_finalize_gate)
  // The Newton systems of the line-preconditioned path in FP32 (Af, and A2f
  // for the one assembled ahead); sys32: the loaded system is Af's
  float *Af = nullptr, *A2f = nullptr;
  bool sys32 = false;
  const float *za_cur = nullptr;  // the up coupling the line apply reads
  // the last calc_numsol converged at a full step whose S and kr at head_new
  // k_conv_state wrote (srat_new, kr_p); read by run_box right after it
  bool conv_state_fresh = false;
  bool dh_unset = false;  // dh is the solve's x0 = 0 but not yet written
  // calc_numsol's tail launched behind each first half-step of the solve
  // (tail_ahead), and valid when the solve ended there (tail_seq: its fold)
  bool tail_ahead = false, tail_valid = false;
  unsigned tail_seq = 0;
  double eta_first = 0.5, eta_max = 0.9;  // Eisenstat-Walker forcing (the reference's values)
  int law = MSG123_LAW_REFERENCE;  // form of the van Genuchten-Mualem evaluation
  int algebraic_slope = 0;     // line-search slope from the Krylov residual instead of F(x+eps v)
  int beta_refine = 1;         // the line search's beta-condition refinement (the reference's)
  // the system just assembled is solved from x0 = 0: its |rhs|^2 is being
  // folded into result slot 8 (sequence r0_seq); r0_value once read
  bool r0_pending = false;
  unsigned r0_seq = 0;
  double r0_value = NAN;
  int amg_setups = 0, amg_fallbacks = 0;
  double *dh0 = nullptr;       // AMG32: the solve's initial guess, kept for a restart
  double *mass_acc = nullptr;  // [24] the run's cumulative mass balance on the device
  int mass_steps = 0;          // steps accumulated into mass_acc
  bool mass_on = false;        // accumulate the mass balance of msg123_run_box's steps
  msg123_amg_opts amg_opt{12, 0.05f, 0.67f, 0.67f, 2, 2, 64, 20, 1, 1};
  // warm start: previous accepted head, candidates, network inputs
  double *h_prev = nullptr, *h_lin = nullptr, *h_try = nullptr, *net_scale = nullptr;
  uint16_t *net_x = nullptr;
  float *lnks = nullptr;
  msg123_net *net = nullptr;
  double rech_ms = 0.0;
  msg123_warm_stats wstats{};
  double *h_log = nullptr;     // mapped pinned: the convergence-log values of one iteration
  double *log_map = nullptr;   // its device view
  std::vector<msg123_iter_rec> iters;  // per Newton iteration of the last calc_numsol
  int end_reason = MSG123_END_CONVERGED;
  double *partials = nullptr, *h_scal = nullptr;
  double *scal = nullptr;  // device view of h_scal (mapped pinned memory)
  double *scal_dev = nullptr;  // the same results in device memory, for kernels to read
  double *zero = nullptr;      // 32 zeros (Geo::zero)
  unsigned *h_flag = nullptr, *flag_map = nullptr;  // reduction-done flag (mapped) and its view
  unsigned flag_seq = 0;
  MaxRec *maxrec = nullptr;
  MaxRec *h_maxrec = nullptr;   // one record, mapped pinned memory
  MaxRec *maxrec_map = nullptr;  // its device view
  double *h_part = nullptr;
  Tile *tiles = nullptr;
  std::vector<int> diag_off, diag_cnt;
  // single-launch sweeps (k_ilu_wave): tiles per axis, ticket counter and
  // error flag [2], per-tile epoch flags, the last epoch used
  int tni = 0, tnj = 0, tnk = 0, ntiles = 0;
  unsigned *wave_ctl = nullptr, *wave_flags = nullptr;
  bool wave_ran = false;  // k_ilu_wave launched since the last wave_check
  unsigned wave_epoch = 0;
  int ilu_launch = MSG123_ILU_DIAGONALS;
  int num_cu = 256;
  double delt = 0.0;
  FILE *conv_log = nullptr;   // reference-format convergence log, when enabled
  int last_maxidx = -1;
  double log_time = 0.0;
  int timing = 0;  // 0 off, 1 every kernel group, 2 the SpMV group only
  std::map<std::string, Timer> timers;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<PendingEvent> pending;   // recorded, not yet resolved
  std::vector<hipEvent_t> event_pool;  // free events
  // time-loop state of msg123_run_box (Time_module.f90 module variables)
  struct {
    int started = 0, conv_flag = 0, timestep_num = 0, out_iter = 0, sk = 0, have_prev = 0;
    int retry_persist = 0;  // the learned start of the last attempt failed: same dt, persistence
    int shadow = 0;         // ... or was a trial: its result is discarded, same dt, persistence
    // learned start chosen by the Newton iterations it costs on trial
    // steps solved from both starts: completed trials, their Newton
    // iterations from the network's start and from persistence, and the
    // decision (-1 while trials run, 0 persistence, 1 network)
    int trials = 0, trial_net = 0, trial_per = 0, net_on = -1;
    int per_prev = 0;  // Newton iterations of the last step accepted from persistence
    float current_t = 0.0f, t_start = 0.0f;
    double delt = 0.0, delt_old1 = 0.0;
    // recharge records: the record in effect, st_step_flag%rech (set when
    // the planned step ends at the record's end) and st_rech%etime
    int rk = 0, rflag = 0;
    float etime = 0.0f;
  } tl;
};

namespace {

// the constitutive-law kernel of the context's form
using VgKernel = decltype(&k_vg<false>);
VgKernel vg_kernel(const msg123_ctx *c) {
  return c->law == MSG123_LAW_LOGEXP ? &k_vg<true> : &k_vg<false>;
}

Mat7F mat7f(msg123_ctx *c) {
  const size_t N = (size_t)c->dm.n;
  const float *A = c->Af;
  return Mat7F{A, A + N, A + 2 * N, A + 3 * N, A + 4 * N, A + 5 * N, A + 6 * N};
}

Mat7 mat7(msg123_ctx *c) {
  const size_t N = (size_t)c->dm.n;
  double *A = c->A;
  return Mat7{A, A + N, A + 2 * N, A + 3 * N, A + 4 * N, A + 5 * N, A + 6 * N};
}
Faces faces(msg123_ctx *c) {
  const size_t N = (size_t)c->dm.n;
  double *f = c->faces;
  return Faces{f, f + N, f + 2 * N, f + 3 * N, f + 4 * N, f + 5 * N,
               c->drain_on ? c->drain : nullptr};
}
Soil soil(msg123_ctx *c) {
  const size_t N = (size_t)c->dm.n;
  float *s = c->soil;
  return Soil{s, s + N, s + 2 * N, s + 3 * N, s + 4 * N};
}
Surface surface(msg123_ctx *c) {
  const size_t S = (size_t)c->ns;
  double *s = c->surf;
  return Surface{s,           s + S,         s + 2 * S,    s + 3 * S,
                 c->surf_head, c->surf_rati, c->calc_rech, c->rl_on ? c->rl : nullptr};
}
BCells bcells(msg123_ctx *c) {
  const size_t M = (size_t)c->bc_nsea;
  return BCells{c->bc_n, c->bc_cell, c->bc_well, c->bc_beg, c->bc_sea, c->bc_sea + M,
                c->bc_sea + 2 * M};
}
bool has_bterms(const msg123_ctx *c) { return c->rl_on || c->bc_n > 0; }
Storage storage(msg123_ctx *c) {
  return Storage{c->h_old, c->srat_old, c->cell_vol, c->soil + 3 * (size_t)c->dm.n,
                 c->scal_dev + kDeltSlot};
}

// Timed launch helper: when timing is enabled, records a pair of HIP events
// on the context's stream around a kernel group; the pairs are resolved
// (host-synchronised) only when the timings are read, so enabling timing
// adds no synchronisation to the timed region.
hipEvent_t take_event(msg123_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void resolve_events(msg123_ctx *c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &p : c->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    Timer &t = c->timers[p.name];
    t.total_ms += ms;
    t.launches += 1;
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
}

struct Scope {
  msg123_ctx *c;
  const char *name;
  hipEvent_t a = nullptr;
  Scope(msg123_ctx *cc, const char *nm) : c(cc), name(nm) {
    if (c->timing == 1 || (c->timing == 2 && std::strcmp(nm, "spmv") == 0)) {
      a = take_event(c);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~Scope() {
    if (a) {
      hipEvent_t b = take_event(c);
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back(PendingEvent{name, a, b});
      if (c->pending.size() > 4096) resolve_events(c);
    }
  }
};

// The last timed launch of group `from`, still pending, did no work (its
// kernels found their run flag off): it is booked under `to` instead.
void retime_last(msg123_ctx *c, const char *from, const char *to) {
  for (auto it = c->pending.rbegin(); it != c->pending.rend(); ++it)
    if (it->name == from) {
      it->name = to;
      return;
    }
}

// Wait until a kernel published `seq` (publish()): a spin on the mapped
// flag, which returns as soon as the results are visible, without waiting
// for the stream's completion signal.  Every 4096 spins the stream is
// queried: an error ends the wait, and a stream that has drained without
// the flag is reported as a device error instead of spinning forever.
int wait_flag(msg123_ctx *c, unsigned seq) {
  volatile unsigned *f = c->h_flag;
  for (unsigned spin = 1;; spin++) {
    if ((int)(*f - seq) >= 0) break;  // this fold or a later one has published
    if ((spin & 4095u) == 0) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e == hipSuccess) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if ((int)(*f - seq) < 0) return MSG123_EDEVICE;
        break;
      }
      if (e != hipErrorNotReady) return MSG123_EDEVICE;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return MSG123_OK;
}

// Fold nv partial sums into slots off.. of the host-mapped results and of
// their device copy; returns the sequence number the fold publishes.
int reduce_launch(msg123_ctx *c, int nslots, int nv, int off, unsigned *seq) {
  if (nv < 1 || nv > kMaxRed || off < 0 || off + nv > kScal) return MSG123_EINVAL;
  *seq = ++c->flag_seq;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFin), 0, c->stream, c->partials, nslots, nv,
                     c->scal + off, c->scal_dev + off, c->flag_map, *seq);
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}

int reduce_finalize(msg123_ctx *c, int nslots, int nv, double *host_out) {
  unsigned seq;
  if (int rc = reduce_launch(c, nslots, nv, 0, &seq)) return rc;
  if (int rc = wait_flag(c, seq)) return rc;
  for (int k = 0; k < nv; k++) host_out[k] = c->h_scal[k];
  return MSG123_OK;
}

int eff_precon(msg123_ctx *c) { return c->cur_precon >= 0 ? c->cur_precon : c->prm.precon; }
bool use_line32(msg123_ctx *c) { return eff_precon(c) == MSG123_PRECON_LINE32; }
bool use_amg(msg123_ctx *c) { return eff_precon(c) == MSG123_PRECON_AMG; }
bool use_amg32(msg123_ctx *c) { return eff_precon(c) == MSG123_PRECON_AMG32; }
bool use_mc32(msg123_ctx *c) { return eff_precon(c) == MSG123_PRECON_MCILU32; }
// the diagonal the convergence log reports (the AMG path: the D-ILU one)
DiagLog diag_log(msg123_ctx *c) {
  if (use_amg(c)) return DiagLog{c->dg, nullptr};
  return c->sys32 ? DiagLog{nullptr, c->Af} : DiagLog{c->A, nullptr};
}
// Newton systems in FP32: the line preconditioner's path (its Jacobian only
// steers Newton, and the FD Jacobian holds about 8 digits)
bool a32_on(msg123_ctx *c) { return use_line32(c); }
int alloc_a32(msg123_ctx *c) {
  if (c->Af) return MSG123_OK;
  const size_t N = (size_t)c->dm.n;
  if (hipMalloc((void **)&c->Af, sizeof(float) * 7 * N) != hipSuccess) return MSG123_ENOMEM;
  return MSG123_OK;
}

// The fused stencil kernel of the context's law form, in the launch shape
// chosen for its kind (residual or assembly; msg123_autotune).  The kernels
// do not write rel_perm and sst (S, kr, Ss stay on chip): readers of those
// arrays refresh them first (refresh_aux).
template <int MODE, bool FAST>
using FusedFn = decltype(&k_fused<MODE, FAST, true, kFTX, kFTY, 1>);

template <int MODE, bool FAST, bool LAY>
FusedFn<MODE, FAST> fused_fn_lay(int minw_idx) {
  switch (minw_idx) {
    case 1: return &k_fused<MODE, FAST, LAY, kFTX, kFTY, 4>;
    case 2: return &k_fused<MODE, FAST, LAY, kFTX, kFTY, 5>;
    default: return &k_fused<MODE, FAST, LAY, kFTX, kFTY, 1>;
  }
}

template <int MODE, bool FAST>
FusedFn<MODE, FAST> fused_fn(bool layered, int minw_idx) {
  return layered ? fused_fn_lay<MODE, FAST, true>(minw_idx)
                 : fused_fn_lay<MODE, FAST, false>(minw_idx);
}

template <int MODE>
int launch_fused_shape(msg123_ctx *c, const HeadIn &hin, const FusedOut &o, int minw_idx,
                       const Tiling &tg) {
  if constexpr (MODE == kJv) {
    if (c->bc_n > 0) {
      // the well and sea cells enter F after the fused kernel: F at the
      // perturbed head with them, then J.v and rhs.jv from it (per block of
      // the residual's slot count, as the fused J.v kernel leaves them)
      FusedOut r = o;
      r.f = c->f_bnd;
      r.partials = nullptr;
      if (int rc = launch_fused_shape<kResid>(c, hin, r, minw_idx, tg)) return rc;
      hipLaunchKernelGGL(k_jv_from_f, dim3(tg.nblocks), dim3(kBlock), 0, c->stream, c->dm.n,
                         c->f_bnd, o.f1, o.eps_inv, o.eps_inv_dev, o.rhs, o.f, o.partials,
                         o.run_if);
      HIPCHK(hipGetLastError());
      return MSG123_OK;
    }
  }
  const bool fast = c->law == MSG123_LAW_LOGEXP;
  constexpr int NT = kFTX * kFTY + 2 * kFTX + 2 * kFTY;
  const bool lay = c->geo.layered != 0;
  const FusedFn<MODE, true> kf = fused_fn<MODE, true>(lay, minw_idx);
  const FusedFn<MODE, false> kr = fused_fn<MODE, false>(lay, minw_idx);
  const size_t lds = lay ? fused_geo_lds(tg) : 0;
  if (fast)
    hipLaunchKernelGGL(kf, dim3(tg.nblocks), dim3(NT), lds, c->stream, c->dm, tg, hin, soil(c),
                       c->geo, faces(c), storage(c), surface(c), c->newper, c->newper_inv,
                       c->prm.consistent_jacobian, o);
  else
    hipLaunchKernelGGL(kr, dim3(tg.nblocks), dim3(NT), lds, c->stream, c->dm, tg, hin, soil(c),
                       c->geo, faces(c), storage(c), surface(c), c->newper, c->newper_inv,
                       c->prm.consistent_jacobian, o);
  HIPCHK(hipGetLastError());
  if constexpr (MODE != kJv) {
    if (c->bc_n > 0) {
      if (fast)
        hipLaunchKernelGGL((k_bcells<MODE, true>), dim3(1), dim3(kBcT), 0, c->stream, bcells(c),
                           hin, c->cell_top, soil(c), c->newper, c->newper_inv, o.f, o.A,
                           o.Af, o.partials, o.run_if);
      else
        hipLaunchKernelGGL((k_bcells<MODE, false>), dim3(1), dim3(kBcT), 0, c->stream,
                           bcells(c), hin, c->cell_top, soil(c), c->newper, c->newper_inv, o.f,
                           o.A, o.Af, o.partials, o.run_if);
      HIPCHK(hipGetLastError());
    }
  }
  c->aux_stale = true;
  return MSG123_OK;
}

template <int MODE>
int launch_fused(msg123_ctx *c, const HeadIn &hin, const FusedOut &o, const char *group) {
  Scope sc(c, group);
  const int kind = MODE == kAssem ? 1 : 0;
  return launch_fused_shape<MODE>(c, hin, o, c->fshape[kind].minw_idx, c->ftg[kind]);
}

// rel_perm and sst of head_new (the reference's calc_func side effects),
// recomputed when a fused kernel left them stale
int refresh_aux(msg123_ctx *c) {
  if (!c->aux_stale) return MSG123_OK;
  hipLaunchKernelGGL(vg_kernel(c), dim3(c->nb), dim3(kBlock), 0, c->stream, c->dm, c->nb, c->h_new,
                     nullptr, nullptr, 0.0, nullptr, c->cell_top, soil(c), 0.0, c->newper, 0,
                     c->srat_p, c->rel_perm, c->sst, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  c->aux_stale = false;
  return MSG123_OK;
}

// F(h_in) into `out` (calc_func, Calc_function.f90:24-101), its norm folded
// into result slot `off` without waiting for it (*seq: the fold's sequence
// number)
int residual_launch(msg123_ctx *c, const double *h_in, double *out, int off, unsigned *seq) {
  const HeadIn hin{h_in, nullptr, nullptr, 0.0, nullptr};
  const FusedOut o{out, nullptr, c->srat_new, nullptr, 0.0, nullptr, c->partials};
  if (int rc = launch_fused<kResid>(c, hin, o, "residual")) return rc;
  return reduce_launch(c, c->ftg[0].nblocks, 1, off, seq);
}

// F(h) into `out`, h = h_in or (from_pre) h_pre + lam*dh written to h_new;
// returns ||F||^2 in *red when red != null.  With f1: calc_vecjacf's
// J.v = -(F(h_in) - f1)*eps_inv (Calc_function.f90:515-621) into `out` and
// rhs.jv in *red.
int residual(msg123_ctx *c, const double *h_in, bool from_pre, double lam, double *out,
             const double *f1, double eps_inv, double *red) {
  const HeadIn hin = from_pre ? HeadIn{nullptr, c->h_pre, c->dh, lam, c->h_new}
                              : HeadIn{h_in, nullptr, nullptr, 0.0, nullptr};
  const FusedOut o{out, nullptr, c->srat_new, f1, eps_inv,
                   (f1 && use_amg(c)) ? c->amg_last : c->rhs, red ? c->partials : nullptr};
  int rc = f1 ? launch_fused<kJv>(c, hin, o, "residual") : launch_fused<kResid>(c, hin, o, "residual");
  if (rc) return rc;
  if (red) return reduce_finalize(c, c->ftg[0].nblocks, 1, red);
  return MSG123_OK;
}

// make_matvec on head_new (Make_linearsystem.f90:24-68); returns ||rhs||^2.
// With for_newton (calc_numsol) head_pre becomes head_new by exchanging the
// two buffers (the Newton update writes every cell of head_new before it is
// read again), and ||rhs||^2 is folded into result slot 8 without a wait:
// the solve from x0 = 0 that follows reads it there.
int assemble(msg123_ctx *c, double *rhs_norm2, bool for_newton = false) {
  // calc_numsol's systems on the line path in FP32, every other one in FP64
  const bool f32 = for_newton && a32_on(c);
  if (f32) {
    if (int rc = alloc_a32(c)) return rc;
  }
  const HeadIn hin{c->h_new, nullptr, nullptr, 0.0, nullptr};
  const FusedOut o{c->rhs, f32 ? nullptr : c->A, c->srat_new, nullptr, 0.0, nullptr, c->partials,
                   nullptr, nullptr, f32 ? c->Af : nullptr};
  if (int rc = launch_fused<kAssem>(c, hin, o, "assemble")) return rc;
  c->sys32 = f32;
  if (!for_newton) return reduce_finalize(c, c->ftg[1].nblocks, 1, rhs_norm2);
  std::swap(c->h_pre, c->h_new);
  c->r0_pending = true;
  return reduce_launch(c, c->ftg[1].nblocks, 1, 8, &c->r0_seq);
}

constexpr int kWaveSpin = 1 << 22;  // polls of a tile wait before it gives up

// after a synchronisation: a single-launch sweep that gave up on a wait
// (k_ilu_wave's error flag) is a device error; the flag is cleared
int wave_check(msg123_ctx *c) {
  if (!c->wave_ctl || !c->wave_ran) return MSG123_OK;
  c->wave_ran = false;
  unsigned e = 0;
  HIPCHK(hipMemcpy(&e, c->wave_ctl + 1, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (e == 0) return MSG123_OK;
  HIPCHK(hipMemsetAsync(c->wave_ctl + 1, 0, sizeof(unsigned), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_EDEVICE;
}

int ilu_sweep(msg123_ctx *c, int mode, const double *r, double *out) {
  const Mat7 A = mat7(c);
  const int nd = (int)c->diag_cnt.size();
  Scope sc(c, mode == 0 ? "ilu_factor" : (mode == 1 ? "ilu_fwd" : "ilu_bwd"));
  if (c->ilu_launch == MSG123_ILU_WAVE) {
    const unsigned epoch = ++c->wave_epoch;
    const int grid = std::min(c->ntiles, 2 * c->num_cu);
    c->wave_ran = true;
    hipLaunchKernelGGL(k_ilu_wave_reset, dim3(1), dim3(64), 0, c->stream, c->wave_ctl);
    if (mode == 0)
      hipLaunchKernelGGL(k_ilu_wave<kIluFactor>, dim3(grid), dim3(kTT), 0, c->stream, c->dm,
                         c->tiles, c->ntiles, c->tni, c->tnj, c->tnk, A, c->dg, r, out, c->skip,
                         c->wave_ctl, c->wave_flags, epoch, kWaveSpin);
    else if (mode == 1)
      hipLaunchKernelGGL(k_ilu_wave<kIluFwd>, dim3(grid), dim3(kTT), 0, c->stream, c->dm,
                         c->tiles, c->ntiles, c->tni, c->tnj, c->tnk, A, c->dg, r, out, c->skip,
                         c->wave_ctl, c->wave_flags, epoch, kWaveSpin);
    else
      hipLaunchKernelGGL(k_ilu_wave<kIluBwd>, dim3(grid), dim3(kTT), 0, c->stream, c->dm,
                         c->tiles, c->ntiles, c->tni, c->tnj, c->tnk, A, c->dg, r, out, c->skip,
                         c->wave_ctl, c->wave_flags, epoch, kWaveSpin);
    HIPCHK(hipGetLastError());
    return MSG123_OK;
  }
  for (int q = 0; q < nd; q++) {
    const int D = (mode == 2) ? nd - 1 - q : q;
    const Tile *tl = c->tiles + c->diag_off[D];
    const int cnt = c->diag_cnt[D];
    if (mode == 0)
      hipLaunchKernelGGL(k_ilu_factor, dim3(cnt), dim3(kTT), 0, c->stream, c->dm, tl, A, c->dg);
    else if (mode == 1)
      hipLaunchKernelGGL(k_ilu_fwd, dim3(cnt), dim3(kTT), 0, c->stream, c->dm, tl, A, c->dg, r,
                         out, c->skip);
    else
      hipLaunchKernelGGL(k_ilu_bwd, dim3(cnt), dim3(kTT), 0, c->stream, c->dm, tl, A, c->dg,
                         out, c->skip);
  }
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}


// AMG path: loop_amg copies its input into the level-1 rhs (Linear_solution.f90:525-527)
int amg_record(msg123_ctx *c, const double *in) {
  if (!use_amg(c)) return MSG123_OK;
  HIPCHK(hipMemcpyAsync(c->amg_last, in, sizeof(double) * c->dm.n, hipMemcpyDeviceToDevice,
                        c->stream));
  return MSG123_OK;
}

// preconditioner set-up for the loaded system
int precon_setup(msg123_ctx *c) {
  if (use_amg32(c)) {
    if (!c->amg32) {
      Scope sc(c, "amg_setup");
      c->amg_setups++;
      return amg_gpu_create(c->dm, c->offind.data(), c->offrow.data(), c->A, c->amg_opt,
                            c->stream, &c->amg32);
    }
    Scope sc(c, "amg_numeric");
    return amg_gpu_numeric(c->amg32, c->A, c->stream);
  }
  if (use_mc32(c)) {
    Scope sc(c, "mc_factor");
    hipLaunchKernelGGL(k_mc_factor, dim3(nblk(c->mc_np)), dim3(kBlock), 0, c->stream, c->dm,
                       mat7(c), c->mc_np, c->mc, c->dg);
    HIPCHK(hipGetLastError());
    return MSG123_OK;
  }
  if (use_line32(c)) {
    Scope sc(c, "line_factor");
    // an FP32 system's up coupling is read in place by the apply
    if (c->sys32) {
      hipLaunchKernelGGL(k_zline_factor<float>, dim3((c->dm.nxy + kCol - 1) / kCol), dim3(kCol),
                         0, c->stream, c->dm, mat7f(c), c->zinv, c->zcp, nullptr);
      c->za_cur = c->Af + (size_t)c->dm.n;
    } else {
      hipLaunchKernelGGL(k_zline_factor<double>, dim3((c->dm.nxy + kCol - 1) / kCol), dim3(kCol),
                         0, c->stream, c->dm, mat7(c), c->zinv, c->zcp, c->za);
      c->za_cur = c->za;
    }
    HIPCHK(hipGetLastError());
    return MSG123_OK;
  }
  return ilu_sweep(c, 0, nullptr, nullptr);
}

int line_apply(msg123_ctx *c, const double *r, double *z) {
  Scope sc(c, "line_apply");
  const size_t lds = sizeof(float) * (size_t)c->dm.nz * kCol;
  hipLaunchKernelGGL(k_zline_apply, dim3((c->dm.nxy + kCol - 1) / kCol), dim3(kCol), lds,
                     c->stream, c->dm, c->zinv, c->zcp, c->za_cur, r, z, c->skip);
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}

// z = M^-1 r with the current D-ILU diagonal
int ilu_apply(msg123_ctx *c, const double *r, double *z) {
  if (use_line32(c)) return line_apply(c, r, z);
  if (use_mc32(c)) {
    Scope sc(c, "mc_apply");
    hipLaunchKernelGGL(k_mc_black, dim3(nblk(c->mc_np)), dim3(kBlock), 0, c->stream, c->dm,
                       c->mc_np, c->mc, r, z, c->skip);
    hipLaunchKernelGGL(k_mc_red, dim3(nblk(c->mc_np)), dim3(kBlock), 0, c->stream, c->dm,
                       c->mc_np, c->mc, r, z, c->skip);
    HIPCHK(hipGetLastError());
    return MSG123_OK;
  }
  if (use_amg32(c)) {
    Scope sc(c, "amg_vcycle");
    return amg_gpu_apply(c->amg32, r, z, c->stream);
  }
  int rc = ilu_sweep(c, 1, r, z);
  if (rc) return rc;
  return ilu_sweep(c, 2, nullptr, z);
}

// BiCGStab of solve_bicgs (Linear_solution.f90:234-493), x in c->dh.
// *breakdown is set when a recurrence scalar turns zero or non-finite (the
// reference has no such check; its D-ILU path does not meet it).
int tail_launch(msg123_ctx *c, unsigned *seq, bool spec, int lconv, const int *gate = nullptr);

int bicgstab_run(msg123_ctx *c, double bnorm, double errtol, int maxinn, double *rnorm_out,
                 int *in_iter_out, int *conv_out, int *breakdown) {
  *breakdown = 0;
  const Dims dm = c->dm;
  const int n = dm.n, nb = c->nb, nbl = nblk(n);
  const Mat7 A = mat7(c);
  double *x = c->dh, *r = c->kr_r, *p = c->kr_p_, *v = c->kr_v, *z = c->kr_z, *t = c->kr_t;
  const double *rs = c->kr_rs;
  double sk;
  int rc;
  unsigned seq;
  // Scalars in the result slots (host-mapped and device copies): 8 rho at
  // the start, 9 v.r-hat, 10 |s|^2, 11-12 t.s and t.t, 13-14 |r|^2 and the
  // next rho.  The update kernels read their coefficient from the device
  // copies, so each half-step needs one wait, for its convergence test.
  // r0 is read in place where it lies: the first x/r update reads r_src
  // and writes r, and the first search direction p = r0 is r-hat (rs), read
  // only, until the first p update writes p
  const double *r_src = r;
  // x0 = 0 left unwritten by calc_numsol (dh_unset): the first update writes
  // x without reading it; a solve that reads x first clears it here
  bool x_unset = c->dh_unset;
  c->dh_unset = false;
  if (x_unset && !c->r0_pending) {
    hipLaunchKernelGGL(k_zero, dim3(std::min(c->nb, 4096)), dim3(kBlock), 0, c->stream,
                       (size_t)c->dm.n, c->dh);
    HIPCHK(hipGetLastError());
    x_unset = false;
  }
  if (c->r0_pending) {
    // x0 = 0 (calc_numsol clears dh): r0 = b - A 0 = b value for value
    // (A 0 sums zeros), so r0 and r-hat are rhs itself (read only), and
    // |r0|^2 is the |rhs|^2 the assembly is folding into slot 8: the first
    // alpha reads it on the device, the host after its first wait
    rs = c->rhs;
    r_src = c->rhs;
    sk = NAN;
    c->r0_pending = false;
  } else {
    {
      Scope sc(c, "resid");
      if (c->sys32)
        hipLaunchKernelGGL(k_resid<float>, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, mat7f(c),
                           x, c->rhs, r, c->kr_rs, c->partials);
      else
        hipLaunchKernelGGL(k_resid<double>, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, A, x,
                           c->rhs, r, c->kr_rs, c->partials);
      HIPCHK(hipGetLastError());
    }
    if ((rc = reduce_launch(c, nb, 1, 8, &seq)) || (rc = wait_flag(c, seq))) return rc;
    sk = c->h_scal[8];
    if (std::isnan(bnorm)) bnorm = sk;
  }
  const double *sk_dev = c->scal_dev + 8;
  double sk0 = DINFI, sk2, alpha = 0.0, beta, omega = 1.0, rnorm = 0.0;
  if ((rc = precon_setup(c))) return rc;
  int conv = 0, it;
  const double *p_src = rs;  // p = r0 = r-hat in the first iteration
  for (it = 1; it <= maxinn; it++) {
    if (it > 1) {
      beta = (sk / sk0) * (alpha / omega);
      hipLaunchKernelGGL(k_pupdate, dim3(nbl), dim3(kBlock), 0, c->stream, n, beta, omega, r,
                         p_src, p, v);
      HIPCHK(hipGetLastError());
      p_src = p;
    }
    if ((rc = ilu_apply(c, p_src, z)) || (rc = amg_record(c, p_src))) return rc;
    {
      Scope sc(c, "spmv");
      if (c->sys32)
        hipLaunchKernelGGL((k_spmv_dot<0, float>), dim3(nb), dim3(kBlock), 0, c->stream, dm, nb,
                           mat7f(c), z, v, rs, c->partials, nullptr);
      else
        hipLaunchKernelGGL(k_spmv_dot<0>, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, A, z, v,
                           rs, c->partials, nullptr);
      HIPCHK(hipGetLastError());
    }
    if ((rc = reduce_launch(c, nb, 1, 9, &seq))) return rc;
    // x += alpha z, s = r - alpha v with alpha = rho / (v.r-hat), and |s|^2
    hipLaunchKernelGGL(k_xr_update, dim3(nbl), dim3(kBlock), 0, c->stream, n, nbl, sk_dev,
                       c->scal_dev + 9, 0, x, r_src, r, z, v, nullptr, c->partials, nullptr,
                       x_unset ? 1 : 0);
    x_unset = false;
    r_src = r;
    HIPCHK(hipGetLastError());
    // With tail_ahead the half-step's convergence test is made on the device
    // by the fold, and calc_numsol's tail is launched behind it, running only
    // if the solve ends here: the GPU starts it without waiting for the host.
    const bool ahead = c->tail_ahead;
    if (ahead) {
      seq = ++c->flag_seq;
      hipLaunchKernelGGL(k_finalize_gate, dim3(1), dim3(kFin), 0, c->stream, c->partials, nbl,
                         c->scal, c->scal_dev, c->flag_map, seq, sk, bnorm, errtol,
                         c->conv_run + 2);
      HIPCHK(hipGetLastError());
      if ((rc = tail_launch(c, &c->tail_seq, true, 1, c->conv_run + 2))) return rc;
      if ((rc = wait_flag(c, seq))) return rc;
    } else if ((rc = reduce_launch(c, nbl, 1, 10, &seq)) || (rc = wait_flag(c, seq))) {
      return rc;
    }
    if (std::isnan(sk)) {  // the assembly's fold finished before this one
      sk = c->h_scal[ahead ? 18 : 8];
      c->r0_value = sk;
      if (std::isnan(bnorm)) bnorm = sk;
    }
    sk2 = c->h_scal[9];
    if (sk2 == 0.0 || !std::isfinite(sk2) || !std::isfinite(sk)) {
      *breakdown = 1;  // the update kernel skipped the update
      if (ahead && c->timing == 1) retime_last(c, "assemble", "assemble_skipped");
      break;
    }
    alpha = sk / sk2;
    rnorm = c->h_scal[10];
    if (ahead) {
      conv = c->h_scal[15] != 0.0 ? 1 : 0;  // the device's test, as the tail saw it
      if (conv) c->tail_valid = true;
      else if (c->timing == 1) retime_last(c, "assemble", "assemble_skipped");
    } else {
      conv = (std::sqrt(rnorm / bnorm) <= errtol) ? 1 : 0;
    }
    if (conv) break;
    if ((rc = ilu_apply(c, r, z)) || (rc = amg_record(c, r))) return rc;
    {
      Scope sc(c, "spmv");
      if (c->sys32)
        hipLaunchKernelGGL((k_spmv_dot<1, float>), dim3(nb), dim3(kBlock), 0, c->stream, dm, nb,
                           mat7f(c), z, t, r, c->partials, nullptr);
      else
        hipLaunchKernelGGL(k_spmv_dot<1>, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, A, z, t,
                           r, c->partials, nullptr);
      HIPCHK(hipGetLastError());
    }
    if ((rc = reduce_launch(c, nb, 2, 11, &seq))) return rc;
    // x += omega z, r = s - omega t with omega = (t.s)/(t.t), |r|^2 and r.r-hat
    hipLaunchKernelGGL(k_xr_update, dim3(nbl), dim3(kBlock), 0, c->stream, n, nbl,
                       c->scal_dev + 11, c->scal_dev + 12, 1, x, r, r, z, t, rs, c->partials,
                       nullptr, 0);
    HIPCHK(hipGetLastError());
    if ((rc = reduce_launch(c, nbl, 2, 13, &seq)) || (rc = wait_flag(c, seq))) return rc;
    omega = c->h_scal[11] / c->h_scal[12];
    if (c->h_scal[12] == 0.0 || omega == 0.0 || !std::isfinite(omega)) {
      *breakdown = 1;  // the update kernel skipped the update
      break;
    }
    rnorm = c->h_scal[13];
    conv = (std::sqrt(rnorm / bnorm) <= errtol) ? 1 : 0;
    if (conv || it == maxinn) break;
    sk0 = sk;
    sk = c->h_scal[14];
    sk_dev = c->scal_dev + 14;
  }
  if (std::isnan(sk)) {  // no iteration ran: the assembly's norm is still read
    if ((rc = wait_flag(c, c->r0_seq))) return rc;
    c->r0_value = c->h_scal[8];
  }
  if (it > maxinn) it = maxinn;
  *rnorm_out = rnorm;
  *in_iter_out = it;
  *conv_out = conv;
  return MSG123_OK;
}

// The linear solve.  With the device AMG, a breakdown or 200 iterations
// without convergence restart the solve from its initial guess with the
// FP32 line preconditioner (D-ILU when the columns exceed the line kernel's
// 640 cells): SA-AMG with Jacobi smoothing is not robust on Jacobians that
// lose the M-matrix sign pattern at wetting fronts, and the solve's answer
// must not depend on the preconditioner.
constexpr int kAmgIterCap = 200;
int bicgstab(msg123_ctx *c, double bnorm, double errtol, int maxinn, double *rnorm_out,
             int *in_iter_out, int *conv_out) {
  int breakdown = 0, rc;
  if (!use_amg32(c))
    return bicgstab_run(c, bnorm, errtol, maxinn, rnorm_out, in_iter_out, conv_out, &breakdown);
  const size_t N = (size_t)c->dm.n;
  if (!c->dh0 && hipMalloc((void **)&c->dh0, 8 * N) != hipSuccess) return MSG123_ENOMEM;
  HIPCHK(hipMemcpyAsync(c->dh0, c->dh, 8 * N, hipMemcpyDeviceToDevice, c->stream));
  const int cap = std::min(maxinn, kAmgIterCap);
  if ((rc = bicgstab_run(c, bnorm, errtol, cap, rnorm_out, in_iter_out, conv_out, &breakdown)))
    return rc;
  if (*conv_out || (!breakdown && cap == maxinn)) return MSG123_OK;
  const int used = *in_iter_out;
  c->amg_fallbacks++;
  if (std::isnan(bnorm)) bnorm = c->r0_value;  // the first attempt resolved it
  HIPCHK(hipMemcpyAsync(c->dh, c->dh0, 8 * N, hipMemcpyDeviceToDevice, c->stream));
  c->cur_precon = c->dm.nz <= 640 ? MSG123_PRECON_LINE32 : MSG123_PRECON_DILU;
  rc = bicgstab_run(c, bnorm, errtol, maxinn, rnorm_out, in_iter_out, conv_out, &breakdown);
  c->cur_precon = -1;
  *in_iter_out += used;
  return rc;
}

// check_abserrmax with optional update; maxvar = signed dh at the argmax,
// maxunk = h at the argmax of |h|
int maxerr_launch(msg123_ctx *c, bool update, unsigned *seq) {
  const int n = c->dm.n, nbl = nblk(n);
  const int nbm = std::min(nbl, kMaxrecBlocks);
  hipLaunchKernelGGL(k_update_maxerr, dim3(nbm), dim3(kBlock), 0, c->stream, n, c->h_pre, c->dh,
                     c->h_new, update ? 1 : 0, c->maxrec, nullptr);
  HIPCHK(hipGetLastError());
  *seq = ++c->flag_seq;
  hipLaunchKernelGGL(k_maxrec_final, dim3(1), dim3(kFinMax), 0, c->stream, nbm, c->maxrec,
                     c->maxrec_map, c->flag_map, *seq);
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}

// the folded record of the last maxerr_launch, after a wait covering it
int maxerr_read(msg123_ctx *c, double *maxvar, double *maxunk, int *nan_flag) {
  const MaxRec m = *c->h_maxrec;
  *nan_flag = m.nan;
  c->last_maxidx = m.ich;
  if (m.nan) {
    *maxvar = 0.0;
    *maxunk = NAN;
    return MSG123_OK;
  }
  *maxvar = (m.ich >= 0) ? m.sch : 0.0;
  *maxunk = (m.ival >= 0) ? m.sval : 0.0;
  return MSG123_OK;
}

int maxerr(msg123_ctx *c, bool update, double *maxvar, double *maxunk, int *nan_flag) {
  unsigned seq;
  if (int rc = maxerr_launch(c, update, &seq)) return rc;
  if (int rc = wait_flag(c, seq)) return rc;
  return maxerr_read(c, maxvar, maxunk, nan_flag);
}

int set_delt(msg123_ctx *c, double delt) {
  c->delt = delt;
  hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, c->stream, c->scal_dev + kDeltSlot, 1.0 / delt);
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}

int dcopy(msg123_ctx *c, double *d0, const double *s0, size_t n0, double *d1, const double *s1,
          size_t n1, double *d2, const double *s2, size_t n2) {
  const size_t nmax = std::max(n0, std::max(n1, n2));
  if (nmax == 0) return MSG123_OK;
  const int grid = (int)std::min<size_t>((nmax + kBlock - 1) / kBlock, 4096);
  CopyJob j{{d0, d1, d2}, {s0, s1, s2}, {n0, n1, n2}};
  hipLaunchKernelGGL(k_copy3, dim3(grid), dim3(kBlock), 0, c->stream, j);
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}

int dots3(msg123_ctx *c, const double *a, const double *b, double out[3]) {
  const int n = c->dm.n, nbl = nblk(n);
  hipLaunchKernelGGL(k_dots3, dim3(nbl), dim3(kBlock), 0, c->stream, n, nbl, a, b, c->partials);
  HIPCHK(hipGetLastError());
  return reduce_finalize(c, nbl, 3, out);
}

struct Newton {
  double l2_new = 0, l2_pre = 0, l2_jnorm = 0, lam = 1, eta = 0.5, gradient = 0, maxstep = 1;
  int back_iter = 0, back_flag = 0, beta_iter = 0, func_evals = 0;
  bool moved = false;  // the line search moved head_new off the full step
  // set_backtr's |head_new|^2 folded into kHnormSlot, not yet read
  bool maxstep_pending = false;
  unsigned maxstep_seq = 0;
};

// set_backtr (Nonlinear_solution.f90:395-412): maxstep = max(1, |h| VARMAX),
// read where the line search first needs it
int maxstep_ready(msg123_ctx *c, Newton &nt) {
  if (!nt.maxstep_pending) return MSG123_OK;
  if (int rc = wait_flag(c, nt.maxstep_seq)) return rc;
  nt.maxstep = std::sqrt(c->h_scal[kHnormSlot]) * VARMAX;
  if (nt.maxstep < 1.0) nt.maxstep = 1.0;
  nt.maxstep_pending = false;
  return MSG123_OK;
}

// calc_funcl2norm (Nonlinear_solution.f90:760-790)
int funcl2norm(msg123_ctx *c, Newton &nt, double lam) {
  nt.func_evals++;
  nt.moved = true;
  return residual(c, nullptr, true, lam, c->fbuf, nullptr, 0.0, &nt.l2_new);
}

// set_eater (Nonlinear_solution.f90:414-439); eta_max is the reference's 0.9
// unless msg123_set_forcing lowered it
void set_eater(Newton &nt, double eta_max) {
  const double eta_min = 1.0e-4, eta_alpha = (1.0 + std::sqrt(5.0)) / 2.0;
  double eta_safe = std::pow(nt.eta, eta_alpha);
  const double minus1 = 1.0 - nt.lam;
  const double lin = std::sqrt(minus1 * minus1 * nt.l2_pre + 2.0 * nt.lam * minus1 * nt.gradient +
                               nt.lam * nt.lam * nt.l2_jnorm);
  nt.eta = std::fabs(std::sqrt(nt.l2_new) - lin) / std::sqrt(nt.l2_pre);
  if (eta_safe < 0.1) eta_safe = 0.0;
  nt.eta = std::fmax(nt.eta, eta_safe);
  nt.eta = std::fmax(nt.eta, eta_min);
  nt.eta = std::fmin(nt.eta, eta_max);
}

// The sums run_backtr starts from, folded into result slots 0..4 and
// kMaxRed - 1 (with the algebraic slope also F(head_new) and its norm, slot
// 5): one pass for l2_pnorm = |dh|^2, the Brown-Saad sums v.v, x.v, |v|_1
// with x = head_pre, v = dh, the algebraic slope's sums, and lam_length's
// max (run_backtr :504-513) for the uncapped step.  *seq: the last fold's
// sequence number.
int backtr_launch(msg123_ctx *c, unsigned *seq, bool spec = false) {
  const int n = c->dm.n, nbl = nblk(n);
  const bool alg = c->algebraic_slope != 0;
  const int nbs = std::min(nbl, kMaxrecBlocks);
  double *maxpart = c->partials + (size_t)(kMaxRed - 1) * nbs;
  if (alg)
    hipLaunchKernelGGL(k_backtr_sums<true>, dim3(nbs), dim3(kBlock), 0, c->stream, n, nbs, c->dh,
                       c->h_pre, c->rhs, c->kr_r, c->h_new, c->partials, maxpart, nullptr);
  else
    hipLaunchKernelGGL(k_backtr_sums<false>, dim3(nbs), dim3(kBlock), 0, c->stream, n, nbs, c->dh,
                       c->h_pre, c->rhs, c->kr_r, c->h_new, c->partials, maxpart, nullptr);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kFinMax), 0, c->stream, nbs, maxpart,
                     c->scal + (kMaxRed - 1));
  HIPCHK(hipGetLastError());
  int rc;
  if ((rc = reduce_launch(c, nbs, alg ? 5 : 3, 0, seq))) return rc;
  // With the algebraic slope nothing needs these sums before calc_func(head_new):
  // both are folded before the one wait.  With `spec`, F(head_new) comes as
  // -rhs of the next iteration's system, assembled ahead into rhs2 and A2
  // (the same F, bit for bit); its norm is folded into slot 8 for that
  // iteration's solve and, last, into slot 5.
  if (alg && spec) {
    const HeadIn hin{c->h_new, nullptr, nullptr, 0.0, nullptr};
    const FusedOut o{c->rhs2, c->sys32 ? nullptr : c->A2, c->srat_new, nullptr, 0.0, nullptr,
                     c->partials, nullptr, nullptr, c->sys32 ? c->A2f : nullptr};
    if ((rc = launch_fused<kAssem>(c, hin, o, "assemble"))) return rc;
    c->spec_r0_seq = *seq = ++c->flag_seq;
    hipLaunchKernelGGL(k_finalize2, dim3(1), dim3(kFin), 0, c->stream, c->partials,
                       c->ftg[1].nblocks, c->scal + 8, c->scal_dev + 8, c->scal + 5,
                       c->scal_dev + 5, c->flag_map, *seq, nullptr);
    HIPCHK(hipGetLastError());
    return MSG123_OK;
  }
  if (alg && (rc = residual_launch(c, c->h_new, c->fbuf, 5, seq))) return rc;
  return MSG123_OK;
}

// The merged tail of a host-driven Newton iteration with the algebraic
// slope: k_newton_tail (update, records, sums), the folds, the log values
// at the cell of the largest change, and F(head_new) (or, with `spec`, the
// next system assembled ahead); *seq: the last fold's sequence number.
// gate (device, may be null): launched ahead of the Krylov solve's end,
// every kernel of the tail runs only where *gate != 0 (k_finalize_gate).
int tail_launch(msg123_ctx *c, unsigned *seq, bool spec, int lconv, const int *gate) {
  const int n = c->dm.n, nbl = nblk(n);
  const int nbs = std::min(nbl, kMaxrecBlocks);
  double *maxpart = c->partials + (size_t)(kMaxRed - 1) * nbs;
  hipLaunchKernelGGL(k_newton_tail, dim3(nbs), dim3(kBlock), 0, c->stream, n, nbs, c->h_pre,
                     c->dh, c->h_new, c->rhs, c->kr_r, c->maxrec, c->partials, maxpart, gate);
  *seq = ++c->flag_seq;
  hipLaunchKernelGGL(k_tail_final, dim3(1), dim3(kFin), 0, c->stream, nbs, c->maxrec,
                     c->maxrec_map, diag_log(c), c->rhs, c->h_new, c->log_map, maxpart,
                     c->scal + (kMaxRed - 1), c->partials, 5, c->scal, c->scal_dev, c->flag_map,
                     *seq, lconv, c->len_scal, c->prm.criteria, spec ? c->conv_run : nullptr,
                     gate);
  HIPCHK(hipGetLastError());
  int rc;
  if (spec) {
    // the assembly ahead runs where the iteration has not converged; where
    // it has, only S(head_new) and the seepage flags are brought up to date
    const HeadIn hin{c->h_new, nullptr, nullptr, 0.0, nullptr};
    const FusedOut o{c->rhs2, c->sys32 ? nullptr : c->A2, c->srat_new, nullptr, 0.0, nullptr,
                     c->partials, c->conv_run, nullptr, c->sys32 ? c->A2f : nullptr};
    if ((rc = launch_fused<kAssem>(c, hin, o, "assemble"))) return rc;
    // |F(head_new)|^2 into slot 8 (the next iteration's r0 norm) and slot 5
    c->spec_r0_seq = *seq = ++c->flag_seq;
    hipLaunchKernelGGL(k_finalize2, dim3(1), dim3(kFin), 0, c->stream, c->partials,
                       c->ftg[1].nblocks, c->scal + 8, c->scal_dev + 8, c->scal + 5,
                       c->scal_dev + 5, c->flag_map, *seq, c->conv_run);
    const int grid = std::min(c->nb, 4 * c->num_cu);
    if (c->law == MSG123_LAW_LOGEXP)
      hipLaunchKernelGGL(k_conv_state<true>, dim3(grid), dim3(kBlock), 0, c->stream, c->dm.n,
                         c->dm.nxy, c->conv_run + 1, c->h_new, c->cell_top, soil(c), surface(c),
                         c->srat_new, c->kr_p);
    else
      hipLaunchKernelGGL(k_conv_state<false>, dim3(grid), dim3(kBlock), 0, c->stream, c->dm.n,
                         c->dm.nxy, c->conv_run + 1, c->h_new, c->cell_top, soil(c), surface(c),
                         c->srat_new, c->kr_p);
    HIPCHK(hipGetLastError());
    return MSG123_OK;
  }
  return residual_launch(c, c->h_new, c->fbuf, 5, seq);
}

// run_backtr (Nonlinear_solution.f90:441-631) with calc_vecjacf (Calc_function.f90:515-621)
int run_backtr(msg123_ctx *c, Newton &nt) {
  const int n = c->dm.n, nbl = nblk(n);
  const double back_alpha = 1.0e-4, back_beta = 0.9;
  double l2_new2 = nt.l2_new, lam2 = 1.0, maxpnorm = 1.0, temp_lam;
  int rc;
  if ((rc = maxstep_ready(c, nt))) return rc;
  nt.lam = 1.0;
  // One pass for l2_pnorm = |dh|^2, the Brown-Saad sums v.v, x.v, |v|_1 with
  // x = head_pre, v = dh, the algebraic slope's sums, and lam_length's max
  // (run_backtr :504-513) for the uncapped step; one synchronisation.
  const bool alg = c->algebraic_slope != 0;
  if (c->backtr_pending) {
    c->backtr_pending = false;  // enqueued with the convergence test and waited for
  } else {
    unsigned seq;
    if ((rc = backtr_launch(c, &seq)) || (rc = wait_flag(c, seq))) return rc;
  }
  double d5[5];
  for (int k = 0; k < 5; k++) d5[k] = c->h_scal[k];
  const double lam_length_uncapped = c->h_scal[kMaxRed - 1];
  const double l2_pnorm = d5[0], l2_v = d5[0], l2_x = d5[1], l1_v = d5[2];
  // calc_vecjacf(head_pre, dh): F1 = F(x), jc = x + eps v, jv = -(F(jc) - F1)/eps.
  // F(head_pre) is the residual this Newton iteration assembled, -rhs, bit for
  // bit (same head, same time level; surf_rati is 1 wherever the seepage
  // term is non-zero), so it is taken from rhs instead of being recomputed
  // (when the difference quotient is needed).
  nt.func_evals += 1;
  const double sgn = (l2_x < 0.0) ? -1.0 : 1.0;
  const double eps = sgn * std::sqrt(MACHI_EPS) * std::fmax(std::fabs(l2_x), l1_v) / l2_v;
  const double eps_inv = 1.0 / eps;
  auto fd_slope = [&](double *out) {  // calc_vecjacf's difference quotient
    hipLaunchKernelGGL(k_scale_copy, dim3(nbl), dim3(kBlock), 0, c->stream, n, -1.0, c->rhs,
                       c->f1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_axpy, dim3(nbl), dim3(kBlock), 0, c->stream, n, c->h_pre, eps, c->dh,
                       c->jvec);
    HIPCHK(hipGetLastError());
    return residual(c, c->jvec, false, 0.0, c->fbuf, c->f1, eps_inv, out);
  };
  double slope0;
  bool state_stale = false;  // the context's state is not head_new's
  if (alg) {
    // J dh = rhs - r for the Krylov residual r left by the solve, so
    // rhs . (-J dh) = -|rhs|^2 + rhs . r: no residual evaluation
    slope0 = -d5[3] + d5[4];
  } else if ((rc = fd_slope(&slope0))) {
    return rc;
  }
  // calc_func(head_new) -> new_f and its norm.  Taken after the J.v
  // evaluation (the two are independent) so that the saturations and
  // permeabilities left in the context always belong to the last head
  // evaluated, which is the accepted head_new on every exit below (with the
  // algebraic slope it was launched with the sums above).
  nt.func_evals++;
  if (alg) nt.l2_new = c->h_scal[5];
  else if ((rc = residual(c, c->h_new, false, 0.0, c->fbuf, nullptr, 0.0, &nt.l2_new))) return rc;
  if (c->algebraic_slope) {
    // When the full step increases the residual norm, the assembled
    // Jacobian may be a poor derivative there (the residual is not smooth at
    // saturation and seepage switches): the backtracking then interpolates
    // with the reference's difference-quotient slope.  That evaluation
    // leaves the state of h_pre + eps dh in the context; the backtracking
    // overwrites it, and where the full step is kept after all, head_new's
    // residual is evaluated again below.
    double sq = std::sqrt(l2_pnorm), mp = 1.0;
    if (sq > nt.maxstep) mp = nt.maxstep / sq;
    if (0.5 * nt.l2_new > 0.5 * nt.l2_pre + back_alpha * slope0 * mp &&
        nt.l2_new > nt.l2_pre) {
      if ((rc = fd_slope(&slope0))) return rc;
      nt.func_evals++;
      state_stale = true;
    }
  }
  double sql2_pnorm = std::sqrt(l2_pnorm);
  if (sql2_pnorm > nt.maxstep) {
    maxpnorm = nt.maxstep / sql2_pnorm;
    hipLaunchKernelGGL(k_scale, dim3(nbl), dim3(kBlock), 0, c->stream, n, maxpnorm, c->dh);
    HIPCHK(hipGetLastError());
    sql2_pnorm = nt.maxstep;
  }
  const double step_len = sql2_pnorm;
  // l2_new = |new_f|^2 was computed with new_f; slope = rhs.jacvec * maxpnorm
  const double slope = slope0 * maxpnorm;
  double lam_length = lam_length_uncapped;
  if (maxpnorm != 1.0) {  // dh was scaled: its ratios again, as the reference takes them
    hipLaunchKernelGGL(k_lam_length, dim3(nbl), dim3(kBlock), 0, c->stream, n, c->dh, c->h_new,
                       c->partials);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kFinMax), 0, c->stream, nbl, c->partials,
                       c->scal);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    lam_length = c->h_scal[0];
  }
  nt.l2_jnorm = nt.l2_jnorm * maxpnorm * maxpnorm;
  const double step_tol = std::pow(MACHI_EPS, 2.0 / 3.0);
  const double lam_min = step_tol / lam_length;
  double alpha_cond = 0.5 * nt.l2_pre + back_alpha * nt.lam * slope, beta_cond;
  nt.gradient = slope;
  for (;;) {
    if (0.5 * nt.l2_new <= alpha_cond) break;
    nt.back_iter++;
    // The first pass evaluates lam = 1, the head already evaluated above
    // (h_pre + 1*dh is head_new bit for bit when dh was not capped): its
    // norm and state are in hand unless the difference quotient replaced
    // the state.
    const bool known = nt.back_iter == 1 && nt.lam == 1.0 && maxpnorm == 1.0 && !state_stale;
    if (!known && (rc = funcl2norm(c, nt, nt.lam))) return rc;
    const double lam = nt.lam;
    if (lam == 1.0) {
      temp_lam = -slope / (2.0 * (0.5 * nt.l2_new - 0.5 * nt.l2_pre - slope));
    } else {
      const double rhs1 = 0.5 * nt.l2_new - 0.5 * nt.l2_pre - lam * slope;
      const double rhs2 = l2_new2 - 0.5 * nt.l2_pre - lam2 * slope;
      const double lam_inv = 1.0 / (lam * lam), lam2_inv = 1.0 / (lam2 * lam2);
      const double del_lam = 1.0 / (lam - lam2);
      const double av = (rhs1 * lam_inv - rhs2 * lam2_inv) * del_lam;
      const double bv = (-lam2 * rhs1 * lam_inv + lam * rhs2 * lam2_inv) * del_lam;
      if (av == 0) {
        temp_lam = -slope / (2.0 * bv);
      } else {
        const double root = bv * bv - 3.0 * av * slope;
        if (root < 0.0) temp_lam = 0.5 * lam;
        else if (bv <= 0.0) temp_lam = (-bv + std::sqrt(root)) / (3.0 * av);
        else temp_lam = -slope / (bv + std::sqrt(root));
        if (temp_lam > 0.5 * lam) temp_lam = 0.5 * lam;
      }
    }
    lam2 = lam;
    l2_new2 = 0.5 * nt.l2_new;
    nt.lam = std::fmax(temp_lam, 0.1 * lam);
    if (nt.lam < lam_min) {
      nt.back_flag = 1;
      return MSG123_OK;
    }
    alpha_cond = 0.5 * nt.l2_pre + back_alpha * nt.lam * slope;
  }
  if (state_stale && nt.back_iter == 0) {  // the full step is kept: its state again
    nt.func_evals++;
    if ((rc = residual(c, c->h_new, false, 0.0, c->fbuf, nullptr, 0.0, &nt.l2_new))) return rc;
  }
  alpha_cond = 0.5 * nt.l2_pre + back_alpha * nt.lam * slope;
  beta_cond = 0.5 * nt.l2_pre + back_beta * nt.lam * slope;
  // Without the refinement the step the alpha condition accepted is kept:
  // each refinement probe costs a residual evaluation, and after a backtrack
  // the probes rarely pay for themselves in Newton iterations.
  if (c->beta_refine && 0.5 * nt.l2_new < beta_cond) {
    if (nt.lam == 1.0 && sql2_pnorm < step_len) {
      const double lam_max = step_len / sql2_pnorm;
      for (;;) {
        if (0.5 * nt.l2_new > alpha_cond || 0.5 * nt.l2_new >= beta_cond || nt.lam >= lam_max)
          break;
        lam2 = nt.lam;
        l2_new2 = 0.5 * nt.l2_new;
        nt.lam = std::fmin(2.0 * nt.lam, lam_max);
        if ((rc = funcl2norm(c, nt, nt.lam))) return rc;
        alpha_cond = 0.5 * nt.l2_pre + back_alpha * nt.lam * slope;
        beta_cond = 0.5 * nt.l2_pre + back_beta * nt.lam * slope;
      }
    }
    if (nt.lam < 1.0 || (nt.lam > 1.0 && 0.5 * nt.l2_new > alpha_cond)) {
      double lam_base = std::fmin(nt.lam, lam2), lam_diff = std::fabs(lam2 - nt.lam);
      for (;;) {
        if (0.5 * nt.l2_new <= alpha_cond && (0.5 * nt.l2_new >= beta_cond || lam_diff <= lam_min))
          break;
        const double lam_incr = 0.5 * lam_diff;
        nt.lam = lam_base + lam_incr;
        if ((rc = funcl2norm(c, nt, nt.lam))) return rc;
        alpha_cond = 0.5 * nt.l2_pre + back_alpha * nt.lam * slope;
        beta_cond = 0.5 * nt.l2_pre + back_beta * nt.lam * slope;
        if (0.5 * nt.l2_new > alpha_cond) {
          lam_diff = lam_incr;
        } else if (0.5 * nt.l2_new < beta_cond) {
          lam_base = nt.lam;
          lam_diff = lam_diff - lam_incr;
        }
        if (lam_diff == 0.0) break;
      }
      if (0.5 * nt.l2_new < beta_cond) {
        if ((rc = funcl2norm(c, nt, lam_base))) return rc;
        nt.beta_iter++;
      }
      if (nt.beta_iter == 10) {
        nt.back_flag = 1;
        return MSG123_OK;
      }
    }
  }
  return MSG123_OK;
}

template <typename T>
int dalloc(T **p, size_t count) {
  if (hipMalloc((void **)p, sizeof(T) * (count ? count : 1)) != hipSuccess) return MSG123_ENOMEM;
  return MSG123_OK;
}

template <typename T>
int h2d(msg123_ctx *c, T *dst, const T *src, size_t count) {
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
  return MSG123_OK;
}

}  // namespace

// ===========================================================================
// Scratch of msg123_calc_mass_noalloc: the per-block partial sums of the
// mass-balance kernel, six per block of kBlock cells, in a file-scope device
// array so that the entry point allocates nothing.  Every GPU the library
// runs on holds its own copy (3 MiB) until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: its device
// address is looked up once per device (scratch_init, called by
// msg123_ctx_create with the context's device current) and only that
// address is handed to kernels.
constexpr int kNoallocMaxBlocks = 65536;  // cells <= 65536 * kBlock = 16,777,216
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[6 * kNoallocMaxBlocks];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return MSG123_EDEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return MSG123_OK;
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double *>(p);
  return MSG123_OK;
}

double *scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace

// ===========================================================================
// C ABI

extern "C" int msg123_ctx_create(const msg123_mesh *mesh, const msg123_params *params,
                                 msg123_ctx **out) {
  if (!mesh || !out) return MSG123_EINVAL;
  *out = nullptr;
  const int nx = mesh->nx, ny = mesh->ny, nz = mesh->nz;
  if (nx < 1 || ny < 1 || nz < 1) return MSG123_EINVAL;
  const long long nll = (long long)nx * ny * nz;
  if (nll != mesh->ncalc || (long long)nx * ny != mesh->ncals || nll > (1LL << 30))
    return MSG123_EINVAL;
  const int n = (int)nll, nxy = nx * ny;
  // the pattern must be the full 7-point box stencil in CSR order up,N,W,E,S,down
  std::vector<double> hE(n, 0.0), aE(n, 0.0), hS(n, 0.0), aS(n, 0.0), hD(n, 0.0), aD(n, 0.0);
  {
    if (mesh->offind[0] != 0) return MSG123_EINVAL;
    for (int i = 0; i < n; i++) {
      const int x = i % nx, y = (i / nx) % ny, z = i / nxy;
      const int nb[6] = {z > 0 ? i - nxy : -1, y > 0 ? i - nx : -1, x > 0 ? i - 1 : -1,
                         x < nx - 1 ? i + 1 : -1, y < ny - 1 ? i + nx : -1,
                         z < nz - 1 ? i + nxy : -1};
      int k = mesh->offind[i];
      for (int f = 0; f < 6; f++) {
        if (nb[f] < 0) continue;
        if (k >= mesh->offind[i + 1] || mesh->offrow[k] != nb[f] + 1 || mesh->dir_conn[k] != f + 1)
          return MSG123_EINVAL;
        const double hy = mesh->hydf_conn[k], ab = mesh->abyd_conn[k];
        if (f == 3) { hE[i] = hy; aE[i] = ab; }
        if (f == 4) { hS[i] = hy; aS[i] = ab; }
        if (f == 5) { hD[i] = hy; aD[i] = ab; }
        k++;
      }
      if (k != mesh->offind[i + 1]) return MSG123_EINVAL;
    }
    if (mesh->offind[n] != mesh->nconn) return MSG123_EINVAL;
    // bitwise symmetry of the face coefficients
    for (int i = 0; i < n; i++) {
      const int x = i % nx, y = (i / nx) % ny, z = i / nxy;
      int k = mesh->offind[i];
      const int nb[3] = {z > 0 ? i - nxy : -1, y > 0 ? i - nx : -1, x > 0 ? i - 1 : -1};
      for (int f = 0; f < 3; f++) {
        if (nb[f] < 0) continue;
        const double hy = mesh->hydf_conn[k], ab = mesh->abyd_conn[k];
        const double *hh = f == 0 ? hD.data() : (f == 1 ? hS.data() : hE.data());
        const double *aa = f == 0 ? aD.data() : (f == 1 ? aS.data() : aE.data());
        if (std::memcmp(&hy, &hh[nb[f]], 8) != 0 || std::memcmp(&ab, &aa[nb[f]], 8) != 0)
          return MSG123_EINVAL;
        k++;
      }
    }
  }
  if (params && (params->precon < MSG123_PRECON_DILU || params->precon > MSG123_PRECON_MCILU32))
    return MSG123_EINVAL;
  if (params && params->precon == MSG123_PRECON_LINE32 && nz > 640) return MSG123_EINVAL;
  msg123_ctx *c = new msg123_ctx();
  c->dm = Dims{nx, ny, nz, n, nxy};
  c->ns = nxy;
  c->nconn = mesh->nconn;
  c->nb = nblk(n);
  c->ftg[0] = c->ftg[1] = tiling_shape<kFTX, kFTY>(c->dm, kFusedBlocks);
  const int max_tiles = tiling_shape<kFTX, kFTY>(c->dm, kFusedMaxBlocks).nblocks;
  c->len_scal = mesh->len_scal;
  c->newper = mesh->newper;
  c->newper_inv = 1.0 / mesh->newper;
  c->prm = params ? *params : msg123_params{20, 10, 1.0e-3, MSG123_PRECON_DILU, 0};
  c->offind.assign(mesh->offind, mesh->offind + n + 1);
  c->hydf_surf_h.assign(mesh->hydf_surf, mesh->hydf_surf + nxy);
  c->offrow.assign(mesh->offrow, mesh->offrow + mesh->nconn);
  c->dir.assign(mesh->dir_conn, mesh->dir_conn + mesh->nconn);
  int rc = MSG123_OK;
  auto fail = [&](int e) {
    msg123_ctx_destroy(c);
    return e;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return fail(MSG123_EDEVICE);
  if ((rc = scratch_init(c->device))) return fail(rc);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(MSG123_EDEVICE);
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return fail(MSG123_EDEVICE);
  const size_t N = (size_t)n, S = (size_t)nxy;
  const int nbl = nblk(n);
  if ((rc = dalloc(&c->cell_top, N)) || (rc = dalloc(&c->cell_vol, N)) ||
      (rc = dalloc(&c->soil, 5 * N)) || (rc = dalloc(&c->faces, 6 * N)) ||
      (rc = dalloc(&c->surf, 4 * S)) || (rc = dalloc(&c->surf_head, S)) ||
      (rc = dalloc(&c->surf_rati, S)) || (rc = dalloc(&c->calc_rech, S)) ||
      (rc = dalloc(&c->surf_old, S)) ||
      (rc = dalloc(&c->h_new, N)) || (rc = dalloc(&c->h_old, N)) || (rc = dalloc(&c->h_pre, N)) ||
      (rc = dalloc(&c->dh, N)) || (rc = dalloc(&c->srat_new, N)) ||
      (rc = dalloc(&c->srat_old, N)) || (rc = dalloc(&c->rel_perm, N)) ||
      (rc = dalloc(&c->sst, N)) || (rc = dalloc(&c->srat_p, N)) || (rc = dalloc(&c->kr_p, N)) ||
      (rc = dalloc(&c->fbuf, N)) || (rc = dalloc(&c->f1, N)) || (rc = dalloc(&c->jvec, N)) ||
      (rc = dalloc(&c->rhs, N)) || (rc = dalloc(&c->A, 7 * N)) || (rc = dalloc(&c->kr_r, N)) ||
      (rc = dalloc(&c->kr_rs, N)) || (rc = dalloc(&c->kr_p_, N)) || (rc = dalloc(&c->kr_v, N)) ||
      (rc = dalloc(&c->kr_z, N)) || (rc = dalloc(&c->kr_t, N)) || (rc = dalloc(&c->dg, N)) ||
      (rc = dalloc(&c->partials, kMaxRed * (size_t)std::max(std::max(nbl, c->nb), max_tiles))) ||
      (rc = dalloc(&c->maxrec, (size_t)nbl)) || (rc = dalloc(&c->scal_dev, kScal)) ||
      (rc = dalloc(&c->zinv, N)) || (rc = dalloc(&c->zcp, N)) || (rc = dalloc(&c->za, N)) ||
      (rc = dalloc(&c->zero, 32)) ||
      (c->prm.precon == MSG123_PRECON_AMG && (rc = dalloc(&c->amg_last, N))))
    return fail(rc);
  c->mc_np = (n + 1) / 2;
  if (c->prm.precon == MSG123_PRECON_MCILU32 && (rc = dalloc(&c->mc, 14 * (size_t)c->mc_np)))
    return fail(rc);
  // the reductions' results: written by k_finalize straight into coherent
  // pinned memory, so that reading one costs no copy launch
  if (hipHostMalloc((void **)&c->h_scal, kScal * sizeof(double),
                    hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->scal, c->h_scal, 0) != hipSuccess ||
      hipHostMalloc((void **)&c->h_flag, 64, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->flag_map, c->h_flag, 0) != hipSuccess ||
      hipHostMalloc((void **)&c->h_log, 4 * sizeof(double),
                    hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->log_map, c->h_log, 0) != hipSuccess ||
      hipHostMalloc((void **)&c->h_maxrec, sizeof(MaxRec),
                    hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->maxrec_map, c->h_maxrec, 0) != hipSuccess ||
      hipMalloc((void **)&c->dn, sizeof(DevNewton)) != hipSuccess ||
      hipHostMalloc((void **)&c->dn_host, sizeof(DevNewton),
                    hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->dn_map, c->dn_host, 0) != hipSuccess)
    return fail(MSG123_ENOMEM);
  *c->h_flag = 0;
  std::vector<float> soilv(5 * N);
  std::memcpy(&soilv[0], mesh->reta, 4 * N);
  std::memcpy(&soilv[N], mesh->retn, 4 * N);
  std::memcpy(&soilv[2 * N], mesh->resi, 4 * N);
  std::memcpy(&soilv[3 * N], mesh->poro, 4 * N);
  std::memcpy(&soilv[4 * N], mesh->ss, 4 * N);
  std::vector<double> facev(6 * N);
  std::memcpy(&facev[0], hE.data(), 8 * N);
  std::memcpy(&facev[N], aE.data(), 8 * N);
  std::memcpy(&facev[2 * N], hS.data(), 8 * N);
  std::memcpy(&facev[3 * N], aS.data(), 8 * N);
  std::memcpy(&facev[4 * N], hD.data(), 8 * N);
  std::memcpy(&facev[5 * N], aD.data(), 8 * N);
  // Face geometry, volumes and cell tops: per-layer tables when every
  // interior face of a layer has the same A/d and every cell of a layer the
  // same volume and top (bit for bit), else the per-cell arrays.
  bool layered = true;
  std::vector<double> gt(5 * (size_t)nz, 0.0);  // aE, aS, aD, vol, top per layer
  for (int z = 0; z < nz && layered; z++) {
    const int b = z * nxy;
    gt[3 * (size_t)nz + z] = mesh->cell_vol[b];
    gt[4 * (size_t)nz + z] = mesh->cell_top[b];
    bool haveE = false, haveS = false, haveD = false;
    for (int q = 0; q < nxy && layered; q++) {
      const int i = b + q, x = q % nx, y = q / nx;
      auto same = [&](double v, double &slot, bool &have) {
        if (!have) { slot = v; have = true; return true; }
        return std::memcmp(&v, &slot, 8) == 0;
      };
      if (std::memcmp(&mesh->cell_vol[i], &gt[3 * (size_t)nz + z], 8) != 0 ||
          std::memcmp(&mesh->cell_top[i], &gt[4 * (size_t)nz + z], 8) != 0)
        layered = false;
      if (x < nx - 1 && !same(aE[i], gt[z], haveE)) layered = false;
      if (y < ny - 1 && !same(aS[i], gt[(size_t)nz + z], haveS)) layered = false;
      if (z < nz - 1 && !same(aD[i], gt[2 * (size_t)nz + z], haveD)) layered = false;
    }
  }
  std::vector<double> surfv(4 * S);
  std::memcpy(&surfv[0], mesh->hydf_surf, 8 * S);
  std::memcpy(&surfv[S], mesh->abyd_surf, 8 * S);
  std::memcpy(&surfv[2 * S], mesh->surf_top, 8 * S);
  std::memcpy(&surfv[3 * S], mesh->surf_bott, 8 * S);
  // ILU tile schedule, grouped by tile diagonal I+J+K
  const int NI = (nx + kTX - 1) / kTX, NJ = (ny + kTY - 1) / kTY, NK = (nz + kTZ - 1) / kTZ;
  std::vector<Tile> tl;
  for (int D = 0; D <= NI + NJ + NK - 3; D++) {
    c->diag_off.push_back((int)tl.size());
    for (int K = 0; K < NK; K++)
      for (int J = 0; J < NJ; J++) {
        const int I = D - J - K;
        if (I >= 0 && I < NI) tl.push_back(Tile{I, J, K, 0});
      }
    c->diag_cnt.push_back((int)tl.size() - c->diag_off.back());
  }
  c->tni = NI;
  c->tnj = NJ;
  c->tnk = NK;
  c->ntiles = (int)tl.size();
  if ((rc = dalloc(&c->tiles, tl.size())) || (rc = dalloc(&c->wave_ctl, 2)) ||
      (rc = dalloc(&c->wave_flags, tl.size())))
    return fail(rc);
  if (hipMemsetAsync(c->wave_ctl, 0, 2 * sizeof(unsigned), c->stream) != hipSuccess ||
      hipMemsetAsync(c->wave_flags, 0, tl.size() * sizeof(unsigned), c->stream) != hipSuccess)
    return fail(MSG123_EDEVICE);
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) ==
            hipSuccess && ncu > 0)
      c->num_cu = ncu;
  }
  if (layered) {
    if ((rc = dalloc(&c->geo_tab, gt.size())) || (rc = h2d(c, c->geo_tab, gt.data(), gt.size())))
      return fail(rc);
    const double *g = c->geo_tab;
    c->geo = Geo{g, g + nz, g + 2 * (size_t)nz, g + 3 * (size_t)nz, g + 4 * (size_t)nz, 1,
                 c->zero};
  } else {
    c->geo = Geo{c->faces + N, c->faces + 3 * N, c->faces + 5 * N, c->cell_vol, c->cell_top, 0,
                 c->zero};
  }
  if ((rc = h2d(c, c->cell_top, mesh->cell_top, N)) || (rc = h2d(c, c->cell_vol, mesh->cell_vol, N)) ||
      (rc = h2d(c, c->soil, soilv.data(), 5 * N)) || (rc = h2d(c, c->faces, facev.data(), 6 * N)) ||
      (rc = h2d(c, c->surf, surfv.data(), 4 * S)) || (rc = h2d(c, c->tiles, tl.data(), tl.size())))
    return fail(rc);
  if (hipMemsetAsync(c->A, 0, sizeof(double) * 7 * N, c->stream) != hipSuccess ||
      hipMemsetAsync(c->surf_rati, 0, sizeof(double) * S, c->stream) != hipSuccess ||
      hipMemsetAsync(c->calc_rech, 0, sizeof(double) * S, c->stream) != hipSuccess ||
      hipMemsetAsync(c->zero, 0, sizeof(double) * 32, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(MSG123_EDEVICE);
  *out = c;
  return MSG123_OK;
}

extern "C" void msg123_ctx_destroy(msg123_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->conv_log) std::fclose(c->conv_log);
  amg_gpu_destroy(c->amg32);
  void *bufs[] = {c->cell_top, c->cell_vol, c->soil, c->faces, c->surf, c->surf_head, c->surf_old,
                  c->surf_rati, c->calc_rech, c->h_new, c->h_old, c->h_pre, c->dh, c->srat_new,
                  c->srat_old, c->rel_perm, c->sst, c->srat_p, c->kr_p, c->fbuf, c->f1, c->jvec,
                  c->rhs, c->A, c->kr_r, c->kr_rs, c->kr_p_, c->kr_v, c->kr_z, c->kr_t, c->dg,
                  c->partials, c->maxrec, c->scal_dev, c->tiles, c->zinv, c->zcp, c->za,
                  c->mc, c->amg_last, c->h_prev, c->h_lin, c->h_try, c->net_scale, c->net_x, c->lnks,
                  c->dh0, c->mass_acc, c->drain, c->geo_tab, c->probe_idx, c->probe_val,
                  c->zero, c->wave_ctl, c->wave_flags, c->rl, c->rl_cols, c->bc_cell,
                  c->bc_beg, c->bc_well, c->bc_sea, c->f_bnd, c->bm, c->A2, c->rhs2,
                  c->conv_run, c->Af, c->A2f};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  if (c->h_log) (void)hipHostFree(c->h_log);
  if (c->h_maxrec) (void)hipHostFree(c->h_maxrec);
  if (c->dn) (void)hipFree(c->dn);
  if (c->dn_host) (void)hipHostFree(c->dn_host);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int msg123_set_state(msg123_ctx *c, const double *head_new, const double *head_old,
                                const double *srat_old, const double *surf_head,
                                const double *surf_rati, const double *calc_rech, double delt) {
  if (!c) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n, S = (size_t)c->ns;
  int rc;
  if (head_new && (rc = h2d(c, c->h_new, head_new, N))) return rc;
  if (head_old && (rc = h2d(c, c->h_old, head_old, N))) return rc;
  if (srat_old && (rc = h2d(c, c->srat_old, srat_old, N))) return rc;
  if (surf_head && (rc = h2d(c, c->surf_head, surf_head, S))) return rc;
  if (surf_rati && (rc = h2d(c, c->surf_rati, surf_rati, S))) return rc;
  if (calc_rech && (rc = h2d(c, c->calc_rech, calc_rech, S))) return rc;
  if (delt > 0.0 && (rc = set_delt(c, delt))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_get_state(msg123_ctx *c, double *head_new, double *srat_new,
                                double *rel_perm, double *surf_head, double *surf_rati) {
  if (!c) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n, S = (size_t)c->ns;
  if (head_new) HIPCHK(hipMemcpyAsync(head_new, c->h_new, 8 * N, hipMemcpyDeviceToHost, c->stream));
  if (srat_new) HIPCHK(hipMemcpyAsync(srat_new, c->srat_new, 8 * N, hipMemcpyDeviceToHost, c->stream));
  if (rel_perm) {
    if (int rc = refresh_aux(c)) return rc;
    HIPCHK(hipMemcpyAsync(rel_perm, c->rel_perm, 8 * N, hipMemcpyDeviceToHost, c->stream));
  }
  if (surf_head) HIPCHK(hipMemcpyAsync(surf_head, c->surf_head, 8 * S, hipMemcpyDeviceToHost, c->stream));
  if (surf_rati) HIPCHK(hipMemcpyAsync(surf_rati, c->surf_rati, 8 * S, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_calc_srat_rperm(msg123_ctx *c, double pertur, const double *pres,
                                      double *srat, double *rperm, double *sstor) {
  if (!c || !pres) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n;
  int rc;
  if ((rc = h2d(c, c->jvec, pres, N))) return rc;
  hipLaunchKernelGGL(vg_kernel(c), dim3(c->nb), dim3(kBlock), 0, c->stream, c->dm, c->nb, c->jvec,
                     nullptr, nullptr, 0.0, nullptr, c->cell_top, soil(c), pertur, 0.0, 0,
                     c->srat_p, c->kr_p, c->f1, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  if (srat) HIPCHK(hipMemcpyAsync(srat, c->srat_p, 8 * N, hipMemcpyDeviceToHost, c->stream));
  if (rperm) HIPCHK(hipMemcpyAsync(rperm, c->kr_p, 8 * N, hipMemcpyDeviceToHost, c->stream));
  if (sstor) HIPCHK(hipMemcpyAsync(sstor, c->f1, 8 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_calc_func(msg123_ctx *c, const double *h, double *funcv) {
  if (!c || !h) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n;
  int rc;
  if (c->delt <= 0.0) return MSG123_EINVAL;
  if ((rc = h2d(c, c->jvec, h, N))) return rc;
  if ((rc = residual(c, c->jvec, false, 0.0, c->fbuf, nullptr, 0.0, nullptr))) return rc;
  if (funcv) HIPCHK(hipMemcpyAsync(funcv, c->fbuf, 8 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

// calc_vecjacf (Calc_function.f90:515-621): the directional derivative of
// the residual at x along v without the Jacobian, jv = -(F(x + eps v) -
// F(x))/eps with Brown and Saad's step eps = sign(x.v) sqrt(eps_SP)
// max(|x.v|, |v|_1) / (v.v).  Like the reference, it leaves srat_new /
// rel_perm / surf_rati of the last evaluation, F(x + eps v), in the context.
extern "C" int msg123_calc_vecjacf(msg123_ctx *c, const double *x, const double *v, double *jv) {
  if (!c || !x || !v || !jv || c->delt <= 0.0) return MSG123_EINVAL;
  const int n = c->dm.n, nbl = nblk(n);
  const size_t N = (size_t)n;
  int rc;
  // x and v in the BiCGStab work vectors (free outside a solve)
  if ((rc = h2d(c, c->kr_t, x, N)) || (rc = h2d(c, c->kr_v, v, N))) return rc;
  if ((rc = residual(c, c->kr_t, false, 0.0, c->f1, nullptr, 0.0, nullptr))) return rc;
  double d3[3];  // v.v, x.v, |v|_1
  if ((rc = dots3(c, c->kr_v, c->kr_t, d3))) return rc;
  const double l2_v = d3[0], l2_x = d3[1], l1_v = d3[2];
  const double sgn = (l2_x < 0.0) ? -1.0 : 1.0;
  const double eps = sgn * std::sqrt(MACHI_EPS) * std::fmax(std::fabs(l2_x), l1_v) / l2_v;
  const double eps_inv = 1.0 / eps;
  hipLaunchKernelGGL(k_axpy, dim3(nbl), dim3(kBlock), 0, c->stream, n, c->kr_t, eps, c->kr_v,
                     c->jvec);
  HIPCHK(hipGetLastError());
  double red;
  if ((rc = residual(c, c->jvec, false, 0.0, c->fbuf, c->f1, eps_inv, &red))) return rc;
  HIPCHK(hipMemcpyAsync(jv, c->fbuf, 8 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

// write_3dbin's value, real(out_val(i)*out_unit, SP) with out_unit =
// len_scal (Write_module.f90:170): the product in FP64, rounded once to FP32
// on the device so that only 4 bytes per cell cross PCIe.
__global__ void k_head_out(int n, const double *__restrict__ h, double len_scal,
                           float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)(h[i] * len_scal);
}

namespace {
// A stream-unformatted file (open_new_wbin, Utility_module.f90:174-175): raw
// little-endian values, no record markers.
int write_raw(const char *path, const char *mode, const void *hdr, size_t hdr_bytes,
              const void *body, size_t body_bytes) {
  FILE *f = std::fopen(path, mode);
  if (!f) return MSG123_EIO;
  const bool ok = std::fwrite(hdr, 1, hdr_bytes, f) == hdr_bytes &&
                  std::fwrite(body, 1, body_bytes, f) == body_bytes;
  return (std::fclose(f) == 0 && ok) ? MSG123_OK : MSG123_EIO;
}
}  // namespace

// write_out_headf (Write_output.f90:336-362): write_header_bin's float32 time
// (Write_module.f90:209-225) followed by write_3dbin's nx*ny*nz float32 heads
// times len_scal, x fastest (Write_module.f90:146-183).  The box mesh holds
// every grid cell in grid order, so calc2calc/get_calc_grid is the identity
// and no cell takes SNOVAL.  append = 0 starts a new file (status='replace').
extern "C" int msg123_write_head(msg123_ctx *c, const char *path, int32_t append, float time) {
  if (!c || !path) return MSG123_EINVAL;
  const int n = c->dm.n;
  float *dev = reinterpret_cast<float *>(c->jvec);  // 8n-byte scratch, free between calls
  hipLaunchKernelGGL(k_head_out, dim3(nblk(n)), dim3(kBlock), 0, c->stream, n, c->h_new,
                     c->len_scal, dev);
  HIPCHK(hipGetLastError());
  std::vector<float> host((size_t)n);
  HIPCHK(hipMemcpyAsync(host.data(), dev, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return write_raw(path, append ? "ab" : "wb", &time, 4, host.data(), 4 * (size_t)n);
}

// write_out_restf (Write_output.f90:415-432): real(time, DP), then
// head_new(i)*len_scal in FP64 for i = 1..ncalc; the file is rewritten.
extern "C" int msg123_write_restart(msg123_ctx *c, const char *path, float time) {
  if (!c || !path) return MSG123_EINVAL;
  const int n = c->dm.n;
  hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(kBlock), 0, c->stream, n, c->len_scal,
                     c->h_new, c->jvec);
  HIPCHK(hipGetLastError());
  std::vector<double> host((size_t)n);
  HIPCHK(hipMemcpyAsync(host.data(), c->jvec, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double t = (double)time;
  return write_raw(path, "wb", &t, 8, host.data(), 8 * (size_t)n);
}

// assign_initv's restart branch (Assign_calculation.f90:416-435): the FP64
// header time and ncalc FP64 values, scaled by len_scal_inv = 1/len_scal
// taken in FP32 as the reference holds it (Read_input.f90:1108).  Host only.
extern "C" int msg123_read_restart(const char *path, int32_t ncalc, double len_scal,
                                   double *heads, float *time) {
  if (!path || ncalc <= 0 || !heads || !(len_scal > 0.0)) return MSG123_EINVAL;
  FILE *f = std::fopen(path, "rb");
  if (!f) return MSG123_EIO;
  double t = 0.0;
  const bool ok = std::fread(&t, 8, 1, f) == 1 &&
                  std::fread(heads, 8, (size_t)ncalc, f) == (size_t)ncalc;
  std::fclose(f);
  if (!ok) return MSG123_EIO;
  const double inv = (double)(1.0f / (float)len_scal);
  for (int32_t i = 0; i < ncalc; i++) heads[i] = heads[i] * inv;
  if (time) *time = (float)t;
  return MSG123_OK;
}

// structured diagonals <-> reference CSR order
static void csr_from_diag(msg123_ctx *c, const std::vector<double> &A7, double *lumat) {
  const int n = c->dm.n;
  const size_t N = (size_t)n;
  for (int i = 0; i < n; i++)
    for (int k = c->offind[i]; k < c->offind[i + 1]; k++)
      lumat[k] = A7[(size_t)c->dir[k] * N + i];
}

extern "C" int msg123_make_matvec(msg123_ctx *c, double *rhs, double *dmat, double *lumat) {
  if (!c || c->delt <= 0.0) return MSG123_EINVAL;
  double nrm;
  int rc = assemble(c, &nrm);
  if (rc) return rc;
  const size_t N = (size_t)c->dm.n;
  if (rhs) HIPCHK(hipMemcpyAsync(rhs, c->rhs, 8 * N, hipMemcpyDeviceToHost, c->stream));
  if (dmat) HIPCHK(hipMemcpyAsync(dmat, c->A, 8 * N, hipMemcpyDeviceToHost, c->stream));
  if (lumat) {
    std::vector<double> A7(7 * N);
    HIPCHK(hipMemcpyAsync(A7.data(), c->A, 8 * 7 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    csr_from_diag(c, A7, lumat);
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_load_system(msg123_ctx *c, const double *rhs, const double *dmat,
                                  const double *lumat) {
  if (!c || !dmat || !lumat) return MSG123_EINVAL;
  const int n = c->dm.n;
  const size_t N = (size_t)n;
  std::vector<double> A7(7 * N, 0.0);
  std::memcpy(A7.data(), dmat, 8 * N);
  for (int i = 0; i < n; i++)
    for (int k = c->offind[i]; k < c->offind[i + 1]; k++)
      A7[(size_t)c->dir[k] * N + i] = lumat[k];
  int rc;
  if ((rc = h2d(c, c->A, A7.data(), 7 * N))) return rc;
  c->sys32 = false;
  if (rhs && (rc = h2d(c, c->rhs, rhs, N))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_precon_dilu(msg123_ctx *c, double *d) {
  if (!c) return MSG123_EINVAL;
  int rc = precon_setup(c);
  if (rc) return rc;
  if (d && !use_line32(c) && !use_amg32(c)) HIPCHK(hipMemcpyAsync(d, c->dg, 8 * (size_t)c->dm.n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return wave_check(c);
}

extern "C" int msg123_solve_ilu(msg123_ctx *c, const double *r, double *z) {
  if (!c || !r || !z) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n;
  int rc;
  if ((rc = h2d(c, c->kr_r, r, N))) return rc;
  if ((rc = ilu_apply(c, c->kr_r, c->kr_z))) return rc;
  HIPCHK(hipMemcpyAsync(z, c->kr_z, 8 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return wave_check(c);
}

extern "C" int msg123_matvec(msg123_ctx *c, const double *x, double *y) {
  if (!c || !x || !y) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n;
  int rc;
  if ((rc = h2d(c, c->kr_z, x, N))) return rc;
  hipLaunchKernelGGL(k_spmv, dim3(c->nb), dim3(kBlock), 0, c->stream, c->dm, c->nb, mat7(c),
                     c->kr_z, c->kr_v);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, c->kr_v, 8 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_solve_linalg(msg123_ctx *c, double init_norm, double errtol, double *dh,
                                   double *last_norm, int32_t *in_iter, int32_t *conv_flag) {
  if (!c || !dh) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n;
  int rc;
  if ((rc = h2d(c, c->dh, dh, N))) return rc;
  double rn;
  int it, conv;
  if ((rc = bicgstab(c, init_norm, errtol, c->prm.maxinn_iter, &rn, &it, &conv))) return rc;
  HIPCHK(hipMemcpyAsync(dh, c->dh, 8 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = wave_check(c))) return rc;
  if (last_norm) *last_norm = rn;
  if (in_iter) *in_iter = it;
  if (conv_flag) *conv_flag = conv;
  return MSG123_OK;
}

int mass_launch(msg123_ctx *c, double *part, bool law_fresh);

namespace {

bool dev_control(msg123_ctx *c) {
  if (c->control == MSG123_CONTROL_HOST || !c->dn || has_bterms(c)) return false;
  const int pc = eff_precon(c);
  return pc == MSG123_PRECON_DILU || pc == MSG123_PRECON_LINE32 || pc == MSG123_PRECON_AMG ||
         pc == MSG123_PRECON_MCILU32;
}

// The head of a device-driven Newton iteration: rhs and the Jacobian at
// head_new (head_pre becomes head_new), |rhs|^2 starting the solve, dh = 0,
// the preconditioner's set-up.
int dev_head(msg123_ctx *c) {
  const bool f32 = a32_on(c);
  if (f32) {
    if (int rc = alloc_a32(c)) return rc;
  }
  const HeadIn hin{c->h_new, nullptr, nullptr, 0.0, nullptr};
  const FusedOut o{c->rhs, f32 ? nullptr : c->A, c->srat_new, nullptr, 0.0, nullptr, c->partials,
                   nullptr, nullptr, f32 ? c->Af : nullptr};
  if (int rc = launch_fused<kAssem>(c, hin, o, "assemble")) return rc;
  c->sys32 = f32;
  std::swap(c->h_pre, c->h_new);
  hipLaunchKernelGGL(k_fold_dev<kStR0>, dim3(1), dim3(kFin), 0, c->stream, c->partials,
                     c->ftg[1].nblocks, 1, c->dn);
  hipLaunchKernelGGL(k_zero, dim3(nblk(c->dm.n)), dim3(kBlock), 0, c->stream, c->dm.n, c->dh);
  HIPCHK(hipGetLastError());
  return precon_setup(c);
}

// BiCGStab iterations it0 .. it0+iters-1 of the device-driven solve from
// x0 = 0 (solve_bicgs, Linear_solution.f90:234-493): r0 = r-hat = rhs read in
// place, the scalars in the control record, every kernel skipping once the
// solve has ended.
int dev_krylov(msg123_ctx *c, int it0, int iters) {
  const Dims dm = c->dm;
  const int n = dm.n, nb = c->nb, nbl = nblk(n);
  const Mat7 A = mat7(c);
  double *x = c->dh, *r = c->kr_r, *p = c->kr_p_, *v = c->kr_v, *z = c->kr_z, *t = c->kr_t;
  const double *rs = c->rhs;
  DevNewton *dn = c->dn;
  const int *done = &dn->done;
  int rc = MSG123_OK;
  c->skip = done;
  for (int k = 0; k < iters && rc == MSG123_OK; k++) {
    const int it = it0 + k;
    if (it > 1) {
      hipLaunchKernelGGL(k_pupdate_dev, dim3(nbl), dim3(kBlock), 0, c->stream, n, dn, r,
                         it == 2 ? rs : p, p, v);
    }
    const double *pin = it == 1 ? rs : p;
    if ((rc = ilu_apply(c, pin, z))) break;
    if (use_amg(c))
      hipLaunchKernelGGL(k_copy_skip, dim3(nbl), dim3(kBlock), 0, c->stream, n, done, pin,
                         c->amg_last);
    {
      Scope sc(c, "spmv");
      if (c->sys32)
        hipLaunchKernelGGL((k_spmv_dot<0, float>), dim3(nb), dim3(kBlock), 0, c->stream, dm, nb,
                           mat7f(c), z, v, rs, c->partials, done);
      else
        hipLaunchKernelGGL(k_spmv_dot<0>, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, A, z, v,
                           rs, c->partials, done);
    }
    hipLaunchKernelGGL(k_fold_dev<kStS1>, dim3(1), dim3(kFin), 0, c->stream, c->partials, nb, 1,
                       dn);
    hipLaunchKernelGGL(k_xr_update, dim3(nbl), dim3(kBlock), 0, c->stream, n, nbl, &dn->rho,
                       &dn->sk2, 0, x, it == 1 ? c->rhs : r, r, z, v, nullptr, c->partials, done,
                       0);
    hipLaunchKernelGGL(k_fold_dev<kStS2>, dim3(1), dim3(kFin), 0, c->stream, c->partials, nbl, 1,
                       dn);
    if ((rc = ilu_apply(c, r, z))) break;
    if (use_amg(c))
      hipLaunchKernelGGL(k_copy_skip, dim3(nbl), dim3(kBlock), 0, c->stream, n, done, r,
                         c->amg_last);
    {
      Scope sc(c, "spmv");
      if (c->sys32)
        hipLaunchKernelGGL((k_spmv_dot<1, float>), dim3(nb), dim3(kBlock), 0, c->stream, dm, nb,
                           mat7f(c), z, t, r, c->partials, done);
      else
        hipLaunchKernelGGL(k_spmv_dot<1>, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, A, z, t,
                           r, c->partials, done);
    }
    hipLaunchKernelGGL(k_fold_dev<kStS3>, dim3(1), dim3(kFin), 0, c->stream, c->partials, nb, 2,
                       dn);
    hipLaunchKernelGGL(k_xr_update, dim3(nbl), dim3(kBlock), 0, c->stream, n, nbl, &dn->ts,
                       &dn->tt, 1, x, r, r, z, t, rs, c->partials, done, 0);
    hipLaunchKernelGGL(k_fold_dev<kStS4>, dim3(1), dim3(kFin), 0, c->stream, c->partials, nbl, 2,
                       dn);
  }
  c->skip = nullptr;
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}

// The tail of a device-driven Newton iteration, run once the solve has
// ended: head_new = head_pre + dh with check_abserrmax's record, run_backtr's
// sums, the difference-quotient slope (calc_vecjacf) unless the slope is
// algebraic, F(head_new) and its norm, the decision; then one wait for the
// published record.
int dev_tail_launch(msg123_ctx *c, unsigned *seq);
int dev_tail(msg123_ctx *c, DevNewton *st) {
  unsigned seq;
  if (int rc = dev_tail_launch(c, &seq)) return rc;
  if (int rc = wait_flag(c, seq)) return rc;
  *st = *c->dn_host;
  return MSG123_OK;
}

// the kernels of dev_tail; *seq: the sequence number k_decide publishes
int dev_tail_launch(msg123_ctx *c, unsigned *seq_out) {
  const int n = c->dm.n, nbl = nblk(n);
  DevNewton *dn = c->dn;
  const int *go = &dn->done;
  const int nbm = std::min(nbl, kMaxrecBlocks);
  const bool alg = c->algebraic_slope != 0;
  hipLaunchKernelGGL(k_update_maxerr, dim3(nbm), dim3(kBlock), 0, c->stream, n, c->h_pre, c->dh,
                     c->h_new, 1, c->maxrec, go);
  hipLaunchKernelGGL(k_maxrec_dev, dim3(1), dim3(kFinMax), 0, c->stream, nbm, c->maxrec, dn);
  double *maxpart = c->partials + (size_t)(kMaxRed - 1) * nbm;
  if (alg)
    hipLaunchKernelGGL(k_backtr_sums<true>, dim3(nbm), dim3(kBlock), 0, c->stream, n, nbm, c->dh,
                       c->h_pre, c->rhs, c->kr_r, c->h_new, c->partials, maxpart, go);
  else
    hipLaunchKernelGGL(k_backtr_sums<false>, dim3(nbm), dim3(kBlock), 0, c->stream, n, nbm, c->dh,
                       c->h_pre, c->rhs, c->kr_r, c->h_new, c->partials, maxpart, go);
  hipLaunchKernelGGL(k_max_dev, dim3(1), dim3(kFinMax), 0, c->stream, nbm, maxpart, dn);
  hipLaunchKernelGGL(k_fold_dev<kStSums>, dim3(1), dim3(kFin), 0, c->stream, c->partials, nbm,
                     alg ? 5 : 3, dn);
  HIPCHK(hipGetLastError());
  if (!alg) {
    hipLaunchKernelGGL(k_fd_prep, dim3(nbl), dim3(kBlock), 0, c->stream, n, dn, c->rhs, c->h_pre,
                       c->dh, c->f1, c->jvec);
    const HeadIn hj{c->jvec, nullptr, nullptr, 0.0, nullptr};
    const FusedOut oj{c->fbuf, nullptr, c->srat_new, c->f1, 0.0,
                      use_amg(c) ? c->amg_last : c->rhs, c->partials, &dn->search_go,
                      &dn->eps_inv};
    if (int rc = launch_fused<kJv>(c, hj, oj, "residual")) return rc;
    hipLaunchKernelGGL(k_fold_dev<kStSlope>, dim3(1), dim3(kFin), 0, c->stream, c->partials,
                       c->ftg[0].nblocks, 1, dn);
  }
  const HeadIn hh{c->h_new, nullptr, nullptr, 0.0, nullptr};
  const FusedOut oh{c->fbuf, nullptr, c->srat_new, nullptr, 0.0, nullptr, c->partials, go,
                    nullptr};
  if (int rc = launch_fused<kResid>(c, hh, oh, "residual")) return rc;
  hipLaunchKernelGGL(k_fold_dev<kStL2>, dim3(1), dim3(kFin), 0, c->stream, c->partials,
                     c->ftg[0].nblocks, 1, dn);
  const unsigned seq = ++c->flag_seq;
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, c->stream, dn, diag_log(c),
                     use_amg(c) ? c->amg_last : c->rhs, c->h_new, c->dn_map, c->flag_map, seq);
  HIPCHK(hipGetLastError());
  *seq_out = seq;
  return MSG123_OK;
}

// MSG123_CONTROL_GRAPH: one chunk of a device-driven Newton iteration (with
// its head when it0 == 1) from a HIP graph.  Everything its kernels take by
// value is in the key, so a graph replays only where the same launches
// would be made; the pointer exchange of head_pre and head_new that the
// head makes on the host is repeated on the host at each replay.  The
// replay ends with a synchronisation (the sequence number k_decide
// publishes is the one of the capture).
bool graph_ok(msg123_ctx *c) {
  if (c->control != MSG123_CONTROL_GRAPH || c->timing != 0) return false;
  const int pc = eff_precon(c);
  return pc == MSG123_PRECON_LINE32 || pc == MSG123_PRECON_MCILU32 ||
         (pc == MSG123_PRECON_DILU && c->ilu_launch != MSG123_ILU_WAVE);
}

int graph_chunk(msg123_ctx *c, int it0, int chunk, DevNewton *st) {
  char key[256];
  std::snprintf(key, sizeof(key), "%d %d %d %d %d %d %d %d %d %d %d %d %p %p %p %p %p",
                it0 == 1 ? 1 : 0, it0, chunk, c->fshape[0].minw_idx, c->ftg[0].nblocks,
                c->fshape[1].minw_idx, c->ftg[1].nblocks, c->law, c->algebraic_slope,
                c->drain_on ? 1 : 0, eff_precon(c), c->ilu_launch, (void *)c->h_new,
                (void *)c->h_pre, (void *)c->A, (void *)c->rhs, (void *)c->Af);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    double *hn = c->h_new, *hp = c->h_pre;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    unsigned seq = 0;
    int rc = MSG123_OK;
    if (it0 == 1) rc = dev_head(c);
    if (!rc) rc = dev_krylov(c, it0, chunk);
    if (!rc) rc = dev_tail_launch(c, &seq);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->h_new = hn;  // the replay below makes the head's exchange
    c->h_pre = hp;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess || !g) return MSG123_EDEVICE;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return MSG123_EDEVICE;
    it = c->graphs.emplace(key, ex).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  if (it0 == 1) std::swap(c->h_pre, c->h_new);
  c->aux_stale = true;  // its fused kernels leave rel_perm and sst stale
  HIPCHK(hipStreamSynchronize(c->stream));
  *st = *c->dn_host;
  return MSG123_OK;
}

}  // namespace

// calc_numsol (Nonlinear_solution.f90:33-294), sim_type = 1
extern "C" int msg123_calc_numsol(msg123_ctx *c, msg123_newton_stats *stats) {
  if (!c || c->delt <= 0.0) return MSG123_EINVAL;
  c->conv_state_fresh = false;
  c->tail_valid = false;
  // the FP32 system before any graph capture (no allocation may happen there)
  if (a32_on(c)) {
    if (int rc = alloc_a32(c)) return rc;
  }
  const double L = c->len_scal;
  int rc;
  Newton nt;
  msg123_newton_stats ns{};
  const bool dev = dev_control(c);
  // set_backtr: maxstep = max(1, |h| * VARMAX).  Device control passes it
  // to the control record now; host control folds |h|^2 and reads it where
  // the line search first needs it, without a wait here.
  if (dev) {
    double d3[3];
    if ((rc = dots3(c, c->h_new, nullptr, d3))) return rc;
    nt.maxstep = std::sqrt(d3[0]) * VARMAX;
    if (nt.maxstep < 1.0) nt.maxstep = 1.0;
  } else {
    const int nbl = nblk(c->dm.n);
    hipLaunchKernelGGL(k_dots3, dim3(nbl), dim3(kBlock), 0, c->stream, c->dm.n, nbl, c->h_new,
                       nullptr, c->partials);
    HIPCHK(hipGetLastError());
    if ((rc = reduce_launch(c, nbl, 1, kHnormSlot, &nt.maxstep_seq))) return rc;
    nt.maxstep_pending = true;
  }
  c->spec_ready = false;
  if (dev) {
    hipLaunchKernelGGL(k_dn_init, dim3(1), dim3(64), 0, c->stream, c->dn, c->eta_first,
                       c->eta_max, c->prm.criteria, L, nt.maxstep, c->prm.maxinn_iter,
                       c->algebraic_slope);
    HIPCHK(hipGetLastError());
  }
  int conv_flag = 0, out_iter;
  double maxvar = 0.0, maxunk = 0.0;
  c->iters.clear();
  c->end_reason = MSG123_END_MAXITER;
  for (out_iter = 1; out_iter <= c->prm.maxout_iter; out_iter++) {
    if (out_iter != 1) nt.l2_pre = nt.l2_new;
    nt.back_flag = 0;
    nt.back_iter = 0;
    nt.beta_iter = 0;
    nt.moved = false;
    if (out_iter == 1) nt.eta = c->eta_first;
    int in_iter = 0, lconv = 0;
    double last = 0.0;
    // device-driven iteration: the record of its end (status), or the host
    // driver from the update on (kNeedHost), or all of it (host control)
    DevNewton st{};
    int status = -1;
    if (dev) {
      const bool graph = graph_ok(c);
      if (!graph && (rc = dev_head(c))) return rc;
      nt.func_evals += 1;
      int it0 = 1, chunk = std::max(1, std::min(c->dev_chunk, c->prm.maxinn_iter));
      for (;;) {
        if (graph) {
          if ((rc = graph_chunk(c, it0, chunk, &st))) return rc;
        } else if ((rc = dev_krylov(c, it0, chunk)) || (rc = dev_tail(c, &st))) {
          return rc;
        }
        if (st.status != kSolving) break;
        it0 += chunk;
        if (it0 > c->prm.maxinn_iter) return MSG123_EDEVICE;  // the solve always ends by then
        chunk = std::min(chunk, c->prm.maxinn_iter - it0 + 1);
      }
      in_iter = std::min(st.it, c->prm.maxinn_iter);
      c->dev_chunk = std::max(1, in_iter);
      lconv = st.conv;
      last = st.rnorm;
      c->r0_value = st.r0;
      if (out_iter == 1) nt.l2_new = st.r0;
      nt.l2_pre = st.l2_pre;
      nt.eta = st.eta;  // the tolerance this solve used (set_eater below may change it)
      status = st.status;
      if (status != kNeedHost) nt.func_evals += (st.alg || !st.searched) ? 1 : 2;
    } else {
      if (c->spec_ready) {
        // the system at head_new was assembled ahead (backtr_launch): take
        // it, as assemble() leaves it for the solve
        std::swap(c->A, c->A2);
        std::swap(c->Af, c->A2f);
        std::swap(c->rhs, c->rhs2);
        std::swap(c->h_pre, c->h_new);
        c->r0_pending = true;
        c->r0_seq = c->spec_r0_seq;
        c->spec_ready = false;
      } else {
        if ((rc = assemble(c, nullptr, true))) return rc;  // also |rhs|^2 pending
        nt.func_evals++;
      }
      // dh = 0: the AMG32 solve copies it first; the others' first update
      // writes it without reading it (dh_unset)
      if (use_amg32(c)) {
        hipLaunchKernelGGL(k_zero, dim3(std::min(c->nb, 4096)), dim3(kBlock), 0, c->stream,
                           (size_t)c->dm.n, c->dh);
        HIPCHK(hipGetLastError());
      } else {
        c->dh_unset = true;
      }
      // the tail launched behind the solve's half-steps (the look-ahead
      // paths, whose tail is tail_launch's): its buffers first
      const bool ahead = c->algebraic_slope != 0 && !use_amg(c) &&
                         (use_line32(c) || use_mc32(c));
      if (ahead && !c->A2) {
        if ((rc = dalloc(&c->A2, 7 * (size_t)c->dm.n)) ||
            (rc = dalloc(&c->rhs2, (size_t)c->dm.n)))
          return rc;
        if (hipMalloc((void **)&c->conv_run, 3 * sizeof(int)) != hipSuccess) return MSG123_ENOMEM;
      }
      if (ahead && c->sys32 && !c->A2f &&
          hipMalloc((void **)&c->A2f, sizeof(float) * 7 * (size_t)c->dm.n) != hipSuccess)
        return MSG123_ENOMEM;
      c->tail_ahead = ahead;
      c->tail_valid = false;
      // the first iteration's reference norm is |rhs|^2 (NaN: the solve reads
      // it from the assembly's fold)
      rc = bicgstab(c, out_iter == 1 ? NAN : nt.l2_pre, nt.eta, c->prm.maxinn_iter, &last,
                    &in_iter, &lconv);
      c->tail_ahead = false;
      if (rc) return rc;
      if (out_iter == 1) {
        nt.l2_new = c->r0_value;
        nt.l2_pre = nt.l2_new;
      }
    }
    nt.l2_jnorm = last;
    ns.in_iter_last = in_iter;
    ns.in_iter_total += in_iter;
    conv_flag = lconv;
    bool searched;
    int mi;
    double rec_dmat, rec_rhs, rec_head;
    if (status == kNaNStep) {
      c->end_reason = MSG123_END_NAN;
      ns.out_iter = out_iter;
      if (stats) *stats = ns;
      return MSG123_ENAN;
    } else if (status >= 0 && status != kNeedHost) {
      // the device ended the iteration: the full step was kept
      const MaxRec m = st.mr;
      c->last_maxidx = m.ich;
      maxvar = (m.ich >= 0) ? m.sch : 0.0;
      maxunk = std::fabs((m.ival >= 0) ? m.sval : 0.0) * L;
      conv_flag = st.conv_flag;
      nt.back_flag = st.back_flag;
      nt.l2_new = st.l2_new;
      nt.gradient = st.gradient;
      nt.lam = 1.0;
      nt.eta = st.eta;
      searched = st.searched != 0;
      mi = m.ich >= 0 ? m.ich : 0;
      rec_dmat = st.rec_dmat;
      rec_rhs = st.rec_rhs;
      rec_head = st.rec_head;
    } else {
      int nan_flag;
      // With the algebraic slope the convergence test's maxima, the log
      // values at their cell, and the line search's sums with F(head_new)
      // are enqueued together and read after one wait: F(head_new) is needed
      // whether the iteration converged (its norm) or not (the line search).
      const bool merged = c->algebraic_slope != 0;
      // the next iteration's system assembled ahead in place of F(head_new)
      // (line and multicolour preconditioners, which keep no pointer to A)
      const bool spec = merged && !use_amg(c) && (use_line32(c) || use_mc32(c));
      if (spec && !c->A2) {
        if ((rc = dalloc(&c->A2, 7 * (size_t)c->dm.n)) ||
            (rc = dalloc(&c->rhs2, (size_t)c->dm.n)))
          return rc;
        if (hipMalloc((void **)&c->conv_run, 3 * sizeof(int)) != hipSuccess) return MSG123_ENOMEM;
      }
      if (spec && c->sys32 && !c->A2f &&
          hipMalloc((void **)&c->A2f, sizeof(float) * 7 * (size_t)c->dm.n) != hipSuccess)
        return MSG123_ENOMEM;
      if (merged) {
        // the reference AMG path (use_amg) takes the algebraic slope only
        // through msg123_set_newton; its log reads the D-ILU diagonal
        unsigned seq;
        if (use_amg(c)) {
          if ((rc = maxerr_launch(c, true, &seq))) return rc;
          hipLaunchKernelGGL(k_pick3_rec, dim3(1), dim3(64), 0, c->stream, c->maxrec_map,
                             DiagLog{c->dg, nullptr},
                             c->amg_last, c->h_new, c->log_map);
          HIPCHK(hipGetLastError());
          if ((rc = backtr_launch(c, &seq))) return rc;
        } else if (c->tail_valid) {
          seq = c->tail_seq;  // launched behind the solve's last half-step
          c->tail_valid = false;
        } else if ((rc = tail_launch(c, &seq, spec, conv_flag))) {
          return rc;
        }
        if ((rc = wait_flag(c, seq))) return rc;
        c->backtr_pending = true;
        if ((rc = maxerr_read(c, &maxvar, &maxunk, &nan_flag))) return rc;
      } else if ((rc = maxerr(c, true, &maxvar, &maxunk, &nan_flag))) {
        return rc;
      }
      const double full_maxvar = maxvar, full_maxunk = maxunk;
      const int full_nan = nan_flag;
      double check_val = std::fabs(maxvar) * L;
      maxunk = std::fabs(maxunk) * L;
      if (conv_flag != 0 && !nan_flag) {
        conv_flag = 0;
        if (check_val <= c->prm.criteria && maxunk < XMAX) conv_flag = 1;
      } else if (nan_flag) {
        conv_flag = 0;
        c->backtr_pending = false;
        c->end_reason = MSG123_END_NAN;
        ns.out_iter = out_iter;
        if (stats) *stats = ns;
        return MSG123_ENAN;
      }
      searched = conv_flag != 1;
      if (!searched) c->backtr_pending = false;
      // converged at the full step: k_conv_state ran (the tail's fold made
      // the same decision) and left S and kr of head_new
      if (!searched && spec) {
        c->conv_state_fresh = true;
        // the assembly ahead was switched off on the device
        if (c->timing == 1) retime_last(c, "assemble", "assemble_skipped");
      }
      if (conv_flag != 1) {
        if ((rc = run_backtr(c, nt))) return rc;
        if (nt.moved) {
          if ((rc = maxerr(c, false, &maxvar, &maxunk, &nan_flag))) return rc;
        } else {
          // head_new is still h_pre + dh: check_abserrmax gives the values of
          // the update above (and the same argmax cell)
          maxvar = full_maxvar;
          maxunk = full_maxunk;
          nan_flag = full_nan;
        }
        check_val = std::fabs(maxvar) * L;
        maxunk = std::fabs(maxunk) * L;
        if (check_val >= VARMAX || maxunk >= XMAX) nt.back_flag = 1;
        else if (check_val <= c->prm.criteria && maxunk < XMAX) conv_flag = 1;
        if (conv_flag != 1 && nt.back_flag != 1) set_eater(nt, c->eta_max);
      }
      // the system assembled ahead holds where head_new is still the full
      // step and the iteration goes on
      c->spec_ready = spec && !nt.moved && conv_flag != 1 && nt.back_flag != 1;
      // the convergence-log values at the cell of the largest change
      // (Nonlinear_solution.f90:236-245); the AMG path reports the D-ILU
      // diagonal and the last preconditioner input there.
      mi = c->last_maxidx >= 0 ? c->last_maxidx : 0;
      // merged and head_new still the full step: the values were picked
      // before the wait
      const bool picked = merged && !(searched && nt.moved);
      if (!picked) {
        hipLaunchKernelGGL(k_pick3, dim3(1), dim3(64), 0, c->stream, mi, diag_log(c),
                           use_amg(c) ? c->amg_last : c->rhs,
                           c->h_new, c->log_map);
        HIPCHK(hipGetLastError());
      }
      if (searched) {
        // the line search's last residual evaluation was at the accepted
        // head_new: its norm and constitutive state are the ones needed here
        if (!picked) HIPCHK(hipStreamSynchronize(c->stream));
      } else if (merged) {
        nt.func_evals++;  // F(head_new), enqueued with the test
        nt.l2_new = c->h_scal[5];
      } else {
        nt.func_evals++;
        if ((rc = residual(c, c->h_new, false, 0.0, c->fbuf, nullptr, 0.0, &nt.l2_new))) return rc;
      }
      rec_dmat = c->h_log[0];
      rec_rhs = c->h_log[1];
      rec_head = c->h_log[2];
      if (dev) {
        // the next device-driven iteration starts from the host's values
        hipLaunchKernelGGL(k_dn_set, dim3(1), dim3(64), 0, c->stream, c->dn, nt.eta, nt.l2_new,
                           nt.l2_pre);
        HIPCHK(hipGetLastError());
      }
    }
    ns.back_iter_total += nt.back_iter;
    const double conv_var = maxvar * L;
    ns.max_change = conv_var;
    msg123_iter_rec rec;
    rec.out_iter = out_iter;
    rec.in_iter = in_iter;
    rec.back_iter = nt.back_iter;
    rec.cell_i = mi % c->dm.nx + 1;
    rec.cell_j = (mi / c->dm.nx) % c->dm.ny + 1;
    rec.cell_k = mi / c->dm.nxy + 1;
    rec.max_change = conv_var;
    rec.dmat = rec_dmat * L * L;
    rec.rhs = rec_rhs * L * L * L;
    rec.head = rec_head * L;
    c->iters.push_back(rec);
    if (c->conv_log) {
      // one line per Newton iteration (Nonlinear_solution.f90:250-252, format 11)
      char cell[48];
      std::snprintf(cell, sizeof(cell), "(%d,%d,%d)", rec.cell_i, rec.cell_j, rec.cell_k);
      std::fprintf(c->conv_log, " %6d%6d%6d%12.3E%18s%12.3E%12.3E%12.3E\n", out_iter, in_iter,
                   nt.back_iter, conv_var, cell, rec.dmat, rec.rhs, rec.head);
    }
    const char *why = nullptr;
    if (conv_flag == 1) {
      // calc_surfw: surf_head = head_new on the top layer
      if ((rc = dcopy(c, c->surf_head, c->h_new, (size_t)c->ns))) return rc;
      c->end_reason = MSG123_END_CONVERGED;
      break;
    } else if (nt.back_flag == 1) {
      c->end_reason = MSG123_END_BACKTRACK;
      why = "Stop due to maximum value or change in backtracking";
    } else if (out_iter == c->prm.maxout_iter) {
      c->end_reason = MSG123_END_MAXITER;
      why = "Stop due to maximum number of nonlinear iteration";
    } else if (std::fabs(conv_var) >= VARMAX || maxunk >= XMAX) {
      c->end_reason = MSG123_END_DIVERGED;
      why = "Didn't converge due to maximum value or change";
    }
    if (why) {
      // termination messages (Nonlinear_solution.f90:265-285, formats 12-14)
      if (c->conv_log) std::fprintf(c->conv_log, " %s\n", why);
      break;
    }
  }
  if (c->conv_log) std::fflush(c->conv_log);
  if (out_iter > c->prm.maxout_iter) out_iter = c->prm.maxout_iter;
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = wave_check(c))) return rc;
  ns.out_iter = out_iter;
  ns.conv_flag = conv_flag;
  ns.func_evals = nt.func_evals;
  ns.l2_new = nt.l2_new;
  if (stats) *stats = ns;
  return MSG123_OK;
}

constexpr int kProbeSteps = 256;  // device rows of probe heads between copies
// out[q] = h[idx[q]] (probe cells of msg123_run_box)
__global__ void k_gather_probe(int nprobe, const int32_t *__restrict__ idx,
                               const double *__restrict__ h, double *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nprobe) out[q] = h[idx[q]];
}

// h + r (h - hp); a cell unsaturated at h (below its top) moves at most half
// way to saturation: an extrapolated head that crosses into saturation sends
// the Newton iteration across the storage law's kink and costs iterations
__global__ void k_extrapolate(int n, const double *__restrict__ h, const double *__restrict__ hp,
                              double r, const double *__restrict__ top,
                              double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = h[i] + r * (h[i] - hp[i]);
  if (h[i] < top[i]) v = fmin(v, h[i] + 0.5 * (top[i] - h[i]));
  out[i] = v;
}

// Newton initial head of a step attempt from h_old (= h_n) and h_prev (=
// h_{n-1}).  MSG123_WARM_LINEAR: the linear extrapolation.  MSG123_WARM_SELECT:
// persistence or the linear extrapolation, whichever has the smaller residual
// norm.  MSG123_WARM_LEARNED: persistence or the network's correction of the
// extrapolation, whichever has the smaller residual norm.
int warm_start(msg123_ctx *c, int mode, double r, bool *took_learned) {
  const int n = c->dm.n, nbl = nblk(n);
  int rc;
  *took_learned = false;
  hipLaunchKernelGGL(k_extrapolate, dim3(nbl), dim3(kBlock), 0, c->stream, n, c->h_old,
                     c->h_prev, r, c->cell_top, c->h_lin);
  HIPCHK(hipGetLastError());
  if (mode == MSG123_WARM_LINEAR) {
    HIPCHK(hipMemcpyAsync(c->h_new, c->h_lin, 8 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    return MSG123_OK;
  }
  const double L = c->len_scal;
  const bool learned = mode == MSG123_WARM_LEARNED;
  if (learned) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if ((rc = net_features(c->net, c->h_old, c->h_prev, c->cell_top, c->lnks, r, L,
                           c->rech_ms * c->delt, c->net_x, c->h_lin, c->net_scale)) ||
        (rc = net_guess(c->net, c->net_x, c->h_lin, c->net_scale, L, c->h_old, c->cell_top,
                        c->h_try)))
      return rc;
    HIPCHK(hipStreamSynchronize(net_stream(c->net)));
  }
  double f_old, f_alt;
  const double *alt = learned ? c->h_try : c->h_lin;
  if ((rc = residual(c, c->h_old, false, 0.0, c->fbuf, nullptr, 0.0, &f_old)) ||
      (rc = residual(c, alt, false, 0.0, c->fbuf, nullptr, 0.0, &f_alt)))
    return rc;
  const bool take = f_alt < f_old;
  const double *pick = take ? alt : c->h_old;
  if (learned) {
    c->wstats.attempts++;
    c->wstats.sum_ratio += std::sqrt(f_alt / f_old);
    if (take) {
      c->wstats.learned_taken++;
      *took_learned = true;
    }
  }
  HIPCHK(hipMemcpyAsync(c->h_new, pick, 8 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
  return MSG123_OK;
}

extern "C" int msg123_set_warm_net(msg123_ctx *c, msg123_net *net, const float *lnks,
                                   double rech_ms) {
  if (!c || !net || !lnks || net_n(net) != c->dm.n) return MSG123_EINVAL;
  const size_t N = (size_t)c->dm.n;
  int rc;
  if ((!c->h_try && (rc = dalloc(&c->h_try, N))) || (!c->h_lin && (rc = dalloc(&c->h_lin, N))) ||
      (!c->h_prev && (rc = dalloc(&c->h_prev, N))) ||
      (!c->net_scale && (rc = dalloc(&c->net_scale, 1))) ||
      (!c->net_x && (rc = dalloc(&c->net_x, 4 * N))) || (!c->lnks && (rc = dalloc(&c->lnks, N))))
    return rc;
  HIPCHK(hipMemcpyAsync(c->lnks, lnks, 4 * N, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->net = net;
  c->rech_ms = rech_ms;
  return MSG123_OK;
}

extern "C" int msg123_warm_stats_get(msg123_ctx *c, msg123_warm_stats *out) {
  if (!c || !out) return MSG123_EINVAL;
  *out = c->wstats;
  out->net_on = c->tl.net_on;
  return MSG123_OK;
}

// Time_module.f90 update_tstep / calc_nextst / set_nextet / set_delt /
// set_nextvar / apply_heuri (:34-330, :770-858, :1522-1543) around
// calc_numsol, sim_type = 1, with a constant recharge or a sequence of
// uniform recharge records.  The loop state lives in the context, so
// consecutive calls continue the same run; run->reset = 1 starts again from
// t = 0.
extern "C" int msg123_run_box(msg123_ctx *c, const msg123_box_mesh *box,
                              const msg123_run_spec *run, double *heads_out, double *times_out,
                              double *dts_out, int32_t *newton_out, int32_t *inner_out) {
  if (!c || !box || !run) return MSG123_EINVAL;
  const int n = c->dm.n, ns = c->ns;
  const size_t N = (size_t)n, S = (size_t)ns;
  int rc;
  const int nrec = run->n_rech;
  if (nrec < 0 || (nrec > 0 && (!run->rech_etime || !run->rech_mmh))) return MSG123_EINVAL;
  for (int k = 0; k < nrec; k++)
    if (!std::isfinite(run->rech_etime[k]) || !std::isfinite(run->rech_mmh[k]) ||
        (k > 0 && !(run->rech_etime[k] > run->rech_etime[k - 1])))
      return MSG123_EINVAL;
  // calc_rech, the recharge volume rate per top cell: update_tstep sets it
  // from read_rech at every accepted step (Time_module.f90:81-99), and
  // read_rech changes only when set_nextvar reads a record, so it is
  // uploaded when the record changes.  A record's rate converts as
  // assign_surfbv converts it (mm/h / len_scal * 1e-3/3600 in float32).
  std::vector<double> rech(ns);
  // (k = nrec: past the last record, read_rech reset to zero)
  auto upload_rech = [&](int k) -> int {
    if (nrec > 0) {
      const float Linv = 1.0f / (float)c->len_scal;
      const float mmh = k < nrec ? run->rech_mmh[k] : 0.0f;
      const float rf = (mmh * Linv) * (1.0e-3f / 3600.0f);
      for (int s = 0; s < ns; s++) rech[s] = (double)rf * box->rech_area[s];
      c->rech_ms = (double)mmh * 1.0e-3 / 3600.0;
    } else {
      for (int s = 0; s < ns; s++) rech[s] = (double)box->read_rech[s] * box->rech_area[s];
    }
    return h2d(c, c->calc_rech, rech.data(), S);
  };
  auto &T = c->tl;
  if (run->reset || !T.started) {
    T = {};
    T.started = 1;
    c->wstats = msg123_warm_stats{};  // kept over the calls of one run
    // the first record's end, clipped to end_time (Open_file.f90:816-818)
    T.etime = (nrec > 0 && run->rech_etime[0] < run->end_time) ? run->rech_etime[0]
                                                                : run->end_time;
  }
  if (nrec > 0 && T.rk > nrec) return MSG123_EINVAL;
  if ((rc = upload_rech(T.rk))) return rc;
  const int warm = run->warm_start;
  if (warm < MSG123_WARM_PERSIST || warm > MSG123_WARM_SELECT) return MSG123_EINVAL;
  if (warm == MSG123_WARM_LEARNED && !c->net) return MSG123_EINVAL;
  if (warm != MSG123_WARM_PERSIST &&
      ((!c->h_prev && (rc = dalloc(&c->h_prev, N))) || (!c->h_lin && (rc = dalloc(&c->h_lin, N)))))
    return rc;
  const int maxout = c->prm.maxout_iter;
  const int incr_num = (int)((float)maxout * 0.4f), decr_num = (int)((float)maxout * 0.8f);
  if (run->max_fail < 0 || run->max_newton < 0) return MSG123_EINVAL;
  if (run->n_probe < 0 || (run->n_probe > 0 && (!run->probe_idx || !run->probe_out)))
    return MSG123_EINVAL;
  if (run->n_probe > 0) {
    for (int q = 0; q < run->n_probe; q++)
      if (run->probe_idx[q] < 0 || run->probe_idx[q] >= n) return MSG123_EINVAL;
    if (run->n_probe > c->n_probe_cap) {
      if (c->probe_idx) (void)hipFree(c->probe_idx);
      if (c->probe_val) (void)hipFree(c->probe_val);
      c->probe_idx = nullptr;
      c->probe_val = nullptr;
      c->n_probe_cap = 0;
      if ((rc = dalloc(&c->probe_idx, (size_t)run->n_probe)) ||
          (rc = dalloc(&c->probe_val, (size_t)run->n_probe * kProbeSteps)))
        return rc;
      c->n_probe_cap = run->n_probe;
    }
    if ((rc = h2d(c, c->probe_idx, run->probe_idx, (size_t)run->n_probe))) return rc;
  }
  const int max_fail = run->max_fail > 0 ? run->max_fail : 64;
  int nstep = 0, attempts = 0, consec_fail = 0;
  long long newton_failed = 0;
  // probe heads are gathered into kProbeSteps device rows and copied out
  // when the rows are full and at the end of the call
  int probe_base = 0;
  auto probe_flush = [&](int upto) {
    if (run->n_probe > 0 && upto > probe_base)
      HIPCHK(hipMemcpyAsync(run->probe_out + (size_t)probe_base * run->n_probe, c->probe_val,
                            8 * (size_t)run->n_probe * (size_t)(upto - probe_base),
                            hipMemcpyDeviceToHost, c->stream));
    probe_base = upto;
    return MSG123_OK;
  };
  if (T.current_t >= run->end_time && T.conv_flag == 1) return 0;
  const bool sched = run->sched_times && run->n_sched > 0;
  for (;;) {
    // calc_nextst
    if (T.current_t == 0.0f) {
      T.timestep_num = 0;
      T.delt_old1 = 0.0;
      T.have_prev = 0;
      T.delt = (double)run->init_step;
      if ((rc = h2d(c, c->h_new, box->h0, N)) || (rc = h2d(c, c->h_old, box->h0, N))) return rc;
      hipLaunchKernelGGL(vg_kernel(c), dim3(c->nb), dim3(kBlock), 0, c->stream, c->dm, c->nb, c->h_new,
                         nullptr, nullptr, 0.0, nullptr, c->cell_top, soil(c), 0.0, c->newper, 0,
                         c->srat_new, c->rel_perm, c->sst, nullptr, nullptr);
      HIPCHK(hipGetLastError());
      if ((rc = dcopy(c, c->srat_old, c->srat_new, N, c->surf_head, c->h_new, S, c->surf_old,
                      c->h_new, S)))
        return rc;
      HIPCHK(hipMemsetAsync(c->surf_rati, 0, 8 * S, c->stream));
    } else if (T.conv_flag == 1) {
      if (T.timestep_num == 0) T.delt_old1 = T.delt;
      T.timestep_num++;
      T.delt_old1 = T.delt;
      // write_bound_change: boundstep >= 1 keeps dt (Time_module.f90:219-234)
      if (T.rflag >= 1) T.delt = T.delt_old1;
      else if (T.out_iter <= incr_num) T.delt = T.delt_old1 * 1.2;
      else if (T.out_iter <= decr_num) T.delt = T.delt_old1;
      else T.delt = T.delt_old1 * 0.8;
      // the previous head for the extrapolating starts (none once the
      // learned start's gate has declined the network)
      if (warm != MSG123_WARM_PERSIST && !(warm == MSG123_WARM_LEARNED && T.net_on == 0)) {
        if ((rc = dcopy(c, c->h_prev, c->h_old, N))) return rc;
        T.have_prev = 1;
      }
      if ((rc = dcopy(c, c->h_old, c->h_new, N, c->srat_old, c->srat_new, N, c->surf_old,
                      c->surf_head, S)))
        return rc;
    } else if (T.retry_persist) {
      // the attempt that started from the network's guess failed, or was a
      // trial: the same step again from the previous head (the reference's
      // start), so the learned start never changes the step sequence
      T.current_t = T.t_start;
      if (T.shadow) c->wstats.shadows++;
      else c->wstats.retries++;
      if ((rc = dcopy(c, c->h_new, c->h_old, N, c->srat_new, c->srat_old, N, c->surf_head,
                      c->surf_old, S)))
        return rc;
    } else {
      T.current_t = T.current_t - (float)T.delt;
      T.delt = T.delt * 0.5;
      T.rflag = 0;  // reset_stepf
      if ((rc = dcopy(c, c->h_new, c->h_old, N, c->srat_new, c->srat_old, N, c->surf_head,
                      c->surf_old, S)))
        return rc;
    }
    float next_time = T.current_t + (float)T.delt;
    if (sched) {
      // prescribed schedule: aim at the next scheduled time; after a failure
      // keep the halved dt (sub-steps) until that time is reached
      if (T.sk >= run->n_sched) break;
      const float target = run->sched_times[T.sk];
      if (T.conv_flag == 1 || T.current_t == 0.0f) T.delt = (double)(target - T.current_t);
      if (T.current_t + (float)T.delt > target) T.delt = (double)(target - T.current_t);
      next_time = T.current_t + (float)T.delt;
    } else if (T.conv_flag == 1) {
      // set_nextet: past a record's end, the next record's end time; after
      // the last record, end_time and a zero recharge (reset_value)
      if (T.rflag == 1) {
        if (T.rk + 1 < nrec) {
          T.etime = run->rech_etime[T.rk + 1];
        } else {
          T.rflag = 0;
          T.etime = run->end_time;
          T.rk = nrec;
          if ((rc = upload_rech(T.rk))) return rc;
        }
      }
      // set_delt: steps end at the recharge record's end time
      const float min_step = T.etime;
      if (min_step == T.current_t) {
      } else if (min_step < next_time && next_time < run->end_time) {
        T.delt = (double)(min_step - T.current_t);
      } else if (next_time > run->end_time) {
        T.delt = (double)(run->end_time - T.current_t);
      }
      if (T.delt > (double)run->max_step) T.delt = (double)run->max_step;
      next_time = T.current_t + (float)T.delt;
      // set_nextvar: the next record's values, or the flag for a step that
      // ends at the record's end
      if (T.rflag == 1) {
        T.rk++;
        T.rflag = 0;
        if ((rc = upload_rech(T.rk))) return rc;
      } else if (T.etime == next_time) {
        T.rflag = 1;
      }
    } else if (next_time > run->end_time) {
      T.delt = (double)(run->end_time - T.current_t);
    }
    if (T.delt < DSMAL) return MSG123_ECONV;  // "Time Step is too small." (Time_module.f90:72-74)
    if (sched && nrec > 0) {
      // a prescribed step takes the record in effect at its start (none,
      // that is zero, past the last record's end)
      int k = T.rk;
      while (k < nrec && run->rech_etime[k] <= T.current_t) k++;
      if (k != T.rk) {
        T.rk = k;
        if ((rc = upload_rech(T.rk))) return rc;
      }
    }
    T.t_start = T.current_t;
    T.current_t = T.current_t + (float)T.delt;
    if ((rc = set_delt(c, T.delt))) return rc;
    bool learned_start = false;
    bool trial = false;  // a learned start tried in the shadow of persistence
    if (warm != MSG123_WARM_PERSIST && T.have_prev && !T.retry_persist) {
      bool use = true;
      if (warm == MSG123_WARM_LEARNED) {
        // A lower first residual does not mean fewer Newton iterations (an
        // extrapolated head that crosses a wetting front is "closer" and
        // harder), and a step converged from another start can leave the
        // surface cells in states that make later steps dearer.  So the
        // network is first tried on two steps in the shadow of persistence:
        // each such step is solved from the network's start, its Newton
        // iterations counted, its result discarded, and the step solved
        // again from persistence, which counts persistence's iterations on
        // the same step.  The network is used afterwards only if its trials
        // cost less than 0.8 of persistence's.  A trial is cut off once it
        // has taken 0.8 of the iterations of persistence's last accepted step
        // (at least one) without converging, and the network is then
        // declined without a second trial: a start that costs that much on
        // one step does not beat persistence by the gate's margin.
        trial = T.net_on < 0;
        use = trial || T.net_on == 1;
      }
      if (use && (rc = warm_start(c, warm, T.delt / T.delt_old1, &learned_start))) return rc;
    }
    if (c->conv_log) {
      // step header (Nonlinear_solution.f90:59-63, format 10)
      std::fprintf(c->conv_log,
                   "\n\n CURRENT TIME : %12.5E (SEC)                    TIME STEP : %12.5E (SEC)\n"
                   " ------------------------------------------------------------------------------------\n"
                   "  OUTER INNER  BACK     MAXIMUM           MAXIMUM    DIAGONAL  RIGHT HAND     UNKNOWN\n"
                   "                         CHANGE              CELL      MATRIX      VECTOR       VALUE\n"
                   " ------------------------------------------------------------------------------------\n",
                   (double)T.current_t, T.delt);
    }
    const int maxout_saved = c->prm.maxout_iter;
    const int trial_cap = std::max(1, (int)(0.8 * T.per_prev));
    const bool capped = trial && learned_start && T.per_prev > 0 && trial_cap < maxout_saved;
    if (capped) c->prm.maxout_iter = trial_cap;
    msg123_newton_stats st{};
    rc = msg123_calc_numsol(c, &st);
    c->prm.maxout_iter = maxout_saved;
    if (rc && rc != MSG123_ENAN) return rc;
    T.conv_flag = (rc == MSG123_OK) ? st.conv_flag : 0;
    T.out_iter = st.out_iter;
    const bool repeat_of_trial = T.retry_persist && T.shadow;
    T.retry_persist = (T.conv_flag != 1 && learned_start) ? 1 : 0;
    T.shadow = 0;
    if (trial) {
      if (learned_start) {
        // the network's count; converged or not, the step is solved again
        T.trial_net += st.out_iter;
        c->wstats.shadow_newton += st.out_iter;
        if (capped && T.conv_flag != 1 && st.out_iter >= trial_cap) T.net_on = 0;
        T.conv_flag = 0;
        T.retry_persist = 1;
        T.shadow = 1;
      } else {
        // persistence had the smaller residual: both starts are this one
        T.trial_net += st.out_iter;
        T.trial_per += st.out_iter;
        T.trials++;
      }
    } else if (repeat_of_trial) {
      T.trial_per += st.out_iter;
      T.trials++;
    }
    if (T.net_on < 0 && T.trials >= 2) T.net_on = T.trial_net < 0.8 * T.trial_per ? 1 : 0;
    // failure budget: consecutive failed attempts (a learned start's trial
    // or retry is not a failure of the step) and Newton iterations per call
    if (T.conv_flag == 1 || T.retry_persist) {
      consec_fail = 0;
    } else {
      consec_fail++;
      newton_failed += st.out_iter;
    }
    if (consec_fail >= max_fail || ++attempts > 100000) return MSG123_ECONV;
    if (run->max_newton > 0 && newton_failed >= run->max_newton) return MSG123_ECONV;
    if (T.conv_flag == 1 && !learned_start) T.per_prev = st.out_iter;
    if (T.conv_flag == 1 && c->mass_on) {
      // msg123_calc_mass's sums added to the device accumulator (no wait)
      if ((rc = mass_launch(c, c->partials, c->conv_state_fresh))) return rc;
      hipLaunchKernelGGL(k_mass_accum, dim3(1), dim3(kFin), 0, c->stream, c->partials, c->nb,
                         has_bterms(c) ? c->bm : nullptr, c->mass_acc);
      HIPCHK(hipGetLastError());
      c->mass_steps++;
    }
    if (T.conv_flag == 1) {
      if (sched && T.current_t >= run->sched_times[T.sk]) T.sk++;
      if (nstep < run->nstep_max) {
        if (heads_out)
          HIPCHK(hipMemcpyAsync(heads_out + (size_t)nstep * N, c->h_new, 8 * N,
                                hipMemcpyDeviceToHost, c->stream));
        if (times_out) times_out[nstep] = (double)T.current_t;
        if (dts_out) dts_out[nstep] = T.delt;
        if (newton_out) newton_out[nstep] = st.out_iter;
        if (inner_out) inner_out[nstep] = st.in_iter_total;
        if (run->n_probe > 0) {
          const int row = nstep - probe_base;
          hipLaunchKernelGGL(k_gather_probe, dim3((run->n_probe + 255) / 256), dim3(256), 0,
                             c->stream, run->n_probe, c->probe_idx, c->h_new,
                             c->probe_val + (size_t)row * run->n_probe);
          HIPCHK(hipGetLastError());
          if (row + 1 == kProbeSteps && (rc = probe_flush(nstep + 1))) return rc;
        }
      }
      nstep++;
      if (T.current_t >= run->end_time || nstep >= run->nstep_max) break;
    }
  }
  if ((rc = probe_flush(std::min(nstep, run->nstep_max)))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return nstep;
}

extern "C" int msg123_set_timing(msg123_ctx *c, int32_t enable) {
  if (!c) return MSG123_EINVAL;
  resolve_events(c);
  if (enable < 0 || enable > 2) return MSG123_EINVAL;
  c->timing = enable;
  c->timers.clear();
  return MSG123_OK;
}

extern "C" int msg123_get_timing(msg123_ctx *c, const char *name, double *total_ms,
                                 int64_t *launches) {
  if (!c || !name) return MSG123_EINVAL;
  resolve_events(c);
  auto it = c->timers.find(name);
  if (it == c->timers.end()) {
    if (total_ms) *total_ms = 0.0;
    if (launches) *launches = 0;
    return MSG123_OK;
  }
  if (total_ms) *total_ms = it->second.total_ms;
  if (launches) *launches = it->second.launches;
  return MSG123_OK;
}

extern "C" void *msg123_stream(msg123_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int msg123_set_amg_opts(msg123_ctx *c, const msg123_amg_opts *o) {
  if (!c) return MSG123_EINVAL;
  msg123_amg_opts v = o ? *o : msg123_amg_opts{12, 0.05f, 0.67f, 0.67f, 2, 2, 64, 20, 1, 1};
  if (v.max_levels < 1 || v.npre < 0 || v.npost < 0 || v.coarse_max < 1 || v.coarse_sweeps < 0 ||
      !(v.theta >= 0.0f) || !(v.omega > 0.0f) || !(v.smooth_omega > 0.0f))
    return MSG123_EINVAL;
  c->amg_opt = v;
  amg_gpu_destroy(c->amg32);
  c->amg32 = nullptr;
  return MSG123_OK;
}

extern "C" int msg123_set_newton(msg123_ctx *c, double eta_first, double eta_max,
                                 int32_t algebraic_slope, int32_t beta_refine) {
  if (!c || !(eta_first > 0.0 && eta_first < 1.0) || !(eta_max >= 1.0e-4 && eta_max < 1.0) ||
      algebraic_slope < 0 || algebraic_slope > 1 || beta_refine < 0 || beta_refine > 1)
    return MSG123_EINVAL;
  if (algebraic_slope && use_amg(c)) return MSG123_EINVAL;  // that path's slope is the reference's quirk
  c->eta_first = eta_first;
  c->eta_max = eta_max;
  c->algebraic_slope = algebraic_slope;
  c->beta_refine = beta_refine;
  return MSG123_OK;
}

extern "C" int msg123_set_ilu_launch(msg123_ctx *c, int32_t mode) {
  if (!c || (mode != MSG123_ILU_WAVE && mode != MSG123_ILU_DIAGONALS)) return MSG123_EINVAL;
  c->ilu_launch = mode;
  return MSG123_OK;
}

extern "C" int msg123_set_control(msg123_ctx *c, int32_t mode) {
  if (!c || mode < MSG123_CONTROL_HOST || mode > MSG123_CONTROL_GRAPH) return MSG123_EINVAL;
  c->control = mode;
  return MSG123_OK;
}

extern "C" int msg123_set_law(msg123_ctx *c, int32_t law) {
  if (!c || (law != MSG123_LAW_REFERENCE && law != MSG123_LAW_LOGEXP)) return MSG123_EINVAL;
  c->law = law;
  return MSG123_OK;
}

extern "C" int msg123_set_bottom_drain(msg123_ctx *c, const double *cond) {
  if (!c) return MSG123_EINVAL;
  if (!cond) {
    c->drain_on = false;
    return MSG123_OK;
  }
  const size_t nxy = (size_t)c->dm.nxy;
  for (size_t k = 0; k < nxy; k++)
    if (!(cond[k] >= 0.0) || !std::isfinite(cond[k])) return MSG123_EINVAL;
  int rc;
  if (!c->drain && (rc = dalloc(&c->drain, nxy))) return rc;
  if ((rc = h2d(c, c->drain, cond, nxy))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));  // cond is the caller's to free on return
  c->drain_on = true;
  return MSG123_OK;
}

// The boundary terms' volumes of the step at head_new (kr of head_new in
// kr_p): k_bmass into c->bm, read by bmass_read into c->bm_last.  Zeros
// without terms.
int bmass_launch(msg123_ctx *c) {
  if (!has_bterms(c)) return MSG123_OK;
  hipLaunchKernelGGL(k_bmass, dim3(1), dim3(kBcT), 0, c->stream, c->dm.nxy, bcells(c), c->rl,
                     c->n_rl_cols, c->rl_cols, c->h_new, c->kr_p, c->delt, c->bm);
  HIPCHK(hipGetLastError());
  return MSG123_OK;
}
int bmass_read(msg123_ctx *c) {
  if (!has_bterms(c)) {
    for (double &v : c->bm_last) v = 0.0;
    return MSG123_OK;
  }
  HIPCHK(hipMemcpyAsync(c->bm_last, c->bm, sizeof(c->bm_last), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MSG123_OK;
}

extern "C" int msg123_set_bterms(msg123_ctx *c, const msg123_bterms *bt) {
  if (!c) return MSG123_EINVAL;
  const int n = c->dm.n, ns = c->ns;
  const bool any = bt && (bt->nrive > 0 || bt->nlake > 0 || bt->nwell_conn > 0 || bt->nseal > 0);
  if (bt) {
    if (bt->nrive < 0 || bt->nlake < 0 || bt->nwell_conn < 0 || bt->nseal < 0 ||
        bt->nseal_levels < 0)
      return MSG123_EINVAL;
    auto in = [](const int32_t *a, int cnt, int hi) {
      for (int k = 0; k < cnt; k++)
        if (a[k] < 1 || a[k] > hi) return false;
      return true;
    };
    if (bt->nrive > 0 && (!bt->rive2cals || !bt->rive_head || !bt->rive_bott || !bt->abyd_rive ||
                          !in(bt->rive2cals, bt->nrive, ns)))
      return MSG123_EINVAL;
    if (bt->nlake > 0 && (!bt->lake2cals || !bt->lake_head || !bt->lake_bott || !bt->abyd_lake ||
                          !in(bt->lake2cals, bt->nlake, ns)))
      return MSG123_EINVAL;
    if (bt->nwell_conn > 0 &&
        (!bt->well_conn || !bt->calc_well || !in(bt->well_conn, bt->nwell_conn, n)))
      return MSG123_EINVAL;
    if (bt->nseal > 0 && (!bt->seal2calc || !bt->seal2seal || !bt->hydf_seal || !bt->abyd_seal ||
                          !bt->read_seal || !in(bt->seal2calc, bt->nseal, n) ||
                          !in(bt->seal2seal, bt->nseal, bt->nseal_levels)))
      return MSG123_EINVAL;
  }
  // the previous terms' arrays (a kernel still reading them has finished:
  // hipFree waits for the device)
  void *old[] = {c->rl, c->rl_cols, c->bc_cell, c->bc_beg, c->bc_well, c->bc_sea};
  for (void *q : old)
    if (q) (void)hipFree(q);
  c->rl = nullptr;
  c->rl_cols = c->bc_cell = c->bc_beg = nullptr;
  c->bc_well = c->bc_sea = nullptr;
  c->rl_on = false;
  c->n_rl_cols = c->bc_n = c->bc_nsea = 0;
  if (!any) return MSG123_OK;
  int rc;
  // a failed step returns after the copies already queued have finished
  // (they read host vectors of this call)
  auto fail = [&](int e) {
    (void)hipStreamSynchronize(c->stream);
    return e;
  };
  if ((!c->f_bnd && (rc = dalloc(&c->f_bnd, (size_t)n))) || (!c->bm && (rc = dalloc(&c->bm, 5))))
    return rc;
  // rivers and lakes per top cell; a cell listed again takes the later entry
  if (bt->nrive > 0 || bt->nlake > 0) {
    std::vector<double> rl(6 * (size_t)ns, 0.0);
    std::vector<char> has(ns, 0);
    const int cnt[2] = {bt->nrive, bt->nlake};
    const int32_t *cell[2] = {bt->rive2cals, bt->lake2cals};
    const double *head[2] = {bt->rive_head, bt->lake_head};
    const double *bott[2] = {bt->rive_bott, bt->lake_bott};
    const double *abyd[2] = {bt->abyd_rive, bt->abyd_lake};
    for (int w = 0; w < 2; w++)
      for (int i = 0; i < cnt[w]; i++) {
        const int sc = cell[w][i] - 1;
        double *t = rl.data() + (size_t)w * 3 * ns;
        t[sc] = c->hydf_surf_h[sc] * abyd[w][i];  // tran, in hydf_surf * abyd order
        t[ns + sc] = head[w][i];
        t[2 * (size_t)ns + sc] = bott[w][i];
        has[sc] = 1;
      }
    std::vector<int> cols;
    for (int k = 0; k < ns; k++)
      if (has[k]) cols.push_back(k);
    if ((rc = dalloc(&c->rl, rl.size())) || (rc = h2d(c, c->rl, rl.data(), rl.size())) ||
        (rc = dalloc(&c->rl_cols, cols.size())) ||
        (rc = h2d(c, c->rl_cols, cols.data(), cols.size())))
      return fail(rc);
    c->rl_on = true;
    c->n_rl_cols = (int)cols.size();
  }
  // well and sea cells, ascending, each with its sea connections in order
  if (bt->nwell_conn > 0 || bt->nseal > 0) {
    std::vector<char> well(n, 0);
    std::vector<std::vector<int>> sea_of;
    std::vector<int> slot(n, -1), cells;
    auto add = [&](int cc) {
      if (slot[cc] < 0) {
        slot[cc] = 0;
        cells.push_back(cc);
      }
    };
    for (int k = 0; k < bt->nwell_conn; k++) {
      well[bt->well_conn[k] - 1] = 1;
      add(bt->well_conn[k] - 1);
    }
    for (int i = 0; i < bt->nseal; i++) add(bt->seal2calc[i] - 1);
    std::sort(cells.begin(), cells.end());
    for (size_t e = 0; e < cells.size(); e++) slot[cells[e]] = (int)e;
    sea_of.resize(cells.size());
    for (int i = 0; i < bt->nseal; i++) sea_of[slot[bt->seal2calc[i] - 1]].push_back(i);
    const size_t M = (size_t)bt->nseal;
    std::vector<int> beg(cells.size() + 1, 0);
    std::vector<double> wv(cells.size(), 0.0), sea(3 * M, 0.0);
    size_t q = 0;
    for (size_t e = 0; e < cells.size(); e++) {
      beg[e] = (int)q;
      if (well[cells[e]]) wv[e] = bt->calc_well[cells[e]];
      for (int i : sea_of[e]) {
        sea[q] = bt->hydf_seal[i];
        sea[M + q] = bt->abyd_seal[i];
        sea[2 * M + q] = (double)bt->read_seal[bt->seal2seal[i] - 1];
        q++;
      }
    }
    beg[cells.size()] = (int)q;
    if ((rc = dalloc(&c->bc_cell, cells.size())) ||
        (rc = h2d(c, c->bc_cell, cells.data(), cells.size())) ||
        (rc = dalloc(&c->bc_beg, beg.size())) || (rc = h2d(c, c->bc_beg, beg.data(), beg.size())) ||
        (rc = dalloc(&c->bc_well, wv.size())) || (rc = h2d(c, c->bc_well, wv.data(), wv.size())) ||
        (M > 0 && ((rc = dalloc(&c->bc_sea, sea.size())) ||
                   (rc = h2d(c, c->bc_sea, sea.data(), sea.size())))))
      return fail(rc);
    c->bc_n = (int)cells.size();
    c->bc_nsea = (int)M;
  }
  HIPCHK(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  return MSG123_OK;
}

// S and kr of head_new into srat_p / kr_p (free outside the assembly; f1
// takes the unused Ss output), then k_mass with the Ss switch of both heads
// evaluated inside it, partial sums into `part`, and the boundary terms'
// sums into c->bm.
int mass_launch(msg123_ctx *c, double *part, bool law_fresh) {
  const Dims dm = c->dm;
  const int nb = c->nb;
  // law_fresh: the Newton iteration that just converged left S(head_new) in
  // srat_new and kr(head_new) in kr_p (k_conv_state), the values this pass
  // would compute
  const double *srat = law_fresh ? c->srat_new : c->srat_p;
  if (!law_fresh)
    hipLaunchKernelGGL(vg_kernel(c), dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, c->h_new,
                       nullptr, nullptr, 0.0, nullptr, c->cell_top, soil(c), 0.0, c->newper, 0,
                       c->srat_p, c->kr_p, c->f1, nullptr, nullptr);
  hipLaunchKernelGGL(k_mass, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, c->h_new, srat,
                     c->kr_p, nullptr, nullptr, c->geo, soil(c).ss, faces(c), storage(c),
                     surface(c), c->surf_old, c->delt, part);
  HIPCHK(hipGetLastError());
  return bmass_launch(c);
}

extern "C" int msg123_calc_mass(msg123_ctx *c, double out[6]) {
  if (!c || !out) return MSG123_EINVAL;
  const int nb = c->nb;
  if (int rc = mass_launch(c, c->partials, false)) return rc;
  if (int rc = reduce_finalize(c, nb, 6, out)) return rc;
  if (int rc = bmass_read(c)) return rc;
  out[4] += c->bm_last[4];
  return MSG123_OK;
}

extern "C" int msg123_calc_mass_bterms(msg123_ctx *c, double out[4]) {
  if (!c || !out) return MSG123_EINVAL;
  hipLaunchKernelGGL(vg_kernel(c), dim3(c->nb), dim3(kBlock), 0, c->stream, c->dm, c->nb,
                     c->h_new, nullptr, nullptr, 0.0, nullptr, c->cell_top, soil(c), 0.0,
                     c->newper, 0, c->srat_p, c->kr_p, c->f1, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  if (int rc = bmass_launch(c)) return rc;
  if (int rc = bmass_read(c)) return rc;
  for (int k = 0; k < 4; k++) out[k] = c->bm_last[k];
  return MSG123_OK;
}

// msg123_calc_mass without any allocation: the constitutive law of head_new
// goes to the context's work arrays allocated with it (srat_p, kr_p; f1
// takes the Ss output), the Ss switch of both heads is evaluated inside
// k_mass, the per-block partials go to the device's `scratch` and the fold
// to the context's mapped result slots.  Same kernels, same launch shape and
// the same fixed-order fold as msg123_calc_mass.
extern "C" int msg123_calc_mass_noalloc(msg123_ctx *c, double out[6]) {
  if (!c || !out) return MSG123_EINVAL;
  if (has_bterms(c)) {
    std::fprintf(stderr, "msg123_calc_mass_noalloc: contexts with boundary terms "
                         "(msg123_set_bterms) need msg123_calc_mass\n");
    return MSG123_EINVAL;
  }
  if (c->nb > kNoallocMaxBlocks) {
    std::fprintf(stderr,
                 "msg123_calc_mass_noalloc: %d cells exceed the scratch limit of %d cells\n",
                 c->dm.n, kNoallocMaxBlocks * kBlock);
    return MSG123_EINVAL;
  }
  double *part = scratch_ptr(c->device);
  if (!part) return MSG123_EDEVICE;
  const Dims dm = c->dm;
  const int nb = c->nb;
  hipLaunchKernelGGL(vg_kernel(c), dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, c->h_new, nullptr,
                     nullptr, 0.0, nullptr, c->cell_top, soil(c), 0.0, c->newper, 0, c->srat_p,
                     c->kr_p, c->f1, nullptr, nullptr);
  hipLaunchKernelGGL(k_mass, dim3(nb), dim3(kBlock), 0, c->stream, dm, nb, c->h_new, c->srat_p,
                     c->kr_p, nullptr, nullptr, c->geo, soil(c).ss, faces(c), storage(c),
                     surface(c), c->surf_old, c->delt, part);
  HIPCHK(hipGetLastError());
  const unsigned seq = ++c->flag_seq;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFin), 0, c->stream, part, nb, 6, c->scal,
                     c->scal_dev, c->flag_map, seq);
  HIPCHK(hipGetLastError());
  if (int rc = wait_flag(c, seq)) return rc;
  for (int k = 0; k < 6; k++) out[k] = c->h_scal[k];
  return MSG123_OK;
}

// Launch-shape tuner of the fused stencil kernels: every candidate (the
// occupancy the compiler is held to x the block count the z chunking aims
// at) runs `reps` times on the context's current state, timed with HIP events
// on the context's stream, and the fastest becomes the context's shape.  All
// candidates compute the same per-cell arithmetic; only the number of block
// partial sums of the norms changes.
extern "C" int msg123_autotune(msg123_ctx *c, int32_t reps, msg123_tune_report *rep) {
  if (!c || reps < 1 || c->delt <= 0.0) return MSG123_EINVAL;
  static const int targets[] = {512, 1024, 2048, 4096, 8192};
  msg123_tune_report r{};
  hipEvent_t e0 = take_event(c), e1 = take_event(c);
  for (int kind = 0; kind < 2; kind++) {
    double best = 1e300, dflt = 0.0;
    FusedShape best_shape;
    std::vector<int> seen;
    int cand = 0;
    for (int w = 0; w < kNumMinWaves; w++) {
      seen.clear();
      for (int target : targets) {
        const Tiling tg = tiling_shape<kFTX, kFTY>(c->dm, target);
        if (std::find(seen.begin(), seen.end(), tg.nblocks) != seen.end()) continue;
        seen.push_back(tg.nblocks);
        auto run = [&]() {
          if (kind == 0) {
            const HeadIn hin{c->h_new, nullptr, nullptr, 0.0, nullptr};
            const FusedOut o{c->fbuf, nullptr, c->srat_new, nullptr, 0.0, nullptr, c->partials};
            return launch_fused_shape<kResid>(c, hin, o, w, tg);
          }
          // the store precision calc_numsol's systems take
          const bool f32 = a32_on(c) && c->Af;
          const HeadIn hin{c->h_new, nullptr, nullptr, 0.0, nullptr};
          const FusedOut o{c->rhs, f32 ? nullptr : c->A, c->srat_new, nullptr, 0.0, nullptr,
                           c->partials, nullptr, nullptr, f32 ? c->Af : nullptr};
          c->sys32 = f32;
          return launch_fused_shape<kAssem>(c, hin, o, w, tg);
        };
        if (int rc = run()) return rc;  // warm-up
        HIPCHK(hipEventRecord(e0, c->stream));
        for (int k = 0; k < reps; k++)
          if (int rc = run()) return rc;
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * (double)ms / reps;
        cand++;
        if (w == 0 && target == kFusedBlocks) dflt = us;
        if (us < best) {
          best = us;
          best_shape.minw_idx = w;
          best_shape.target = target;
        }
      }
    }
    c->fshape[kind] = best_shape;
    c->ftg[kind] = tiling_shape<kFTX, kFTY>(c->dm, best_shape.target);
    if (kind == 0) {
      r.residual_waves = kMinWaves[best_shape.minw_idx];
      r.residual_blocks = c->ftg[0].nblocks;
      r.residual_us = best;
      r.residual_default_us = dflt;
    } else {
      r.assembly_waves = kMinWaves[best_shape.minw_idx];
      r.assembly_blocks = c->ftg[1].nblocks;
      r.assembly_us = best;
      r.assembly_default_us = dflt;
    }
    r.candidates += cand;
  }
  c->event_pool.push_back(e0);
  c->event_pool.push_back(e1);
  if (rep) *rep = r;
  return MSG123_OK;
}

extern "C" int msg123_set_mass_balance(msg123_ctx *c, int32_t enable) {
  if (!c) return MSG123_EINVAL;
  if (enable && !c->mass_acc && hipMalloc((void **)&c->mass_acc, sizeof(double) * kMassAcc) != hipSuccess)
    return MSG123_ENOMEM;
  c->mass_on = enable != 0;
  c->mass_steps = 0;
  if (c->mass_acc) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, c->stream, kMassAcc, c->mass_acc);
    HIPCHK(hipGetLastError());
  }
  return MSG123_OK;
}

extern "C" int msg123_mass_balance_get(msg123_ctx *c, msg123_mass_balance *out) {
  if (!c || !out) return MSG123_EINVAL;
  msg123_mass_balance m{};
  if (c->mass_acc) {
    double a[kMassAcc];
    HIPCHK(hipMemcpyAsync(a, c->mass_acc, sizeof(a), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    m.storage = a[0];
    m.connection = a[1];
    m.recharge = a[2];
    m.surface = a[3];
    m.well = a[4];
    m.river = a[5];
    m.lake = a[6];
    m.sea = a[7];
    m.total = a[8];
    m.residual = a[9];
    m.recharge_abs = a[10];
  }
  m.steps = c->mass_steps;
  *out = m;
  return MSG123_OK;
}

extern "C" int msg123_amg_refresh(msg123_ctx *c) {
  if (!c) return MSG123_EINVAL;
  if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
  amg_gpu_destroy(c->amg32);
  c->amg32 = nullptr;
  return MSG123_OK;
}

extern "C" int msg123_amg_device_info(msg123_ctx *c, int32_t *nlevel, int32_t *n, int32_t *nnz,
                                      int32_t *np, int32_t cap) {
  if (!c || !nlevel) return MSG123_EINVAL;
  const int L = amg_gpu_nlevel(c->amg32);
  *nlevel = L;
  for (int l = 0; l < L && l < cap; l++)
    amg_gpu_level_size(c->amg32, l, n ? n + l : nullptr, nnz ? nnz + l : nullptr,
                       np ? np + l : nullptr);
  return MSG123_OK;
}

extern "C" int msg123_amg_stats(msg123_ctx *c, int32_t *setups, int32_t *fallbacks) {
  if (!c) return MSG123_EINVAL;
  if (setups) *setups = c->amg_setups;
  if (fallbacks) *fallbacks = c->amg_fallbacks;
  return MSG123_OK;
}

extern "C" int msg123_amg_device_level(msg123_ctx *c, int32_t level, int32_t *offind,
                                       int32_t *offrow, double *dmat, double *lumat,
                                       int32_t *pindex, int32_t *poffrow, double *pval) {
  if (!c || !c->amg32) return MSG123_EINVAL;
  return amg_gpu_level_get(c->amg32, level - 1, offind, offrow, dmat, lumat, pindex, poffrow,
                           pval, c->stream);
}

extern "C" int msg123_iter_log(msg123_ctx *c, msg123_iter_rec *recs, int32_t cap,
                               int32_t *count, int32_t *end_reason) {
  if (!c || cap < 0 || (cap > 0 && !recs)) return MSG123_EINVAL;
  const int k = std::min<int>(cap, (int)c->iters.size());
  for (int i = 0; i < k; i++) recs[i] = c->iters[i];
  if (count) *count = (int32_t)c->iters.size();
  if (end_reason) *end_reason = c->end_reason;
  return MSG123_OK;
}

extern "C" int msg123_set_conv_log(msg123_ctx *c, const char *path) {
  if (!c) return MSG123_EINVAL;
  if (c->conv_log) {
    std::fclose(c->conv_log);
    c->conv_log = nullptr;
  }
  if (path && path[0]) {
    c->conv_log = std::fopen(path, "a");
    if (!c->conv_log) return MSG123_EINVAL;
  }
  return MSG123_OK;
}
