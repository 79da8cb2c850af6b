// amg_setup.cpp -- smoothed-aggregation AMG hierarchy on the host.
//
// Builds the same hierarchy as the reference's make_amgmat
// (Make_amg_matrix.f90:20-65), element for element, run serially (the
// reference's OpenMP loops in the setup race; its single-thread result is
// the one defined):
//   strength filter    make_grapfilt      (:67-111)
//   aggregation        make_aggregation   (:113-222) with the binary
//                      max-heap remove_que / enter_que / remove_enter_quedeep
//                      (:633-768)
//   prolongator        make_prolomat      (:224-371), Jacobi-smoothed, omega
//   restriction        trans_prolo2restr  (:373-425), R = P^T
//   Galerkin A_c = RAP make_galerkin      (:427-631), multi-pass with at most
//                      seven coarse partial sums per fine column
// Indices inside are 1-based, as in the reference, so the control flow can be
// compared line by line; the C ABI hands out 1-based column ids (like the
// reference's CSR arrays) and 0-based C arrays.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "amg.h"
#include "msg123_gpu.h"

using namespace msg123;

namespace {

struct Level {
  int n = 0;                    // unknowns
  std::vector<int> offind;      // [n+1] row ends
  std::vector<int> offrow;      // 1-based columns
  std::vector<double> dmat, lumat;
  // P (rows = finer level) and R = P^T (rows = this level)
  int nfine = 0;
  std::vector<int> pindex, poffrow, rindex, roffrow;
  std::vector<double> pval, rval;
};

// binary max-heap of the aggregation (1-based positions, holes have flag 0)
struct Heap {
  std::vector<int> flag, glob;  // priority and node at each position
  std::vector<int> &pos;        // node -> position (aggr_luflag), 0 when not queued
  int size = 0;
  explicit Heap(std::vector<int> &p) : pos(p) {}

  void remove(int n) {  // remove_que
    int q = pos[n];
    flag[q] = 0;
    glob[q] = 0;
    int c = 2 * q;
    while (c <= size) {
      if (c + 1 <= size && flag[c + 1] > flag[c]) c = c + 1;
      if (flag[c] == 0) break;
      const int g = glob[c];
      flag[q] = flag[c];
      glob[q] = g;
      pos[g] = q;
      flag[c] = 0;
      glob[c] = 0;
      q = c;
      c = 2 * q;
    }
  }
  void raise(int n) {  // enter_que: sift up after the priority grew
    int q = pos[n];
    int p = q / 2;
    while (q > 1) {
      if (flag[p] >= flag[q]) break;
      std::swap(flag[p], flag[q]);
      const int g1 = glob[q], g2 = glob[p];
      glob[p] = g1;
      glob[q] = g2;
      pos[g2] = q;
      pos[g1] = p;
      q = p;
      p = q / 2;
    }
  }
};

}  // namespace

void msg123::amg_aggregate(int nf, const int *offind, const int *offrow, const double *dmat,
                           const double *lumat, double theta, AmgAggregation &a) {
  // --- strength filter
  std::vector<double> &tdmat = a.tdmat;
  tdmat.assign(nf + 1, 0.0);
  std::vector<int> luflag(nf + 1, 0);
  std::vector<int> &sidx = a.sidx, &slu = a.slu;  // slu: 1-based CSR positions
  sidx.assign(nf + 1, 0);
  slu.clear();
  slu.reserve((size_t)nf * 6);
  const double th2 = theta * theta;
  for (int i = 1; i <= nf; i++) {
    tdmat[i] = dmat[i - 1];
    luflag[i] = 1;
    for (int k = offind[i - 1] + 1; k <= offind[i]; k++) {
      const int j = offrow[k - 1];
      const double dd = dmat[i - 1] * dmat[j - 1];
      const double lu = lumat[k - 1];
      if (lu * lu > std::fabs(dd * th2) && dd * lu > 0.0) {
        slu.push_back(k);
      } else {
        tdmat[i] = tdmat[i] + lu;
      }
    }
    sidx[i] = (int)slu.size();
    if (sidx[i] == sidx[i - 1]) luflag[i] = 0;
  }
  auto scol = [&](int s) { return offrow[slu[s - 1] - 1]; };  // column of strong entry s
  // --- aggregation
  std::vector<int> &aggr_num = a.aggr;
  aggr_num.assign(nf + 1, 0);
  std::vector<int> glob_num(nf + 1, 0), lay2(nf + 1, 0);
  Heap hp(luflag);
  hp.flag.assign(nf + 2, 0);
  hp.glob.assign(nf + 2, 0);
  for (int i = 1; i <= nf; i++) {
    if (luflag[i] == 1) {
      hp.size++;
      luflag[i] = hp.size;
      hp.flag[hp.size] = 1;
      hp.glob[hp.size] = i;
    }
  }
  int nc = 0;
  while (hp.flag[1] != 0) {
    const int i = hp.glob[1];
    nc++;
    aggr_num[i] = nc;
    for (int s = sidx[i - 1] + 1; s <= sidx[i]; s++) {
      const int j = scol(s);
      aggr_num[j] = nc;
      if (luflag[j] > 0) hp.remove(j);
      luflag[j] = 0;
      glob_num[j] = i;
    }
    hp.remove(i);
    luflag[i] = 0;
    glob_num[i] = i;
    // remove_enter_quedeep: second ring leaves the queue, third ring gains priority
    int n2 = 0;
    for (int s = sidx[i - 1] + 1; s <= sidx[i]; s++) {
      const int j = scol(s);
      for (int t = sidx[j - 1] + 1; t <= sidx[j]; t++) {
        const int l2 = scol(t);
        if (glob_num[l2] != i) {
          glob_num[l2] = i;
          lay2[++n2] = l2;
          if (luflag[l2] > 0) {
            hp.remove(l2);
            luflag[l2] = 0;
          }
        }
      }
    }
    for (int q = 1; q <= n2; q++) {
      const int j = lay2[q];
      for (int t = sidx[j - 1] + 1; t <= sidx[j]; t++) {
        const int l3 = scol(t);
        if (glob_num[l3] != i) {
          glob_num[l3] = i;
          if (luflag[l3] > 0) {
            hp.flag[luflag[l3]] = hp.flag[luflag[l3]] + 1;
            hp.raise(l3);
          }
        }
      }
    }
  }
  // orphan fix-up: the first unqueued node with an aggregated strong
  // neighbour takes that neighbour's aggregate, then the scan stops
  for (int i = 1; i <= nf; i++) {
    if (luflag[i] != 0) continue;
    bool done = false;
    for (int s = sidx[i - 1] + 1; s <= sidx[i]; s++) {
      const int j = scol(s);
      if (aggr_num[j] > 0) {
        aggr_num[i] = aggr_num[j];
        done = true;
        break;
      }
    }
    if (done) break;
  }
  a.nc = nc;
}

void msg123::amg_complete_orphans(int nf, const int *offrow, AmgAggregation &a) {
  for (bool changed = true; changed;) {
    changed = false;
    for (int i = 1; i <= nf; i++) {
      if (a.aggr[i] != 0) continue;
      for (int s = a.sidx[i - 1] + 1; s <= a.sidx[i]; s++) {
        const int g = a.aggr[offrow[a.slu[s - 1] - 1]];
        if (g > 0) {
          a.aggr[i] = g;
          changed = true;
          break;
        }
      }
    }
  }
}

bool msg123::amg_prolong_pattern(int nf, const AmgAggregation &a, const int *offrow, int width,
                                 std::vector<int> &pptr, std::vector<int> &pcol) {
  pptr.assign(nf + 1, 0);
  pcol.clear();
  std::vector<int> row;
  for (int i = 1; i <= nf; i++) {
    row.clear();
    auto add = [&](int g) {
      for (int q : row)
        if (q == g) return true;
      if ((int)row.size() >= width) return false;
      row.push_back(g);
      return true;
    };
    for (int s = a.sidx[i - 1] + 1; s <= a.sidx[i]; s++) {
      const int g = a.aggr[offrow[a.slu[s - 1] - 1]];
      if (g > 0 && !add(g)) return false;
    }
    if (a.aggr[i] > 0 && !add(a.aggr[i])) return false;
    pcol.insert(pcol.end(), row.begin(), row.end());
    pptr[i] = (int)pcol.size();
  }
  return true;
}

void msg123::amg_transpose(int nf, int nc, const std::vector<int> &pptr,
                           const std::vector<int> &pcol, std::vector<int> &rindex,
                           std::vector<int> &roffrow, std::vector<int> &rpos) {
  rindex.assign(nc + 1, 0);
  for (int q = 0; q < pptr[nf]; q++) rindex[pcol[q]]++;
  for (int k = 1; k <= nc; k++) rindex[k] += rindex[k - 1];
  std::vector<int> fill(rindex.begin(), rindex.end() - 1);
  roffrow.assign(pptr[nf], 0);
  rpos.assign(pptr[nf], 0);
  for (int i = 1; i <= nf; i++)
    for (int q = pptr[i - 1]; q < pptr[i]; q++) {
      const int k = fill[pcol[q] - 1]++;
      roffrow[k] = i;
      rpos[k] = q;
    }
}

void msg123::amg_galerkin_pattern(int nf, const int *offind, const int *offrow, int nc,
                                  const std::vector<int> &rindex,
                                  const std::vector<int> &roffrow, std::vector<int> &coff,
                                  std::vector<int> &ccol, std::vector<char> &hasdiag) {
  std::vector<int> tsize(nf + 1, 0), tnode(nc + 1, 0), taggr(nc + 1, 0);
  std::vector<int> f2c((size_t)8 * (nf + 1), 0);
  auto F2C = [&](int k, int i) -> int & { return f2c[(size_t)i * 8 + k]; };
  std::vector<int> new_row, new_col;
  coff.assign(nc + 1, 0);
  hasdiag.assign(nc + 1, 0);
  auto touch = [&](int fl, int i) {
    int kk = tsize[fl];
    if (kk == 0 || F2C(kk, fl) != i) {
      kk = kk + 1;
      tsize[fl] = kk;
      F2C(kk, fl) = i;
    }
  };
  bool fin = false;
  while (!fin) {
    fin = true;
    std::fill(tsize.begin(), tsize.end(), 0);
    for (int i = 1; i <= nc; i++) {
      if (taggr[i] > 0) continue;
      bool cancel = false;
      for (int j = rindex[i - 1] + 1; j <= rindex[i] && !cancel; j++) {
        const int fn = roffrow[j - 1];
        if (fn <= nf && tsize[fn] >= 7) cancel = true;
      }
      if (!cancel) {
        for (int j = rindex[i - 1] + 1; j <= rindex[i] && !cancel; j++) {
          const int fn = roffrow[j - 1];
          for (int k = offind[fn - 1] + 1; k <= offind[fn]; k++) {
            const int fl = offrow[k - 1];
            if (fl <= nf && tsize[fl] >= 7) {
              cancel = true;
              break;
            }
          }
        }
      }
      if (cancel) {
        fin = false;
        continue;
      }
      taggr[i] = 1;
      for (int j = rindex[i - 1] + 1; j <= rindex[i]; j++) {
        const int fn = roffrow[j - 1];
        if (fn <= nf) touch(fn, i);
        for (int k = offind[fn - 1] + 1; k <= offind[fn]; k++) {
          const int fl = offrow[k - 1];
          if (fl <= nf) touch(fl, i);
        }
      }
    }
    for (int i = 1; i <= nc; i++) {
      int ns = 0;
      for (int j = rindex[i - 1] + 1; j <= rindex[i]; j++) {
        const int fn = roffrow[j - 1];
        for (int k = 1; k <= tsize[fn]; k++) {
          const int ca = F2C(k, fn);
          int q = 1;
          for (; q <= ns; q++)
            if (tnode[q] == ca) break;
          if (q > ns) tnode[++ns] = ca;
        }
      }
      for (int q = 1; q <= ns; q++) {
        if (i != tnode[q]) coff[tnode[q]]++;  // off-diagonal count of row tnode[q]
        new_row.push_back(tnode[q]);
        new_col.push_back(i);
      }
    }
  }
  // row ends; entries of a row keep the order in which the passes produced them
  for (int i = 1; i <= nc; i++) coff[i] = coff[i] + coff[i - 1];
  ccol.assign(coff[nc], 0);
  std::vector<int> fill(coff.begin(), coff.end() - 1);  // fill[r-1] = start of row r
  for (size_t j = 0; j < new_row.size(); j++) {
    const int row = new_row[j];
    if (row != new_col[j]) {
      ccol[fill[row - 1]++] = new_col[j];
    } else {
      hasdiag[row] = 1;
    }
  }
}

namespace {

// one coarsening step from `f` (level l-1) to `c` (level l); returns 0 or an error
int coarsen(const Level &f, Level &c, int lev, double theta, double omega) {
  const int nf = f.n;
  AmgAggregation ag;
  amg_aggregate(nf, f.offind.data(), f.offrow.data(), f.dmat.data(), f.lumat.data(), theta, ag);
  const std::vector<int> &slu = ag.slu, &sidx = ag.sidx, &aggr_num = ag.aggr;
  const std::vector<double> &tdmat = ag.tdmat;
  const int nc = ag.nc;
  auto scol = [&](int s) { return f.offrow[slu[s - 1] - 1]; };
  // aggregate membership lists (aggr_inresult / aggr_belongnum)
  std::vector<int> agg_cnt(nc + 1, 0), agg_end(nc + 1, 0), agg_mem(nf + 1, 0);
  for (int i = 1; i <= nf; i++)
    if (aggr_num[i] > 0) agg_cnt[aggr_num[i]]++;
  for (int k = 1; k <= nc; k++) {
    agg_cnt[k] = agg_cnt[k] + agg_cnt[k - 1];
    agg_end[k] = agg_cnt[k - 1];
  }
  for (int i = 1; i <= nf; i++) {
    const int j = aggr_num[i];
    if (j > 0) {
      const int k = agg_end[j] + 1;
      agg_end[j] = k;
      agg_mem[k] = i;
    }
  }
  // --- prolongator
  int width = 30 + 80 * (lev - 2);
  if (width > nc) width = nc;
  std::vector<int> g2a(nf + 1, 0), prow(nf + 1, 0);
  std::vector<int> pag((size_t)(width > 0 ? width : 1) * (nf + 1), 0);
  std::vector<double> pv((size_t)(width > 0 ? width : 1) * (nf + 1), 0.0);
  auto PAG = [&](int k, int i) -> int & { return pag[(size_t)i * width + (k - 1)]; };
  auto PV = [&](int k, int i) -> double & { return pv[(size_t)i * width + (k - 1)]; };
  for (int k = 1; k <= nc; k++)
    for (int q = agg_end[k - 1] + 1; q <= agg_end[k]; q++) g2a[agg_mem[q]] = k;
  for (int i = 1; i <= nf; i++) {
    for (int s = sidx[i - 1] + 1; s <= sidx[i]; s++) {
      const int j = scol(s);
      if (g2a[j] == 0) continue;
      int k = 1;
      for (; k <= prow[i]; k++)
        if (PAG(k, i) == g2a[j]) break;
      if (k > prow[i]) {
        if (k > width) return MSG123_EINVAL;  // reference stops: prolongation size
        prow[i] = k;
        PAG(k, i) = g2a[j];
        PV(k, i) = 0.0;
      }
    }
    if (g2a[i] == 0) continue;
    int k = 1;
    for (; k <= prow[i]; k++)
      if (PAG(k, i) == g2a[i]) break;
    if (k > prow[i]) {
      if (k > width) return MSG123_EINVAL;
      prow[i] = k;
      PAG(k, i) = g2a[i];
      PV(k, i) = 0.0;
    }
  }
  for (int i = 1; i <= nf; i++) {
    if (prow[i] == 0) continue;
    for (int s = sidx[i - 1] + 1; s <= sidx[i]; s++) {
      const int j = scol(s);
      if (g2a[j] == 0) continue;
      int k = 1;
      for (; k <= prow[i]; k++)
        if (PAG(k, i) == g2a[j]) break;
      PV(k, i) = PV(k, i) - f.lumat[slu[s - 1] - 1];
    }
  }
  for (int i = 1; i <= nf; i++)
    for (int k = 1; k <= prow[i]; k++) PV(k, i) = PV(k, i) * omega / tdmat[i];
  for (int i = 1; i <= nf; i++) {
    if (g2a[i] == 0) continue;
    int k = 1;
    for (; k <= prow[i]; k++)
      if (PAG(k, i) == g2a[i]) break;
    PV(k, i) = PV(k, i) + 1.0 - omega;
  }
  c.n = nc;
  c.nfine = nf;
  c.pindex.assign(nf + 1, 0);
  for (int i = 1; i <= nf; i++) c.pindex[i] = c.pindex[i - 1] + prow[i];
  for (int i = 1; i <= nf; i++)
    for (int k = 1; k <= prow[i]; k++) {
      c.poffrow.push_back(PAG(k, i));
      c.pval.push_back(PV(k, i));
    }
  // --- restriction R = P^T
  c.rindex.assign(nc + 1, 0);
  for (int i = 1; i <= nf; i++)
    for (int j = c.pindex[i - 1] + 1; j <= c.pindex[i]; j++) {
      const int k = c.poffrow[j - 1] + 1;
      if (k <= nc) c.rindex[k] = c.rindex[k] + 1;
    }
  c.roffrow.assign(c.pindex[nf], 0);
  c.rval.assign(c.pindex[nf], 0.0);
  for (int i = 1; i <= nc; i++) c.rindex[i] = c.rindex[i] + c.rindex[i - 1];
  for (int i = 1; i <= nf; i++)
    for (int j = c.pindex[i - 1] + 1; j <= c.pindex[i]; j++) {
      const int col = c.poffrow[j - 1];
      const int k = c.rindex[col] + 1;
      c.rindex[col] = k;
      c.roffrow[k - 1] = i;
      c.rval[k - 1] = c.pval[j - 1];
    }
  // --- Galerkin product, multi-pass
  std::vector<int> tsize(nf + 1, 0), tnode(nc + 1, 0), taggr(nc + 1, 0);
  std::vector<int> f2c((size_t)8 * (nf + 1), 0);
  std::vector<double> tval((size_t)8 * (nf + 1), 0.0), tlu(nc + 1, 0.0);
  auto F2C = [&](int k, int i) -> int & { return f2c[(size_t)i * 8 + k]; };
  auto TV = [&](int k, int i) -> double & { return tval[(size_t)i * 8 + k]; };
  std::vector<int> new_row, new_col;
  std::vector<double> new_lu;
  std::vector<int> coff(nc + 1, 0);
  int ntc = 0;
  bool fin = false;
  while (!fin) {
    fin = true;
    std::fill(tsize.begin(), tsize.end(), 0);
    for (int i = 1; i <= nc; i++) {
      if (taggr[i] > 0) continue;
      bool cancel = false;
      for (int j = c.rindex[i - 1] + 1; j <= c.rindex[i] && !cancel; j++) {
        const int fn = c.roffrow[j - 1];
        if (fn <= nf && tsize[fn] >= 7) cancel = true;
      }
      if (!cancel) {
        for (int j = c.rindex[i - 1] + 1; j <= c.rindex[i] && !cancel; j++) {
          const int fn = c.roffrow[j - 1];
          for (int k = f.offind[fn - 1] + 1; k <= f.offind[fn]; k++) {
            const int fl = f.offrow[k - 1];
            if (fl <= nf && tsize[fl] >= 7) {
              cancel = true;
              break;
            }
          }
        }
      }
      if (cancel) {
        fin = false;
        continue;
      }
      taggr[i] = 1;
      for (int j = c.rindex[i - 1] + 1; j <= c.rindex[i]; j++) {
        const int fn = c.roffrow[j - 1];
        const double rv = c.rval[j - 1];
        if (fn <= nf) {
          int kk = tsize[fn];
          if (kk == 0 || F2C(kk, fn) != i) {
            kk = kk + 1;
            tsize[fn] = kk;
            F2C(kk, fn) = i;
            TV(kk, fn) = rv * f.dmat[fn - 1];
          } else {
            TV(kk, fn) = TV(kk, fn) + rv * f.dmat[fn - 1];
          }
        }
        for (int k = f.offind[fn - 1] + 1; k <= f.offind[fn]; k++) {
          const int fl = f.offrow[k - 1];
          if (fl > nf) continue;
          int kk = tsize[fl];
          if (kk == 0 || F2C(kk, fl) != i) {
            kk = kk + 1;
            tsize[fl] = kk;
            F2C(kk, fl) = i;
            TV(kk, fl) = rv * f.lumat[k - 1];
          } else {
            TV(kk, fl) = TV(kk, fl) + rv * f.lumat[k - 1];
          }
        }
      }
    }
    for (int i = 1; i <= nc; i++) {
      int ns = 0;
      for (int j = c.rindex[i - 1] + 1; j <= c.rindex[i]; j++) {
        const int fn = c.roffrow[j - 1];
        const double rv = c.rval[j - 1];
        for (int k = 1; k <= tsize[fn]; k++) {
          const int ca = F2C(k, fn);
          int q = 1;
          for (; q <= ns; q++)
            if (tnode[q] == ca) break;
          if (q > ns) tnode[++ns] = ca;
          tlu[ca] = tlu[ca] + rv * TV(k, fn);
        }
      }
      for (int q = 1; q <= ns; q++) {
        if (i != tnode[q]) {
          ntc++;
          const int k = tnode[q] + 1;
          if (k <= nc) coff[k] = coff[k] + 1;
        }
        new_row.push_back(tnode[q]);
        new_col.push_back(i);
        new_lu.push_back(tlu[tnode[q]]);
        tlu[tnode[q]] = 0.0;
      }
    }
  }
  for (int i = 1; i <= nc; i++) coff[i] = coff[i] + coff[i - 1];
  c.offrow.assign(ntc, 0);
  c.lumat.assign(ntc, 0.0);
  c.dmat.assign(nc, 1.0);
  for (size_t j = 0; j < new_row.size(); j++) {
    const int row = new_row[j];
    if (row != new_col[j]) {
      const int k = coff[row] + 1;
      coff[row] = k;
      c.offrow[k - 1] = new_col[j];
      c.lumat[k - 1] = new_lu[j];
    } else {
      c.dmat[row - 1] = new_lu[j];
    }
  }
  c.offind = coff;
  return MSG123_OK;
}

template <typename T>
T *dup(const std::vector<T> &v) {
  T *p = static_cast<T *>(std::malloc(v.size() * sizeof(T) + 1));
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

}  // namespace

extern "C" int msg123_amg_setup(const int32_t *offind, const int32_t *offrow, const double *dmat,
                                const double *lumat, int32_t n, int32_t amg_nlevel,
                                float amg_theta, float jac_omega, msg123_amg *out) {
  if (!offind || !offrow || !dmat || !lumat || !out || n < 1 || amg_nlevel < 1)
    return MSG123_EINVAL;
  std::memset(out, 0, sizeof(*out));
  std::vector<Level> lv(1);
  Level &l1 = lv[0];
  l1.n = n;
  l1.offind.assign(offind, offind + n + 1);
  l1.offrow.assign(offrow, offrow + offind[n]);
  l1.dmat.assign(dmat, dmat + n);
  l1.lumat.assign(lumat, lumat + offind[n]);
  int nlevel = 1;
  // amg_theta and jac_omega are float32 in the reference (Initial_module.f90:211)
  const double theta = (double)amg_theta, omega = (double)jac_omega;
  for (int lev = 2; lev <= amg_nlevel; lev++) {
    if (lv.back().n == 0) break;
    nlevel = lev;
    Level c;
    const int rc = coarsen(lv.back(), c, lev, theta, omega);
    if (rc) return rc;
    lv.push_back(std::move(c));
  }
  out->nlevel = nlevel;
  out->levels = static_cast<msg123_amg_level *>(std::calloc(nlevel, sizeof(msg123_amg_level)));
  for (int l = 0; l < nlevel; l++) {
    const Level &L = lv[l];
    msg123_amg_level &o = out->levels[l];
    o.n = L.n;
    o.nnz = L.offind.empty() ? 0 : L.offind[L.n];
    o.offind = dup(L.offind);
    o.offrow = dup(L.offrow);
    o.dmat = dup(L.dmat);
    o.lumat = dup(L.lumat);
    o.nfine = L.nfine;
    o.pindex = dup(L.pindex);
    o.poffrow = dup(L.poffrow);
    o.pval = dup(L.pval);
    o.rindex = dup(L.rindex);
    o.roffrow = dup(L.roffrow);
    o.rval = dup(L.rval);
  }
  return MSG123_OK;
}

extern "C" void msg123_amg_free(msg123_amg *a) {
  if (!a || !a->levels) return;
  for (int l = 0; l < a->nlevel; l++) {
    msg123_amg_level &o = a->levels[l];
    std::free(o.offind); std::free(o.offrow); std::free(o.dmat); std::free(o.lumat);
    std::free(o.pindex); std::free(o.poffrow); std::free(o.pval);
    std::free(o.rindex); std::free(o.roffrow); std::free(o.rval);
  }
  std::free(a->levels);
  std::memset(a, 0, sizeof(*a));
}
