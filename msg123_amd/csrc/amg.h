// amg.h -- host pieces of the smoothed-aggregation AMG set-up shared by the
// reference-order hierarchy (msg123_amg_setup) and the device hierarchy.
#pragma once
#include <vector>

namespace msg123 {

// Strength filter and greedy aggregation of one level: make_grapfilt and
// make_aggregation with its binary max-heap (Make_amg_matrix.f90:67-222,
// 633-768), serial, in the reference's order.  CSR as the reference's:
// offind[0..nf] row ends, offrow 1-based; every index inside is 1-based.
struct AmgAggregation {
  int nc = 0;                 // aggregates
  std::vector<int> slu;       // strong entries: 1-based CSR positions, row by row
  std::vector<int> sidx;      // [nf+1] end of row i's strong entries in slu
  std::vector<double> tdmat;  // [nf+1] filtered diagonal d + sum of weak entries
  std::vector<int> aggr;      // [nf+1] aggregate of node i, 0 when none
};
void amg_aggregate(int nf, const int *offind, const int *offrow, const double *dmat,
                   const double *lumat, double theta, AmgAggregation &a);

// Completes the aggregation: every node still outside an aggregate that has
// an aggregated strong neighbour joins the first such neighbour's aggregate,
// in node order, repeated until nothing changes.  This departs from the
// reference, whose fix-up (Make_amg_matrix.f90:173-186) stops after the first
// node; the nodes it leaves out are never corrected by the coarse levels and
// the V-cycle barely contracts.
void amg_complete_orphans(int nf, const int *offrow, AmgAggregation &a);

// Pattern of the Jacobi-smoothed prolongator (make_prolomat,
// Make_amg_matrix.f90:224-300): row i holds the aggregates of its strong
// neighbours in order of first appearance, then its own.  pcol is 1-based.
// Returns false when a row needs more than `width` entries (the reference
// stops there).
bool amg_prolong_pattern(int nf, const AmgAggregation &a, const int *offrow, int width,
                         std::vector<int> &pptr, std::vector<int> &pcol);

// R = P^T by coarse rows (trans_prolo2restr, Make_amg_matrix.f90:373-425):
// rindex[0..nc] row ends, roffrow the 1-based fine rows in ascending order,
// rpos the 0-based position of each entry in P.
void amg_transpose(int nf, int nc, const std::vector<int> &pptr, const std::vector<int> &pcol,
                   std::vector<int> &rindex, std::vector<int> &roffrow, std::vector<int> &rpos);

// Pattern of the Galerkin operator A_c = R A P in the reference's entry order
// (make_galerkin's passes with at most seven coarse partial sums per fine
// row, Make_amg_matrix.f90:427-631), without the values: coff[0..nc] row
// ends and the 1-based columns of the off-diagonal entries, and whether a
// row receives a diagonal entry (the reference leaves 1 in rows without).
void amg_galerkin_pattern(int nf, const int *offind, const int *offrow, int nc,
                          const std::vector<int> &rindex, const std::vector<int> &roffrow,
                          std::vector<int> &coff, std::vector<int> &ccol,
                          std::vector<char> &hasdiag);

}  // namespace msg123
