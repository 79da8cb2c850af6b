// mesh_box.cpp -- static inputs of the Newton-Krylov step for a uniform box.
//
// Builds, on the host, the arrays the reference's set-up produces for a
// grid of type 7 (Read_input.f90:542-566) covering the whole box
// (calc_type = 0): node coordinates scaled by len_scal (:1079-1108), cell
// geometry (Make_cell.f90:74-398), the level-1 CSR pattern with face
// coefficients (Set_condition.f90:1584-2158), scaled soil parameters
// (Assign_calculation.f90:117,259-260), recharge (Assign_boundary.f90:216)
// and the init_type 7 initial head (Assign_calculation.f90:655-680).  The
// floating-point operation order follows those routines so the arrays are
// bit-identical to the reference's; a Fortran caller that already owns
// these arrays passes its own instead.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "msg123_gpu.h"

namespace {

struct Grid {
  std::vector<double> gx, gy, gz;
  double X(int xn) const { return gx[xn - 1]; }
  double Y(int yn) const { return gy[yn - 1]; }
  double Z(int zn) const { return gz[zn - 1]; }
};

inline double sq(double x) { return x * x; }

// two-triangle quadrilateral area (Make_cell.f90 fa_2tri)
double quad_area(double x1, double x2, double x3, double x4, double y1, double y2, double y3,
                 double y4, double z1, double z2, double z3, double z4) {
  double a1 = x1 - x2, a2 = y1 - y2, a3 = z1 - z2;
  double b1 = x3 - x2, b2 = y3 - y2, b3 = z3 - z2;
  double t1 = sq(a2 * b3 - a3 * b2) + sq(a3 * b1 - a1 * b3) + sq(a1 * b2 - a2 * b1);
  a1 = x1 - x4; a2 = y1 - y4; a3 = z1 - z4;
  b1 = x3 - x4; b2 = y3 - y4; b3 = z3 - z4;
  double t2 = sq(a2 * b3 - a3 * b2) + sq(a3 * b1 - a1 * b3) + sq(a1 * b2 - a2 * b1);
  return (std::sqrt(t1) + std::sqrt(t2)) * 0.5;
}

// tetrahedron volume |(v1 x v2) . v3| / 6 (Make_cell.f90 vol)
double tet_volume(double x1, double x2, double x3, double x4, double y1, double y2, double y3,
                  double y4, double z1, double z2, double z3, double z4) {
  const double p[3] = {x2 - x3, y2 - y3, z2 - z3};
  const double q[3] = {x4 - x3, y4 - y3, z4 - z3};
  const double r[3] = {x1 - x3, y1 - y3, z1 - z3};
  const double c[3] = {p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2],
                       p[0] * q[1] - p[1] * q[0]};
  double d = 0.0;
  for (int k = 0; k < 3; k++) d += c[k] * r[k];
  return std::fabs(d) / 6.0;
}

// Heron area of the top quadrilateral (Make_cell.f90 rarea)
double plan_area(double x1, double x2, double x3, double x4, double y1, double y2, double y3,
                 double y4) {
  double s1 = std::sqrt(sq(x1 - x2) + sq(y1 - y2)), s2 = std::sqrt(sq(x2 - x3) + sq(y2 - y3));
  double s3 = std::sqrt(sq(x3 - x4) + sq(y3 - y4)), s4 = std::sqrt(sq(x4 - x1) + sq(y4 - y1));
  double dg = std::sqrt(sq(x1 - x3) + sq(y1 - y3));
  double p1 = (s1 + s2 + dg) * 0.5, p2 = (s3 + s4 + dg) * 0.5;
  return std::sqrt(p1 * (p1 - s1) * (p1 - s2) * (p1 - dg)) +
         std::sqrt(p2 * (p2 - s3) * (p2 - s4) * (p2 - dg));
}

template <typename T>
T *dup(const std::vector<T> &v) {
  T *p = static_cast<T *>(std::malloc(v.size() * sizeof(T) + 1));
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

}  // namespace

extern "C" int msg123_mesh_box(const msg123_box_spec *spec, msg123_box_mesh *out) {
  if (!spec || !out) return MSG123_EINVAL;
  std::memset(out, 0, sizeof(*out));
  const int nx = spec->nx, ny = spec->ny, nz = spec->nz;
  if (nx < 1 || ny < 1 || nz < 1) return MSG123_EINVAL;
  const long long nll = (long long)nx * ny * nz;
  if (nll > 0x7fffffffLL / 6) return MSG123_EINVAL;
  const int n = (int)nll, ns = nx * ny, nxy = nx * ny;

  // grid type 7 node coordinates in float32, promoted to double
  Grid g;
  g.gx.resize(nx + 1); g.gy.resize(ny + 1); g.gz.resize(nz + 1);
  const float dx = spec->dx, dy = spec->dy, zw = -spec->dz;
  for (int i = 1; i <= nx + 1; i++) g.gx[i - 1] = (double)(float)(0.0f + dx * (float)(i - 1));
  for (int j = 1; j <= ny + 1; j++) g.gy[j - 1] = (double)(float)(0.0f - dy * (float)(j - 1));
  for (int k = 1; k <= nz + 1; k++) g.gz[k - 1] = (double)(float)(0.0f + zw * (float)(k - 1));

  // length scale: 10**ceil(log10 max|z|), float32
  double minz = 9999.0, maxz = -9999.0;
  for (double z : g.gz) { minz = std::fmin(minz, z); maxz = std::fmax(maxz, z); }
  const double rs = std::log10(std::fmax(std::fabs(maxz), std::fabs(minz)));
  const int ip = (int)rs;
  const int e = (rs - (double)ip > 0.0) ? ip + 1 : ip;
  long long p10 = 1;
  for (int q = 0; q < e; q++) p10 *= 10;
  const float L = e < 0 ? 1.0f : (float)p10;
  const float Linv = 1.0f / L;
  for (auto &v : g.gx) v = v * (double)Linv;
  for (auto &v : g.gy) v = v * (double)Linv;
  for (auto &v : g.gz) v = v * (double)Linv;

  std::vector<double> dis((size_t)n * 6), fa((size_t)n * 6), vol(n), top(n), area_r(ns);
  for (int i = 0; i < n; i++) {
    const int xn = i % nx + 1, yn = (i / nx) % ny + 1, zn = i / nxy + 1;
    const double X0 = g.X(xn), X1 = g.X(xn + 1), Y0 = g.Y(yn), Y1 = g.Y(yn + 1);
    const double Z0 = g.Z(zn), Z1 = g.Z(zn + 1);
    // face points and cell centre, in the reference's summation order
    const double fp_xw = 0.5 * (X0 + X0), fp_yw = 0.5 * (Y0 + Y1);
    const double fp_xe = 0.5 * (X1 + X1), fp_ye = 0.5 * (Y0 + Y1);
    const double fp_xn = 0.5 * (X0 + X1), fp_yn = 0.5 * (Y0 + Y0);
    const double fp_xs = 0.5 * (X0 + X1), fp_ys = 0.5 * (Y1 + Y1);
    const double cx = 0.25 * (X0 + X0 + X1 + X1), cy = 0.25 * (Y0 + Y1 + Y0 + Y1);
    const double fp_zw = 0.25 * (Z0 + Z1 + Z0 + Z1), fp_ze = 0.25 * (Z0 + Z0 + Z1 + Z1);
    const double fp_zn = 0.25 * (Z0 + Z1 + Z0 + Z1), fp_zs = 0.25 * (Z0 + Z0 + Z1 + Z1);
    const double zt = 0.25 * (Z0 + Z0 + Z0 + Z0), zb = 0.25 * (Z1 + Z1 + Z1 + Z1);
    const double cz = 0.5 * (zt + zb);
    double *d = &dis[(size_t)i * 6];
    d[0] = zt - cz;
    d[5] = cz - zb;
    d[2] = std::sqrt(sq(cx - fp_xw) + sq(cy - fp_yw) + sq(cz - fp_zw));
    d[3] = std::sqrt(sq(cx - fp_xe) + sq(cy - fp_ye) + sq(cz - fp_ze));
    d[1] = std::sqrt(sq(cx - fp_xn) + sq(cy - fp_yn) + sq(cz - fp_zn));
    d[4] = std::sqrt(sq(cx - fp_xs) + sq(cy - fp_ys) + sq(cz - fp_zs));
    const double x1 = X0, x2 = X1, x3 = X1, x4 = X0, y1 = Y0, y2 = Y0, y3 = Y1, y4 = Y1;
    const double z1 = Z0, z2 = Z0, z3 = Z0, z4 = Z0, z5 = Z1, z6 = Z1, z7 = Z1, z8 = Z1;
    double *f = &fa[(size_t)i * 6];
    f[0] = quad_area(x1, x2, x3, x4, y1, y2, y3, y4, z1, z2, z3, z4);
    f[5] = quad_area(x1, x2, x3, x4, y1, y2, y3, y4, z5, z6, z7, z8);
    f[2] = quad_area(x1, x1, x4, x4, y1, y1, y4, y4, z1, z5, z8, z4);
    f[3] = quad_area(x2, x2, x3, x3, y2, y2, y3, y3, z2, z6, z7, z3);
    f[1] = quad_area(x1, x2, x2, x1, y1, y2, y2, y1, z1, z2, z6, z5);
    f[4] = quad_area(x4, x3, x3, x4, y4, y3, y3, y4, z4, z3, z7, z8);
    // twelve tetrahedra around the centre
    double v = tet_volume(cx, x1, x4, x4, cy, y1, y4, y4, cz, z1, z4, z8);
    v = v + tet_volume(cx, x4, x1, x1, cy, y4, y1, y1, cz, z8, z5, z1);
    v = v + tet_volume(cx, x2, x3, x3, cy, y2, y3, y3, cz, z2, z3, z7);
    v = v + tet_volume(cx, x3, x2, x2, cy, y3, y2, y2, cz, z7, z6, z2);
    v = v + tet_volume(cx, x1, x1, x2, cy, y1, y1, y2, cz, z1, z5, z6);
    v = v + tet_volume(cx, x2, x2, x1, cy, y2, y2, y1, cz, z6, z2, z1);
    v = v + tet_volume(cx, x4, x4, x3, cy, y4, y4, y3, cz, z4, z8, z7);
    v = v + tet_volume(cx, x3, x3, x4, cy, y3, y3, y4, cz, z7, z3, z4);
    v = v + tet_volume(cx, x1, x4, x3, cy, y1, y4, y3, cz, z1, z4, z3);
    v = v + tet_volume(cx, x3, x2, x1, cy, y3, y2, y1, cz, z3, z2, z1);
    v = v + tet_volume(cx, x1, x4, x3, cy, y1, y4, y3, cz, z5, z8, z7);
    v = v + tet_volume(cx, x3, x2, x1, cy, y3, y2, y1, cz, z7, z6, z5);
    vol[i] = v;
    top[i] = zt;
    if (i < ns) area_r[i] = plan_area(x1, x2, x3, x4, y1, y2, y3, y4);
  }

  // soil parameters, scaled in float32
  std::vector<float> reta(n), retn(n), resi(n), poro(n), ss(n), ck(n);
  for (int i = 0; i < n; i++) {
    reta[i] = spec->alpha[i] * L;
    retn[i] = spec->vg_n[i];
    resi[i] = spec->theta_r[i];
    poro[i] = spec->theta_s[i];
    ss[i] = spec->ss[i] * L;
    ck[i] = spec->ks[i] * Linv;
  }

  // CSR pattern: neighbour order up, N, W, E, S, down; dir 1..6
  std::vector<int> offind(n + 1, 0), offrow, dir;
  std::vector<double> hydf, abyd;
  offrow.reserve((size_t)n * 6); dir.reserve((size_t)n * 6);
  hydf.reserve((size_t)n * 6); abyd.reserve((size_t)n * 6);
  std::vector<double> hydf_surf(ns), abyd_surf(ns), surf_top(ns), surf_bott(ns), rech_area(ns);
  for (int i = 0; i < n; i++) {
    const int in = i % nx + 1, jn = (i / nx) % ny + 1, kn = i / nxy + 1;
    int nb[6] = {-1, -1, -1, -1, -1, -1};
    if (kn != 1) nb[0] = i - nxy;
    if (jn != 1) nb[1] = i - nx;
    if (in != 1) nb[2] = i - 1;
    if (in != nx) nb[3] = i + 1;
    if (jn != ny) nb[4] = i + nx;
    if (kn != nz) nb[5] = i + nxy;
    for (int f = 0; f < 6; f++) {
      if (nb[f] < 0) {
        if (f == 0 && i < ns) {
          rech_area[i] = area_r[i];
          hydf_surf[i] = (double)ck[i];
          abyd_surf[i] = fa[(size_t)i * 6] / dis[(size_t)i * 6];
        }
        continue;
      }
      const int j = nb[f];
      const double tks = (double)ck[i], aks = (double)ck[j];
      const double d1 = dis[(size_t)i * 6 + f], d2 = dis[(size_t)j * 6 + (5 - f)];
      const double d12 = d1 + d2;
      const double c1 = tks * d12, c2 = aks * d12;
      const double cd = 1.0 / (d1 * c2 + d2 * c1);
      offrow.push_back(j + 1);  // 1-based column, as crs_index(1)%offrow
      dir.push_back(f + 1);
      hydf.push_back(c1 * c2 * cd);
      abyd.push_back(fa[(size_t)i * 6 + f] / d12);
    }
    offind[i + 1] = (int)offrow.size();
  }
  for (int s = 0; s < ns; s++) { surf_top[s] = top[s]; surf_bott[s] = top[s]; }

  // recharge rate (mm/h -> L/s in float32), then h0 = surface - depth
  const float uni_conv = 1.0e-3f / 3600.0f;
  std::vector<float> rech(ns, (spec->recharge_mmh * Linv) * uni_conv);
  const float depth = spec->wt_depth * Linv;
  std::vector<double> h0(n);
  for (int i = 0; i < n; i++) h0[i] = top[i % ns] - (double)depth;

  msg123_mesh &m = out->mesh;
  m.nx = nx; m.ny = ny; m.nz = nz; m.ncalc = n; m.ncals = ns; m.nconn = (int)offrow.size();
  m.len_scal = (double)L;
  m.newper = 1.1920928955078125e-07 * (double)Linv;  // epsilon(1._SP) * len_scal_inv
  m.offind = dup(offind); m.offrow = dup(offrow); m.dir_conn = dup(dir);
  m.hydf_conn = dup(hydf); m.abyd_conn = dup(abyd);
  m.cell_vol = dup(vol); m.cell_top = dup(top);
  m.reta = dup(reta); m.retn = dup(retn); m.resi = dup(resi); m.poro = dup(poro); m.ss = dup(ss);
  m.hydf_surf = dup(hydf_surf); m.abyd_surf = dup(abyd_surf);
  m.surf_top = dup(surf_top); m.surf_bott = dup(surf_bott);
  out->rech_area = dup(rech_area);
  out->read_rech = dup(rech);
  out->h0 = dup(h0);
  return MSG123_OK;
}

extern "C" void msg123_box_mesh_free(msg123_box_mesh *b) {
  if (!b) return;
  msg123_mesh &m = b->mesh;
  std::free((void *)m.offind); std::free((void *)m.offrow); std::free((void *)m.dir_conn);
  std::free((void *)m.hydf_conn); std::free((void *)m.abyd_conn);
  std::free((void *)m.cell_vol); std::free((void *)m.cell_top);
  std::free((void *)m.reta); std::free((void *)m.retn); std::free((void *)m.resi);
  std::free((void *)m.poro); std::free((void *)m.ss);
  std::free((void *)m.hydf_surf); std::free((void *)m.abyd_surf);
  std::free((void *)m.surf_top); std::free((void *)m.surf_bott);
  std::free(b->rech_area); std::free(b->read_rech); std::free(b->h0);
  std::memset(b, 0, sizeof(*b));
}
