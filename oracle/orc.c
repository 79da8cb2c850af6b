/* orc.c -- CPU restatement (oracle) of the reference Richards Newton-Krylov step.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  Every function cites the reference
 * routine it restates; arithmetic order, float32 promotion points and
 * loop orders follow the reference so that, compiled with
 * -ffp-contract=off, results match the reference build bit for bit where
 * the same libm is used.  Indices are 0-based.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DZERO 0.0
#define DONE 1.0
#define DHALF 0.5
#define DTWO 2.0
#define DQUA 0.25
#define MACHI_EPS ((double)1.1920928955078125e-07) /* epsilon(1._SP), Constval_module.f90:27 */
#define VARMAX 1.0e3
#define XMAX 1.0e4
#define DSMAL 1.0e-20
#define DINFI 1.0e20

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* ------------------------------------------------------------------ */
/* Geometry: Read_input.f90:542-566 (grid type 7), :1026-1115 (len_scal),
 * Make_cell.f90:74-398, Set_condition.f90:1584-2158.                   */

typedef struct {
  int nx, ny, nz;
  double *gx, *gy, *gz; /* node coordinates (scaled) */
} grid_t;

static double GX(const grid_t *g, int xn, int yn) { (void)yn; return g->gx[xn - 1]; }
static double GY(const grid_t *g, int xn, int yn) { (void)xn; return g->gy[yn - 1]; }
static double GZ(const grid_t *g, int xn, int yn, int zn) {
  (void)xn; (void)yn; return g->gz[zn - 1];
}

static double sq(double x) { return x * x; }

/* fa_2tri, Make_cell.f90:206-232 */
static double fa_2tri(double x1, double x2, double x3, double x4, double y1, double y2,
                      double y3, double y4, double z1, double z2, double z3, double z4) {
  double a1 = x1 - x2, a2 = y1 - y2, a3 = z1 - z2;
  double b1 = x3 - x2, b2 = y3 - y2, b3 = z3 - z2;
  double t1 = sq(a2 * b3 - a3 * b2) + sq(a3 * b1 - a1 * b3) + sq(a1 * b2 - a2 * b1);
  a1 = x1 - x4; a2 = y1 - y4; a3 = z1 - z4;
  b1 = x3 - x4; b2 = y3 - y4; b3 = z3 - z4;
  double t2 = sq(a2 * b3 - a3 * b2) + sq(a3 * b1 - a1 * b3) + sq(a1 * b2 - a2 * b1);
  return (sqrt(t1) + sqrt(t2)) * DHALF;
}

/* vol, Make_cell.f90:306-327 */
static double tvol(double x1, double x2, double x3, double x4, double y1, double y2,
                   double y3, double y4, double z1, double z2, double z3, double z4) {
  double v1[3] = {x2 - x3, y2 - y3, z2 - z3};
  double v2[3] = {x4 - x3, y4 - y3, z4 - z3};
  double v3[3] = {x1 - x3, y1 - y3, z1 - z3};
  double c[3];
  c[0] = v1[1] * v2[2] - v1[2] * v2[1];
  c[1] = v1[2] * v2[0] - v1[0] * v2[2];
  c[2] = v1[0] * v2[1] - v1[1] * v2[0];
  double d = 0.0;
  for (int k = 0; k < 3; k++) d += c[k] * v3[k];
  return fabs(d) / 6.0;
}

/* rarea, Make_cell.f90:400-421 */
static double rarea(double x1, double x2, double x3, double x4, double y1, double y2,
                    double y3, double y4) {
  double s1 = sqrt(sq(x1 - x2) + sq(y1 - y2)), s2 = sqrt(sq(x2 - x3) + sq(y2 - y3));
  double s3 = sqrt(sq(x3 - x4) + sq(y3 - y4)), s4 = sqrt(sq(x4 - x1) + sq(y4 - y1));
  double dg = sqrt(sq(x1 - x3) + sq(y1 - y3));
  double p1 = (s1 + s2 + dg) * DHALF, p2 = (s3 + s4 + dg) * DHALF;
  double t1 = sqrt(p1 * (p1 - s1) * (p1 - s2) * (p1 - dg));
  double t2 = sqrt(p2 * (p2 - s3) * (p2 - s4) * (p2 - dg));
  return t1 + t2;
}

int orc_mesh_box(orc_mesh *m, int nx, int ny, int nz, float dx, float dy, float dz,
                 const float *alpha, const float *vg_n, const float *theta_r,
                 const float *theta_s, const float *ks, const float *ss, float wt_depth,
                 float recharge_mmh) {
  memset(m, 0, sizeof(*m));
  const int n = nx * ny * nz, ns = nx * ny;
  m->nx = nx; m->ny = ny; m->nz = nz; m->n = n; m->ns = ns;

  /* grid type 7 nodes in float32, then promoted (Read_input.f90:550-564) */
  grid_t g = {nx, ny, nz, xcalloc(nx + 1, 8), xcalloc(ny + 1, 8), xcalloc(nz + 1, 8)};
  const float xc = 0.0f, yc = 0.0f, zc = 0.0f, zw = -dz;
  for (int i = 1; i <= nx + 1; i++) g.gx[i - 1] = (double)(float)(xc + dx * (float)(i - 1));
  for (int j = 1; j <= ny + 1; j++) g.gy[j - 1] = (double)(float)(yc - dy * (float)(j - 1));
  for (int k = 1; k <= nz + 1; k++) g.gz[k - 1] = (double)(float)(zc + zw * (float)(k - 1));

  /* check_grid_loc: len_scal = 10**ceil(log10 max|z|) in float32 (Read_input.f90:1079-1108) */
  double minz = 9999.0, maxz = -9999.0;
  for (int k = 0; k <= nz; k++) {
    if (g.gz[k] < minz) minz = g.gz[k];
    if (g.gz[k] > maxz) maxz = g.gz[k];
  }
  double real_scal = log10(fmax(fabs(maxz), fabs(minz)));
  int ip = (int)real_scal;
  long p10 = 1;
  int e = (real_scal - (double)ip > 0.0) ? ip + 1 : ip;
  for (int q = 0; q < e; q++) p10 *= 10;
  float len_scal = (float)p10;
  if (e < 0) len_scal = 1.0f; /* integer 10**negative is 0 in Fortran; not reached for these grids */
  float len_scal_inv = 1.0f / len_scal;
  m->len_scal = len_scal;
  m->len_scal_inv = len_scal_inv;
  for (int i = 0; i <= nx; i++) g.gx[i] = g.gx[i] * (double)len_scal_inv;
  for (int j = 0; j <= ny; j++) g.gy[j] = g.gy[j] * (double)len_scal_inv;
  for (int k = 0; k <= nz; k++) g.gz[k] = g.gz[k] * (double)len_scal_inv;
  /* newper = MACHI_EPS*len_scal_inv (Prep_calculation.f90:79) */
  m->newper = MACHI_EPS * (double)len_scal_inv;
  m->newper_inv = DONE / m->newper;

  /* geometry per cell */
  double *dis = xcalloc((size_t)n * 6, 8), *fa = xcalloc((size_t)n * 6, 8);
  m->cell_vol = xcalloc(n, 8);
  m->cell_top = xcalloc(n, 8);
  double *area_r = xcalloc(ns, 8), *fp_zt = xcalloc(n, 8);
  for (int i = 0; i < n; i++) {
    int xn = i % nx + 1, yn = (i / nx) % ny + 1, zn = i / (nx * ny) + 1;
    /* make_cent_point, Make_cell.f90:74-124 */
    double fp_xw = DHALF * (GX(&g, xn, yn) + GX(&g, xn, yn + 1));
    double fp_yw = DHALF * (GY(&g, xn, yn) + GY(&g, xn, yn + 1));
    double fp_xe = DHALF * (GX(&g, xn + 1, yn) + GX(&g, xn + 1, yn + 1));
    double fp_ye = DHALF * (GY(&g, xn + 1, yn) + GY(&g, xn + 1, yn + 1));
    double fp_xn = DHALF * (GX(&g, xn, yn) + GX(&g, xn + 1, yn));
    double fp_yn = DHALF * (GY(&g, xn, yn) + GY(&g, xn + 1, yn));
    double fp_xs = DHALF * (GX(&g, xn, yn + 1) + GX(&g, xn + 1, yn + 1));
    double fp_ys = DHALF * (GY(&g, xn, yn + 1) + GY(&g, xn + 1, yn + 1));
    double cp_x = DQUA * (GX(&g, xn, yn) + GX(&g, xn, yn + 1) + GX(&g, xn + 1, yn) +
                          GX(&g, xn + 1, yn + 1));
    double cp_y = DQUA * (GY(&g, xn, yn) + GY(&g, xn, yn + 1) + GY(&g, xn + 1, yn) +
                          GY(&g, xn + 1, yn + 1));
    double fp_zw = DQUA * (GZ(&g, xn, yn, zn) + GZ(&g, xn, yn, zn + 1) + GZ(&g, xn, yn + 1, zn) +
                           GZ(&g, xn, yn + 1, zn + 1));
    double fp_ze = DQUA * (GZ(&g, xn + 1, yn, zn) + GZ(&g, xn + 1, yn + 1, zn) +
                           GZ(&g, xn + 1, yn, zn + 1) + GZ(&g, xn + 1, yn + 1, zn + 1));
    double fp_zn = DQUA * (GZ(&g, xn, yn, zn) + GZ(&g, xn, yn, zn + 1) + GZ(&g, xn + 1, yn, zn) +
                           GZ(&g, xn + 1, yn, zn + 1));
    double fp_zs = DQUA * (GZ(&g, xn, yn + 1, zn) + GZ(&g, xn + 1, yn + 1, zn) +
                           GZ(&g, xn, yn + 1, zn + 1) + GZ(&g, xn + 1, yn + 1, zn + 1));
    double zt = DQUA * (GZ(&g, xn, yn, zn) + GZ(&g, xn, yn + 1, zn) + GZ(&g, xn + 1, yn, zn) +
                        GZ(&g, xn + 1, yn + 1, zn));
    double zb = DQUA * (GZ(&g, xn, yn, zn + 1) + GZ(&g, xn, yn + 1, zn + 1) +
                        GZ(&g, xn + 1, yn, zn + 1) + GZ(&g, xn + 1, yn + 1, zn + 1));
    double cp_z = DHALF * (zt + zb);
    fp_zt[i] = zt;
    /* make_dis_point2face, Make_cell.f90:126-162; faces 1=top 2=N 3=W 4=E 5=S 6=bottom */
    double *d = dis + (size_t)i * 6;
    d[0] = zt - cp_z;
    d[5] = cp_z - zb;
    double d3 = sq(cp_x - fp_xw) + sq(cp_y - fp_yw) + sq(cp_z - fp_zw);
    double d4 = sq(cp_x - fp_xe) + sq(cp_y - fp_ye) + sq(cp_z - fp_ze);
    double d2 = sq(cp_x - fp_xn) + sq(cp_y - fp_yn) + sq(cp_z - fp_zn);
    double d5 = sq(cp_x - fp_xs) + sq(cp_y - fp_ys) + sq(cp_z - fp_zs);
    d[2] = sqrt(d3); d[3] = sqrt(d4); d[1] = sqrt(d2); d[4] = sqrt(d5);
    /* make_face_area, Make_cell.f90:164-203 */
    double x1 = GX(&g, xn, yn), x2 = GX(&g, xn + 1, yn), x3 = GX(&g, xn + 1, yn + 1),
           x4 = GX(&g, xn, yn + 1);
    double y1 = GY(&g, xn, yn), y2 = GY(&g, xn + 1, yn), y3 = GY(&g, xn + 1, yn + 1),
           y4 = GY(&g, xn, yn + 1);
    double z1 = GZ(&g, xn, yn, zn), z2 = GZ(&g, xn + 1, yn, zn), z3 = GZ(&g, xn + 1, yn + 1, zn),
           z4 = GZ(&g, xn, yn + 1, zn);
    double z5 = GZ(&g, xn, yn, zn + 1), z6 = GZ(&g, xn + 1, yn, zn + 1),
           z7 = GZ(&g, xn + 1, yn + 1, zn + 1), z8 = GZ(&g, xn, yn + 1, zn + 1);
    double *f = fa + (size_t)i * 6;
    f[0] = fa_2tri(x1, x2, x3, x4, y1, y2, y3, y4, z1, z2, z3, z4);
    f[5] = fa_2tri(x1, x2, x3, x4, y1, y2, y3, y4, z5, z6, z7, z8);
    f[2] = fa_2tri(x1, x1, x4, x4, y1, y1, y4, y4, z1, z5, z8, z4);
    f[3] = fa_2tri(x2, x2, x3, x3, y2, y2, y3, y3, z2, z6, z7, z3);
    f[1] = fa_2tri(x1, x2, x2, x1, y1, y2, y2, y1, z1, z2, z6, z5);
    f[4] = fa_2tri(x4, x3, x3, x4, y4, y3, y3, y4, z4, z3, z7, z8);
    /* make_cell_vol, Make_cell.f90:255-304 */
    double cx = cp_x, cy = cp_y, cz = cp_z;
    double v = tvol(cx, x1, x4, x4, cy, y1, y4, y4, cz, z1, z4, z8);
    v = v + tvol(cx, x4, x1, x1, cy, y4, y1, y1, cz, z8, z5, z1);
    v = v + tvol(cx, x2, x3, x3, cy, y2, y3, y3, cz, z2, z3, z7);
    v = v + tvol(cx, x3, x2, x2, cy, y3, y2, y2, cz, z7, z6, z2);
    v = v + tvol(cx, x1, x1, x2, cy, y1, y1, y2, cz, z1, z5, z6);
    v = v + tvol(cx, x2, x2, x1, cy, y2, y2, y1, cz, z6, z2, z1);
    v = v + tvol(cx, x4, x4, x3, cy, y4, y4, y3, cz, z4, z8, z7);
    v = v + tvol(cx, x3, x3, x4, cy, y3, y3, y4, cz, z7, z3, z4);
    v = v + tvol(cx, x1, x4, x3, cy, y1, y4, y3, cz, z1, z4, z3);
    v = v + tvol(cx, x3, x2, x1, cy, y3, y2, y1, cz, z3, z2, z1);
    v = v + tvol(cx, x1, x4, x3, cy, y1, y4, y3, cz, z5, z8, z7);
    v = v + tvol(cx, x3, x2, x1, cy, y3, y2, y1, cz, z7, z6, z5);
    m->cell_vol[i] = v;
    m->cell_top[i] = zt; /* make_cell_surfelev, Make_cell.f90:329-365 */
    if (i < ns) area_r[i] = rarea(x1, x2, x3, x4, y1, y2, y3, y4); /* Make_cell.f90:367-398 */
  }

  /* scaled soil parameters, float32 arithmetic (Assign_calculation.f90:117,259-260) */
  m->reta = xcalloc(n, 4); m->retn = xcalloc(n, 4); m->resi = xcalloc(n, 4);
  m->poro = xcalloc(n, 4); m->ss = xcalloc(n, 4);
  float *ck = xcalloc(n, 4);
  for (int i = 0; i < n; i++) {
    m->reta[i] = alpha[i] * len_scal;
    m->retn[i] = vg_n[i];
    m->resi[i] = theta_r[i];
    m->poro[i] = theta_s[i];
    m->ss[i] = ss[i] * len_scal;
    ck[i] = ks[i] * len_scal_inv;
  }

  /* set_connect, Set_condition.f90:1584-2026 (isotropic Kx = Ky = Kz) */
  int *offind = xcalloc(n + 1, 4), *offrow = xcalloc((size_t)n * 6, 4), *dir = xcalloc((size_t)n * 6, 4);
  double *hydf = xcalloc((size_t)n * 6, 8), *abyd = xcalloc((size_t)n * 6, 8);
  m->hydf_surf = xcalloc(ns, 8); m->abyd_surf = xcalloc(ns, 8);
  m->surf_top = xcalloc(ns, 8); m->surf_bott = xcalloc(ns, 8); m->rech_area = xcalloc(ns, 8);
  int t = 0;
  const int nxy = nx * ny;
  for (int i = 0; i < n; i++) {
    int in = i % nx + 1, jn = (i / nx) % ny + 1, kn = i / nxy + 1;
    int nb[6] = {-1, -1, -1, -1, -1, -1};
    if (kn != 1) nb[0] = i - nxy;
    if (jn != 1) nb[1] = i - nx;
    if (in != 1) nb[2] = i - 1;
    if (in != nx) nb[3] = i + 1;
    if (jn != ny) nb[4] = i + nx;
    if (kn != nz) nb[5] = i + nxy;
    for (int f = 0; f < 6; f++) {
      if (nb[f] < 0) {
        if (f == 0 && i < ns) { /* top layer: surface coefficients (:1770-1774) */
          m->rech_area[i] = area_r[i];
          m->hydf_surf[i] = (double)ck[i];
          /* set_chabyd: abyd_surf = surf_area/surf_dis (:2092-2109) */
          m->abyd_surf[i] = fa[(size_t)i * 6 + 0] / dis[(size_t)i * 6 + 0];
        }
        continue;
      }
      int j = nb[f];
      double tks = (double)ck[i], aks = (double)ck[j];
      /* set_dis_adj (:2111-2132) */
      double d1 = dis[(size_t)i * 6 + f], d2 = dis[(size_t)j * 6 + (5 - f)], d12 = d1 + d2;
      /* set_sat_hyd (:2134-2158) */
      double dd = d1 + d2;
      double c1 = tks * dd, c2 = aks * dd;
      double cd12 = d1 * c2 + d2 * c1;
      double cd = DONE / cd12;
      offrow[t] = j;
      dir[t] = f + 1;
      hydf[t] = c1 * c2 * cd;
      abyd[t] = fa[(size_t)i * 6 + f] / d12;
      t++;
    }
    offind[i + 1] = t;
  }
  m->nconn = t;
  m->offind = offind; m->offrow = offrow; m->dir = dir; m->hydf = hydf; m->abyd = abyd;

  /* surface elevations without a geography file (Prep_calculation.f90 set_geog_info) */
  for (int s = 0; s < ns; s++) { m->surf_top[s] = fp_zt[s]; m->surf_bott[s] = fp_zt[s]; }

  /* recharge: mm/h * len_scal_inv * (1e-3/3600), float32 (Assign_boundary.f90:216,
   * Open_file.f90:744), then stored as float32 again (Calc_boundary.f90:74-78) */
  const float hoursec = 3600.0f;
  float uni_conv = 1.0e-3f / hoursec;
  m->read_rech = xcalloc(ns, 4);
  for (int s = 0; s < ns; s++) m->read_rech[s] = (float)(double)((recharge_mmh * len_scal_inv) * uni_conv);

  /* init_type 7: h = surface elevation - depth (Assign_calculation.f90:655-680) */
  float depth = wt_depth * len_scal_inv;
  m->h0 = xcalloc(n, 8);
  for (int i = 0; i < n; i++) m->h0[i] = fp_zt[i % ns] - (double)depth;

  free(g.gx); free(g.gy); free(g.gz); free(dis); free(fa); free(area_r); free(fp_zt); free(ck);
  return 0;
}

void orc_mesh_free(orc_mesh *m) {
  free(m->offind); free(m->offrow); free(m->dir); free(m->hydf); free(m->abyd);
  free(m->cell_vol); free(m->cell_top);
  free(m->reta); free(m->retn); free(m->resi); free(m->poro); free(m->ss);
  free(m->hydf_surf); free(m->abyd_surf); free(m->surf_top); free(m->surf_bott);
  free(m->rech_area); free(m->read_rech); free(m->h0);
  memset(m, 0, sizeof(*m));
}

int orc_state_init(orc_state *s, const orc_mesh *m) {
  memset(s, 0, sizeof(*s));
  int n = m->n, ns = m->ns;
  s->head_new = xcalloc(n, 8); s->head_old = xcalloc(n, 8); s->head_pre = xcalloc(n, 8);
  s->head_change = xcalloc(n, 8); s->srat_new = xcalloc(n, 8); s->srat_old = xcalloc(n, 8);
  s->rel_perm = xcalloc(n, 8); s->surf_head = xcalloc(ns, 8); s->surf_old = xcalloc(ns, 8);
  s->surf_rati = xcalloc(ns, 8); s->calc_rech = xcalloc(ns, 8); s->head_old2 = xcalloc(n, 8);
  s->rhs = xcalloc(n, 8); s->dmat = xcalloc(n, 8); s->lumat = xcalloc(m->nconn, 8);
  return 0;
}

void orc_state_free(orc_state *s) {
  free(s->head_new); free(s->head_old); free(s->head_pre); free(s->head_change);
  free(s->srat_new); free(s->srat_old); free(s->rel_perm); free(s->surf_head);
  free(s->surf_old); free(s->surf_rati); free(s->calc_rech); free(s->head_old2);
  free(s->rhs); free(s->dmat); free(s->lumat);
  memset(s, 0, sizeof(*s));
}

/* ------------------------------------------------------------------ */
/* calc_srat_rperm, Calc_parameter.f90:15-77 */
void orc_srat_rperm(const orc_mesh *m, double pertur, const double *pres, double *srat,
                    double *rperm, double *sstor) {
  const double phead0 = DZERO;
  for (int i = 0; i < m->n; i++) {
    double per_phead = pres[i] - m->cell_top[i] + pertur;
    double kr, ss;
    if (per_phead < DZERO) {
      double rn = (double)m->retn[i];
      double retm = DONE - DONE / rn;
      double beta, theta;
      if (per_phead <= phead0) {
        beta = pow(fabs(per_phead * (double)m->reta[i]), rn);
        theta = (double)m->resi[i] + (double)(m->poro[i] - m->resi[i]) * pow(DONE + beta, -retm);
      } else {
        beta = DZERO;
        theta = (double)m->poro[i];
      }
      srat[i] = theta / (double)m->poro[i];
      double se = pow(DONE + beta, -retm);
      double a = pow(se, DONE / retm);
      double c = DONE - pow(DONE - a, retm);
      kr = pow(se, DHALF) * (c * c);
      ss = DZERO;
    } else {
      srat[i] = DONE;
      kr = DONE;
      ss = (double)m->ss[i];
    }
    rperm[i] = kr;
    if (sstor) sstor[i] = ss;
  }
}

/* calc_hyd_upwind, Calc_parameter.f90:107-125 */
static inline double upwind(double head_dif, double k1, double k2) {
  return head_dif >= DZERO ? k1 : k2;
}

/* func_surfterm, Calc_function.f90:304-391 (no geography file: surf_reli = 0) */
static void func_surfterm(const orc_mesh *m, orc_state *s, const double *h, const double *sh,
                          double *surf) {
  for (int i = 0; i < m->ns; i++) {
    double delh;
    if (h[i] >= m->surf_top[i]) {
      delh = (sh[i] > m->surf_top[i]) ? sh[i] - h[i] : m->surf_top[i] - h[i];
      s->surf_rati[i] = DONE;
    } else if (h[i] >= m->surf_bott[i]) {
      delh = (sh[i] > m->surf_bott[i]) ? sh[i] - h[i] : m->surf_bott[i] - h[i];
      s->surf_rati[i] = DONE; /* surf_reli == 0 branch */
    } else if (sh[i] > m->surf_bott[i]) {
      delh = sh[i] - m->surf_bott[i];
      s->surf_rati[i] = DONE;
    } else {
      delh = DZERO;
    }
    surf[i] = m->hydf_surf[i] * m->abyd_surf[i] * delh * s->rel_perm[i] * s->surf_rati[i];
  }
}

/* func_riveterm / func_laketerm, Calc_function.f90:393-465: per entry,
 * assigned to its top cell (a later entry of the same cell replaces it) */
static void func_rlterm(const orc_mesh *m, const orc_state *s, int cnt, const int *cell,
                        const double *head, const double *bott, const double *abyd,
                        const double *h, double *out) {
  for (int i = 0; i < cnt; i++) {
    const int c = cell[i];
    double delh;
    if (h[c] >= bott[i]) delh = head[i] - h[c];
    else if (head[i] > bott[i]) delh = head[i] - bott[i];
    else delh = DZERO;
    out[c] = m->hydf_surf[c] * abyd[i] * delh * s->rel_perm[c];
  }
}

/* func_wellterm :277-302 and func_sealterm :467-513 into welf, seaf [n] */
static void func_cellterms(const orc_mesh *m, const orc_state *s, const double *h, double *welf,
                           double *seaf) {
  const orc_bterms *bt = m->bt;
  for (int k = 0; k < bt->nwell; k++) welf[bt->well_conn[k]] = bt->calc_well[bt->well_conn[k]];
  for (int i = 0; i < bt->nseal; i++) {
    const int c = bt->seal_cell[i];
    const double delhead = (double)bt->read_seal[bt->seal_sea[i]] - h[c];
    seaf[c] = seaf[c] + bt->seal_hydf[i] * s->rel_perm[c] * bt->seal_abyd[i] * delhead;
  }
}

/* calc_func, Calc_function.f90:24-101; func_stochn :176-210; func_connflow :212-254;
 * func_rechterm :256-275; the well, river, lake and sea terms when m->bt is set. */
void orc_calc_func(const orc_mesh *m, orc_state *s, const double *h, double *f) {
  const int n = m->n, ns = m->ns;
  double *alp_ss = xcalloc(n, 8), *surf = xcalloc(ns, 8);
  double *rivf = xcalloc(ns, 8), *lakf = xcalloc(ns, 8);
  double *welf = xcalloc(n, 8), *seaf = xcalloc(n, 8);
  orc_srat_rperm(m, DZERO, h, s->srat_new, s->rel_perm, alp_ss);
  func_surfterm(m, s, h, s->surf_head, surf);
  if (m->bt) {
    const orc_bterms *bt = m->bt;
    func_rlterm(m, s, bt->nrive, bt->rive_cell, bt->rive_head, bt->rive_bott, bt->rive_abyd, h,
                rivf);
    func_rlterm(m, s, bt->nlake, bt->lake_cell, bt->lake_head, bt->lake_bott, bt->lake_abyd, h,
                lakf);
    func_cellterms(m, s, h, welf, seaf);
  }
  for (int i = 0; i < n; i++) {
    /* func_stochn with alp_old = alp_new = alp_ss (Calc_function.f90:53) */
    double ch1 = alp_ss[i] * s->srat_old[i] * s->head_old[i];
    double ch2 = alp_ss[i] * s->srat_new[i] * h[i];
    double ch = (ch2 - ch1) + (double)m->poro[i] * (s->srat_new[i] - s->srat_old[i]);
    double stof = -ch * s->delt_inv * m->cell_vol[i];
    /* func_connflow */
    double conn = DZERO;
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++) {
      int j = m->offrow[k];
      double delhead = h[j] - h[i];
      double relat = upwind(-delhead, s->rel_perm[i], s->rel_perm[j]);
      conn = conn + m->hydf[k] * relat * m->abyd[k] * delhead;
    }
    double fvs = DZERO;
    if (i < ns) fvs = s->calc_rech[i] + surf[i] + rivf[i] + lakf[i];
    f[i] = fvs + stof + conn + welf[i] + seaf[i];
  }
  free(alp_ss); free(surf); free(rivf); free(lakf); free(welf); free(seaf);
}

/* set_rivebound / set_lakebound, Make_linearsystem.f90:287-408 (form_switch
 * == 1): -tran*kr, whatever the branch, plus the FD derivative of kr where
 * the exchange depends on the cell head; assigned per entry */
static void set_rlbound(const orc_mesh *m, const orc_state *s, const double *per_relp, int cnt,
                        const int *cell, const double *head, const double *bott,
                        const double *abyd, double *out) {
  const double *h = s->head_new;
  for (int i = 0; i < cnt; i++) {
    const int c = cell[i];
    double delh, over;
    if (h[c] >= bott[i]) {
      delh = head[i] - h[c];
      over = DONE;
    } else if (head[i] > bott[i]) {
      delh = head[i] - bott[i];
      over = DZERO;
    } else {
      delh = DZERO;
      over = DONE;
    }
    const double tran = m->hydf_surf[c] * abyd[i];
    const double deri = -tran * s->rel_perm[c];
    const double deri_ks = (per_relp[c] - s->rel_perm[c]) * m->newper_inv * delh * tran * over;
    out[c] = deri + deri_ks;
  }
}

/* make_matvec / make_matrix / form_stochn / form_connflow / set_surfbound,
 * Make_linearsystem.f90:24-285,410-497 (form_switch == 1); set_rivebound,
 * set_lakebound, set_seabound (:287-408,499-547) when m->bt is set */
void orc_make_matvec(const orc_mesh *m, orc_state *s) {
  const int n = m->n, ns = m->ns, nc = m->nconn;
  double *tmp = xcalloc(n, 8);
  orc_calc_func(m, s, s->head_new, tmp);
  for (int i = 0; i < n; i++) s->rhs[i] = -tmp[i];
  free(tmp);

  double *per_srat = xcalloc(n, 8), *per_relp = xcalloc(n, 8), *ss_alp = xcalloc(n, 8);
  orc_srat_rperm(m, m->newper, s->head_new, per_srat, per_relp, ss_alp);
  orc_srat_rperm(m, DZERO, s->head_new, s->srat_new, s->rel_perm, ss_alp);
  const double *h = s->head_new;
  double *rivd = xcalloc(ns, 8), *lakd = xcalloc(ns, 8), *sead = xcalloc(n, 8);
  if (m->bt) {
    const orc_bterms *bt = m->bt;
    set_rlbound(m, s, per_relp, bt->nrive, bt->rive_cell, bt->rive_head, bt->rive_bott,
                bt->rive_abyd, rivd);
    set_rlbound(m, s, per_relp, bt->nlake, bt->lake_cell, bt->lake_head, bt->lake_bott,
                bt->lake_abyd, lakd);
    /* set_seabound, :499-547: accumulated per connection in order */
    for (int i = 0; i < bt->nseal; i++) {
      const int c = bt->seal_cell[i];
      const double delh = (double)bt->read_seal[bt->seal_sea[i]] - h[c];
      const double per = per_relp[c] * bt->seal_hydf[i];
      const double rel = s->rel_perm[c] * bt->seal_hydf[i];
      const double deri = -rel * bt->seal_abyd[i];
      const double deri_ks = (per - rel) * m->newper_inv * delh * bt->seal_abyd[i];
      sead[c] = sead[c] + deri + deri_ks;
    }
  }
  const double np = m->newper, npi = m->newper_inv;
  for (int i = 0; i < n; i++) {
    /* form_stochn */
    double deri_srat = (per_srat[i] - s->srat_new[i]) * npi;
    double deri_stor = ss_alp[i] * deri_srat * h[i];
    double stod = -((double)m->poro[i] * deri_srat + ss_alp[i] * s->srat_new[i] + deri_stor) *
                  s->delt_inv * m->cell_vol[i];
    /* form_connflow: three passes merged per row; arithmetic per connection unchanged */
    double cond = DZERO;
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++) {
      int j = m->offrow[k];
      double delhead = h[j] - h[i];
      double relat = upwind(-delhead, s->rel_perm[i], s->rel_perm[j]);
      double rel_hyd = m->hydf[k] * relat;
      double deri_dcon = -rel_hyd * m->abyd[k];
      double deri_lucon = rel_hyd * m->abyd[k];
      double ph1 = -delhead + np, ph2 = -delhead - np;
      double dh1 = m->hydf[k] * upwind(ph1, per_relp[i], s->rel_perm[j]);
      double dh2 = m->hydf[k] * upwind(ph2, s->rel_perm[i], per_relp[j]);
      double con1 = (dh1 - rel_hyd) * npi * delhead * m->abyd[k];
      double con2 = (dh2 - rel_hyd) * npi * delhead * m->abyd[k];
      s->lumat[k] = deri_lucon + con2;
      cond = cond + deri_dcon + con1;
    }
    double dmats = DZERO;
    if (i < ns) {
      /* set_surfbound (:410-497); rivd = lakd = 0 */
      double delh, over;
      const double *sh = s->surf_head;
      if (h[i] >= m->surf_top[i]) {
        delh = (sh[i] > m->surf_top[i]) ? sh[i] - h[i] : m->surf_top[i] - h[i];
        over = DONE;
      } else if (h[i] >= m->surf_bott[i]) {
        delh = (sh[i] > m->surf_bott[i]) ? sh[i] - h[i] : m->surf_bott[i] - h[i];
        over = DONE;
      } else if (sh[i] > m->surf_bott[i]) {
        delh = sh[i] - m->surf_bott[i];
        over = DZERO;
      } else {
        delh = DZERO;
        over = DONE;
      }
      double tran = m->hydf_surf[i] * m->abyd_surf[i];
      double deri_s = -tran * s->rel_perm[i];
      double deri_ks = (per_relp[i] - s->rel_perm[i]) * npi * delh * tran * over;
      double surd = deri_s + deri_ks;
      dmats = surd + rivd[i] + lakd[i];
    }
    s->dmat[i] = dmats + stod + cond + sead[i];
  }
  (void)nc;
  free(per_srat); free(per_relp); free(ss_alp); free(rivd); free(lakd); free(sead);
}

/* ------------------------------------------------------------------ */
double orc_dot(int n, const double *a, const double *b) {
  double sacc = 0.0;
  for (int i = 0; i < n; i++) sacc += a[i] * b[i];
  return sacc;
}

/* calc_matvec, Linear_solution.f90:917-951 */
void orc_matvec(const orc_mesh *m, const double *dmat, const double *lumat, const double *x,
                double *y) {
  for (int i = 0; i < m->n; i++) {
    double t = dmat[i] * x[i];
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++) t = t + lumat[k] * x[m->offrow[k]];
    y[i] = t;
  }
}

/* calc_resi, Calc_simulation.f90:36-73 */
static void orc_resi(const orc_mesh *m, const double *dmat, const double *lumat, const double *x,
                     const double *b, double *r) {
  for (int i = 0; i < m->n; i++) {
    double t = dmat[i] * x[i];
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++) t = t + lumat[k] * x[m->offrow[k]];
    r[i] = b[i] - t;
  }
}

/* precon_dilu, Linear_solution.f90:742-782 */
void orc_precon_dilu(const orc_mesh *m, const double *dmat, const double *lumat, double *d) {
  for (int i = 0; i < m->n; i++) d[i] = dmat[i];
  for (int i = 0; i < m->n; i++) {
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++) {
      int r = m->offrow[k];
      if (3 >= m->dir[k]) {
        double d_invk = DONE / d[r];
        for (int j = m->offind[r]; j < m->offind[r + 1]; j++) {
          if (3 < m->dir[j] && m->offrow[j] == i) d[i] = d[i] - lumat[k] * d_invk * lumat[j];
        }
      }
    }
  }
}

/* solve_ilu, Linear_solution.f90:803-855 */
void orc_solve_ilu(const orc_mesh *m, const double *r, const double *d, const double *lumat,
                   double *z) {
  const int n = m->n;
  for (int i = 0; i < n; i++) z[i] = r[i];
  for (int i = 0; i < n; i++) {
    double t = z[i];
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++)
      if (3 >= m->dir[k]) t = t - lumat[k] * z[m->offrow[k]];
    z[i] = t / d[i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double t = DZERO;
    for (int k = m->offind[i]; k < m->offind[i + 1]; k++)
      if (3 < m->dir[k]) t = t + lumat[k] * z[m->offrow[k]];
    z[i] = z[i] - t / d[i];
  }
}

/* solve_bicgs, Linear_solution.f90:234-493 (precon_type = 0); check_insol,
 * Check_simulation.f90:16-35 */
int orc_bicgstab(const orc_mesh *m, const double *dmat, const double *lumat, const double *rhs,
                 double bnorm, double errtol, int maxinn, double *x, double *rnorm_out,
                 int *conv) {
  return orc_bicgstab_ex(m, dmat, lumat, rhs, bnorm, errtol, maxinn, x, rnorm_out, conv, NULL,
                         NULL);
}

int orc_bicgstab_ex(const orc_mesh *m, const double *dmat, const double *lumat,
                    const double *rhs, double bnorm, double errtol, int maxinn, double *x,
                    double *rnorm_out, int *conv, double *d_out, double *last_in) {
  const int n = m->n;
  double *d = xcalloc(n, 8), *z = xcalloc(n, 8), *p = xcalloc(n, 8), *v = xcalloc(n, 8);
  double *rs = xcalloc(n, 8), *t = xcalloc(n, 8), *resi = xcalloc(n, 8);
  double sk = 0, sk0 = DINFI, sk2 = 0, alpha = 0, beta = 0, omega = DONE, ts = 0, tt = 0;
  double rnorm = 0;
  int conv_flag = 0, in_iter;
  orc_resi(m, dmat, lumat, x, rhs, resi);
  orc_precon_dilu(m, dmat, lumat, d);
  memcpy(rs, resi, (size_t)n * 8);
  for (in_iter = 1; in_iter <= maxinn; in_iter++) {
    sk = orc_dot(n, resi, rs);
    if (in_iter == 1) {
      memcpy(p, resi, (size_t)n * 8);
    } else {
      beta = (sk / sk0) * (alpha / omega);
      for (int i = 0; i < n; i++) p[i] = resi[i] + beta * (p[i] - omega * v[i]);
    }
    orc_solve_ilu(m, p, d, lumat, z);
    if (last_in) memcpy(last_in, p, (size_t)n * 8);
    orc_matvec(m, dmat, lumat, z, v);
    sk2 = orc_dot(n, v, rs);
    alpha = sk / sk2;
    for (int i = 0; i < n; i++) {
      x[i] = x[i] + alpha * z[i];
      resi[i] = resi[i] - alpha * v[i];
    }
    rnorm = orc_dot(n, resi, resi);
    conv_flag = (sqrt(rnorm / bnorm) <= errtol) ? 1 : 0;
    if (conv_flag == 1) break;
    orc_solve_ilu(m, resi, d, lumat, z);
    if (last_in) memcpy(last_in, resi, (size_t)n * 8);
    orc_matvec(m, dmat, lumat, z, t);
    ts = orc_dot(n, t, resi);
    tt = orc_dot(n, t, t);
    omega = ts / tt;
    for (int i = 0; i < n; i++) {
      x[i] = x[i] + omega * z[i];
      resi[i] = resi[i] - omega * t[i];
    }
    rnorm = orc_dot(n, resi, resi);
    conv_flag = (sqrt(rnorm / bnorm) <= errtol) ? 1 : 0;
    if (conv_flag == 1 || in_iter == maxinn) break;
    sk0 = sk;
  }
  if (in_iter > maxinn) in_iter = maxinn;
  if (d_out) memcpy(d_out, d, (size_t)n * 8);
  free(d); free(z); free(p); free(v); free(rs); free(t); free(resi);
  *rnorm_out = rnorm;
  *conv = conv_flag;
  return in_iter;
}

/* check_abserrmax, Check_simulation.f90:37-91 (last index wins ties) */
static void check_abserrmax(int n, const double *xn, const double *xp, double *chmax,
                            double *xmax, int *nunmax) {
  int nan_flag = 0, nun = 0;
  double aval = 0, bval = 0, bch = 0, ach = 0, change = 0, absolute = 0;
  for (int i = 0; i < n; i++) {
    if (xn[i] != xn[i]) nan_flag = 1;
    double xa = fabs(xn[i]);
    double xd = xn[i] - xp[i];
    double xda = fabs(xd);
    if (xda >= ach) { bch = xd; ach = xda; nun = i; }
    if (xa >= aval) { aval = xa; bval = xn[i]; }
  }
  if (nan_flag == 0) {
    if (ach >= change) { *chmax = bch; *nunmax = nun; change = ach; }
    if (aval >= absolute) { *xmax = bval; absolute = aval; }
  } else {
    *chmax = 0.0;
    *xmax = NAN;
    *nunmax = 0;
  }
}

/* calc_vecjacf, Calc_function.f90:515-621 (Brown-Saad step) */
static void calc_vecjacf(const orc_mesh *m, orc_state *s, const double *x, const double *v,
                         double *out) {
  const int n = m->n;
  double *f1 = xcalloc(n, 8), *f2 = xcalloc(n, 8), *jc = xcalloc(n, 8);
  orc_calc_func(m, s, x, f1);
  double l2_v = orc_dot(n, v, v);
  double l2_x = orc_dot(n, x, v);
  double l1_v = 0.0;
  for (int i = 0; i < n; i++) l1_v = l1_v + fabs(v[i]);
  double sign = (l2_x < DZERO) ? -DONE : DONE;
  double eps = sign * sqrt(MACHI_EPS) * fmax(fabs(l2_x), l1_v) / l2_v;
  double eps_inv = DONE / eps;
  for (int i = 0; i < n; i++) jc[i] = x[i] + eps * v[i];
  orc_calc_func(m, s, jc, f2);
  for (int i = 0; i < n; i++) out[i] = -(f2[i] - f1[i]) * eps_inv;
  free(f1); free(f2); free(jc);
}

void orc_calc_vecjacf(const orc_mesh *m, orc_state *s, const double *x, const double *v,
                      double *out) {
  calc_vecjacf(m, s, x, v, out);
}

typedef struct {
  double l2_new, l2_pre, l2_jnorm, lam, eta, gradient, maxstep;
  int back_iter, back_flag, beta_iter;
  double *new_f;
} newton_t;

/* calc_funcl2norm, Nonlinear_solution.f90:760-790 */
static void funcl2norm(const orc_mesh *m, orc_state *s, newton_t *nt, double lam) {
  for (int i = 0; i < m->n; i++) s->head_new[i] = s->head_pre[i] + lam * s->head_change[i];
  orc_calc_func(m, s, s->head_new, nt->new_f);
  nt->l2_new = orc_dot(m->n, nt->new_f, nt->new_f);
}

/* set_eater, Nonlinear_solution.f90:414-439 (Eisenstat-Walker choice 1) */
static void set_eater(newton_t *nt) {
  const double eta_max = 0.9, eta_min = 1.0e-4;
  const double eta_alpha = (DONE + sqrt(5.0)) / DTWO;
  double eta_safe = pow(nt->eta, eta_alpha);
  double minus1 = DONE - nt->lam;
  double lin = sqrt(minus1 * minus1 * nt->l2_pre + DTWO * nt->lam * minus1 * nt->gradient +
                    nt->lam * nt->lam * nt->l2_jnorm);
  nt->eta = fabs(sqrt(nt->l2_new) - lin) / sqrt(nt->l2_pre);
  if (eta_safe < 0.1) eta_safe = DZERO;
  nt->eta = fmax(nt->eta, eta_safe);
  nt->eta = fmax(nt->eta, eta_min);
  nt->eta = fmin(nt->eta, eta_max);
}

/* run_backtr, Nonlinear_solution.f90:441-631 */
static void run_backtr(const orc_mesh *m, orc_state *s, newton_t *nt) {
  const int n = m->n;
  const double back_alpha = 1.0e-4, back_beta = 0.9;
  double l2_new2 = nt->l2_new, lam2 = DONE, maxpnorm = DONE, lam_length = DONE;
  double temp_lam, lam_max, lam_base, lam_diff, lam_incr, alpha_cond, beta_cond;
  nt->lam = DONE;
  double *jacvec = xcalloc(n, 8);
  orc_calc_func(m, s, s->head_new, nt->new_f);
  double l2_pnorm = orc_dot(n, s->head_change, s->head_change);
  calc_vecjacf(m, s, s->head_pre, s->head_change, jacvec);
  double sql2_pnorm = sqrt(l2_pnorm);
  if (sql2_pnorm > nt->maxstep) {
    maxpnorm = nt->maxstep / sql2_pnorm;
    for (int i = 0; i < n; i++) s->head_change[i] = s->head_change[i] * maxpnorm;
    sql2_pnorm = nt->maxstep;
  }
  double step_len = sql2_pnorm;
  nt->l2_new = orc_dot(n, nt->new_f, nt->new_f);
  double slope = orc_dot(n, s->rhs, jacvec) * maxpnorm;
  for (int i = 0; i < n; i++) {
    if (s->head_new[i] != DZERO) {
      temp_lam = fabs(s->head_change[i]) / fabs(s->head_new[i]);
      if (temp_lam > lam_length) lam_length = temp_lam;
    }
  }
  nt->l2_jnorm = nt->l2_jnorm * maxpnorm * maxpnorm;
  free(jacvec);
  double step_tol = pow(MACHI_EPS, 2.0 / 3.0);
  double lam_min = step_tol / lam_length;
  alpha_cond = DHALF * nt->l2_pre + back_alpha * nt->lam * slope;
  nt->gradient = slope;
  for (;;) {
    if (DHALF * nt->l2_new <= alpha_cond) break;
    nt->back_iter++;
    funcl2norm(m, s, nt, nt->lam);
    double lam = nt->lam;
    if (lam == DONE) {
      temp_lam = -slope / (DTWO * (DHALF * nt->l2_new - DHALF * nt->l2_pre - slope));
    } else {
      double rhs1 = DHALF * nt->l2_new - DHALF * nt->l2_pre - lam * slope;
      double rhs2 = l2_new2 - DHALF * nt->l2_pre - lam2 * slope;
      double lam_inv = DONE / (lam * lam), lam2_inv = DONE / (lam2 * lam2);
      double del_lam = DONE / (lam - lam2);
      double av = (rhs1 * lam_inv - rhs2 * lam2_inv) * del_lam;
      double bv = (-lam2 * rhs1 * lam_inv + lam * rhs2 * lam2_inv) * del_lam;
      if (av == 0) {
        temp_lam = -slope / (DTWO * bv);
      } else {
        double root = bv * bv - 3.0 * av * slope;
        if (root < DZERO) temp_lam = DHALF * lam;
        else if (bv <= DZERO) temp_lam = (-bv + sqrt(root)) / (3.0 * av);
        else temp_lam = -slope / (bv + sqrt(root));
        if (temp_lam > DHALF * lam) temp_lam = DHALF * lam;
      }
    }
    lam2 = lam;
    l2_new2 = DHALF * nt->l2_new;
    nt->lam = fmax(temp_lam, 0.1 * lam);
    if (nt->lam < lam_min) { nt->back_flag = 1; return; }
    alpha_cond = DHALF * nt->l2_pre + back_alpha * nt->lam * slope;
  }
  alpha_cond = DHALF * nt->l2_pre + back_alpha * nt->lam * slope;
  beta_cond = DHALF * nt->l2_pre + back_beta * nt->lam * slope;
  if (DHALF * nt->l2_new < beta_cond) {
    if (nt->lam == DONE && sql2_pnorm < step_len) {
      lam_max = step_len / sql2_pnorm;
      for (;;) {
        if (DHALF * nt->l2_new > alpha_cond || DHALF * nt->l2_new >= beta_cond || nt->lam >= lam_max)
          break;
        lam2 = nt->lam; l2_new2 = DHALF * nt->l2_new; nt->lam = fmin(DTWO * nt->lam, lam_max);
        funcl2norm(m, s, nt, nt->lam);
        alpha_cond = DHALF * nt->l2_pre + back_alpha * nt->lam * slope;
        beta_cond = DHALF * nt->l2_pre + back_beta * nt->lam * slope;
      }
    }
    if (nt->lam < DONE || (nt->lam > DONE && DHALF * nt->l2_new > alpha_cond)) {
      lam_base = fmin(nt->lam, lam2); lam_diff = fabs(lam2 - nt->lam);
      for (;;) {
        if (DHALF * nt->l2_new <= alpha_cond && (DHALF * nt->l2_new >= beta_cond || lam_diff <= lam_min))
          break;
        lam_incr = DHALF * lam_diff; nt->lam = lam_base + lam_incr;
        funcl2norm(m, s, nt, nt->lam);
        alpha_cond = DHALF * nt->l2_pre + back_alpha * nt->lam * slope;
        beta_cond = DHALF * nt->l2_pre + back_beta * nt->lam * slope;
        if (DHALF * nt->l2_new > alpha_cond) lam_diff = lam_incr;
        else if (DHALF * nt->l2_new < beta_cond) { lam_base = nt->lam; lam_diff = lam_diff - lam_incr; }
        if (lam_diff == DZERO) break;
      }
      if (DHALF * nt->l2_new < beta_cond) {
        funcl2norm(m, s, nt, lam_base);
        nt->beta_iter++;
      }
      if (nt->beta_iter == 10) { nt->back_flag = 1; return; }
    }
  }
}

/* calc_numsol, Nonlinear_solution.f90:33-294 (sim_type = 1, precon_type = 0) */
void orc_calc_numsol(const orc_mesh *m, orc_state *s, const orc_settings *cfg) {
  const int n = m->n;
  newton_t nt;
  memset(&nt, 0, sizeof(nt));
  nt.new_f = xcalloc(n, 8);
  /* set_backtr (:386-412) */
  double l2x = orc_dot(n, s->head_new, s->head_new);
  nt.maxstep = sqrt(l2x) * VARMAX;
  if (nt.maxstep < DONE) nt.maxstep = DONE;
  s->in_iter_total = 0; s->back_total = 0;
  double L = m->len_scal;
  for (s->out_iter = 1; s->out_iter <= cfg->maxout_iter; s->out_iter++) {
    if (s->out_iter != 1) nt.l2_pre = nt.l2_new;
    nt.back_flag = 0; nt.back_iter = 0; nt.beta_iter = 0;
    memset(s->lumat, 0, (size_t)m->nconn * 8);
    memset(s->dmat, 0, (size_t)n * 8);
    memset(s->rhs, 0, (size_t)n * 8);
    orc_make_matvec(m, s);
    if (s->out_iter == 1) {
      nt.l2_new = orc_dot(n, s->rhs, s->rhs);
      nt.l2_pre = nt.l2_new;
      nt.eta = DHALF;
    }
    memcpy(s->head_pre, s->head_new, (size_t)n * 8);
    memset(s->head_change, 0, (size_t)n * 8);
    int conv = 0;
    double last;
    int in_iter;
    if (cfg->precon_type == 1) {
      /* loop_amg leaves the D-ILU diagonal in array_var(1)%dmat and its last
       * input in %rhs (Linear_solution.f90:517-529,718); run_backtr's slope
       * then reads that rhs (Nonlinear_solution.f90:500) */
      double *d_amg = xcalloc(n, 8), *r_amg = xcalloc(n, 8);
      in_iter = orc_bicgstab_ex(m, s->dmat, s->lumat, s->rhs, nt.l2_pre, nt.eta,
                                cfg->maxinn_iter, s->head_change, &last, &conv, d_amg, r_amg);
      memcpy(s->dmat, d_amg, (size_t)n * 8);
      memcpy(s->rhs, r_amg, (size_t)n * 8);
      free(d_amg);
      free(r_amg);
    } else {
      in_iter = orc_bicgstab(m, s->dmat, s->lumat, s->rhs, nt.l2_pre, nt.eta, cfg->maxinn_iter,
                             s->head_change, &last, &conv);
    }
    nt.l2_jnorm = last;
    s->last_in_iter = in_iter;
    s->in_iter_total += in_iter;
    s->conv_flag = conv;
    for (int i = 0; i < n; i++) s->head_new[i] = s->head_pre[i] + s->head_change[i];
    double maxvar = 0, maxunk = 0;
    int maxnun = 0;
    check_abserrmax(n, s->head_new, s->head_pre, &maxvar, &maxunk, &maxnun);
    double check_val = fabs(maxvar) * L;
    maxunk = fabs(maxunk) * L;
    if (s->conv_flag != 0 && maxunk == maxunk) {
      s->conv_flag = 0;
      if (check_val <= cfg->criteria && maxunk < XMAX) s->conv_flag = 1;
    } else if (maxunk != maxunk) {
      s->conv_flag = 0; /* "Nan detected." */
      break;
    }
    if (s->conv_flag != 1) {
      run_backtr(m, s, &nt);
      check_abserrmax(n, s->head_new, s->head_pre, &maxvar, &maxunk, &maxnun);
      check_val = fabs(maxvar) * L;
      maxunk = fabs(maxunk) * L;
      if (check_val >= VARMAX || maxunk >= XMAX) nt.back_flag = 1;
      else if (check_val <= cfg->criteria && maxunk < XMAX) s->conv_flag = 1;
      if (s->conv_flag != 1 && nt.back_flag != 1) set_eater(&nt);
    }
    s->back_total += nt.back_iter;
    s->last_back_iter = nt.back_iter;
    orc_calc_func(m, s, s->head_new, nt.new_f);
    nt.l2_new = orc_dot(n, nt.new_f, nt.new_f);
    double conv_var = maxvar * L;
    if (s->conv_flag == 1) {
      /* calc_surfw (:326-384): surf_head = head_new on the top layer */
      for (int i = 0; i < m->ns; i++) s->surf_head[i] = s->head_new[i];
      break;
    } else if (nt.back_flag == 1) {
      break;
    } else if (s->out_iter == cfg->maxout_iter) {
      break;
    } else if (fabs(conv_var) >= VARMAX || maxunk >= XMAX) {
      break;
    }
  }
  if (s->out_iter > cfg->maxout_iter) s->out_iter = cfg->maxout_iter;
  free(nt.new_f);
}

/* ------------------------------------------------------------------ */
/* update_tstep / calc_nextst / set_nextet / set_delt / set_nextvar /
 * apply_heuri, Time_module.f90 */
static void calc_nextst(const orc_mesh *m, orc_state *s, const orc_settings *cfg) {
  const int n = m->n;
  if (s->current_t == 0.0f) {
    s->timestep_num = 0;
    s->delt_old1 = DZERO;
    memcpy(s->head_old2, m->h0, (size_t)n * 8);
    s->delt = (double)cfg->init_step;
    double *cs = xcalloc(n, 8);
    orc_srat_rperm(m, DZERO, m->h0, cs, s->rel_perm, NULL);
    for (int i = 0; i < n; i++) {
      s->head_old[i] = m->h0[i]; s->head_new[i] = m->h0[i];
      s->srat_old[i] = cs[i]; s->srat_new[i] = cs[i];
    }
    for (int i = 0; i < m->ns; i++) s->surf_head[i] = m->h0[i];
    for (int i = 0; i < m->ns; i++) s->surf_old[i] = s->surf_head[i];
    free(cs);
  } else if (s->conv_flag == 1) {
    if (s->timestep_num == 0) s->delt_old1 = s->delt;
    s->timestep_num++;
    s->delt_old2 = s->delt_old1;
    s->delt_old1 = s->delt;
    /* write_bound_change: boundstep >= 1 keeps delt (:219-234); else
     * apply_heuri (:1522-1543) */
    int incr_num = (int)((float)cfg->maxout_iter * 0.4f);
    int decr_num = (int)((float)cfg->maxout_iter * 0.8f);
    if (s->rflag >= 1) s->delt = s->delt_old1;
    else if (s->out_iter <= incr_num) s->delt = s->delt_old1 * 1.2;
    else if (s->out_iter <= decr_num) s->delt = s->delt_old1;
    else s->delt = s->delt_old1 * 0.8;
    memcpy(s->head_old2, s->head_old, (size_t)n * 8);
    memcpy(s->head_old, s->head_new, (size_t)n * 8);
    memcpy(s->srat_old, s->srat_new, (size_t)n * 8);
    memcpy(s->surf_old, s->surf_head, (size_t)m->ns * 8);
  } else {
    s->current_t = s->current_t - (float)s->delt;
    s->delt = s->delt * DHALF;
    s->rflag = 0; /* reset_stepf (:1430-1500) */
    memcpy(s->head_new, s->head_old, (size_t)n * 8);
    memcpy(s->srat_new, s->srat_old, (size_t)n * 8);
    memcpy(s->surf_head, s->surf_old, (size_t)m->ns * 8);
  }
  s->next_time = s->current_t + (float)s->delt;
}

void orc_update_tstep(const orc_mesh *m, orc_state *s, const orc_settings *cfg) {
  calc_nextst(m, s, cfg);
  if (s->conv_flag == 1) {
    /* set_nextet (:274-330): past a record's end, read the next end time;
     * at the end of the file, flag 0 and end_time (read_next,
     * Read_module.f90:984-1065) */
    if (s->rflag == 1) {
      if (s->rk + 1 < cfg->n_rech) {
        s->etime = cfg->rech_etime[s->rk + 1];
      } else {
        /* and reset_value: read_rech = 0 (:341-350, 1502-1520) */
        s->rflag = 0;
        s->etime = cfg->end_time;
        s->rk = cfg->n_rech;
      }
    }
    /* set_delt (:770-800): recharge is the only boundary with end times */
    float min_step = s->etime;
    if (min_step == s->current_t) {
      /* delt unchanged */
    } else if (min_step < s->next_time && s->next_time < cfg->end_time) {
      s->delt = (double)(min_step - s->current_t);
    } else if (s->next_time > cfg->end_time) {
      s->delt = (double)(cfg->end_time - s->current_t);
    }
    if (s->delt > (double)cfg->max_step) s->delt = (double)cfg->max_step;
    s->next_time = s->current_t + (float)s->delt;
    /* set_nextvar (:802-858): read the next record's values, or flag the
     * step that ends at the record's end */
    if (s->rflag == 1) {
      s->rk++;
      s->rflag = 0;
    } else if (s->etime == s->next_time) {
      s->rflag = 1;
    }
  } else if (s->next_time > cfg->end_time) {
    s->delt = (double)(cfg->end_time - s->current_t);
  }
  s->current_t = s->current_t + (float)s->delt;
  s->delt_inv = DONE / s->delt;
  if (s->conv_flag == 1 || s->timestep_num == 0) {
    if (cfg->n_rech > 0) {
      /* assign_surfbv's conversion: mm/h * len_scal_inv * uni_conv, float32
       * (Assign_boundary.f90:216, Open_file.f90:744) */
      const float mmh = s->rk < cfg->n_rech ? cfg->rech_mmh[s->rk] : 0.0f;
      const float rate = (mmh * (float)m->len_scal_inv) * (1.0e-3f / 3600.0f);
      for (int i = 0; i < m->ns; i++) s->calc_rech[i] = (double)rate * m->rech_area[i];
    } else {
      for (int i = 0; i < m->ns; i++)
        s->calc_rech[i] = (double)m->read_rech[i] * m->rech_area[i];
    }
  }
}

int orc_run(const orc_mesh *m, const orc_settings *cfg, int nstep_max, double *heads_out,
            double *times_out, double *dts_out, int *newton_out, int *inner_out) {
  orc_state s;
  orc_state_init(&s, m);
  /* the first record's end time, clipped to end_time (Open_file.f90:816-818) */
  s.etime = (cfg->n_rech > 0 && cfg->rech_etime[0] < cfg->end_time) ? cfg->rech_etime[0]
                                                                     : cfg->end_time;
  int nstep = 0, attempts = 0;
  for (;;) {
    orc_update_tstep(m, &s, cfg);
    if (s.delt < DSMAL) { nstep = -1; break; }
    orc_calc_numsol(m, &s, cfg);
    if (++attempts > 100000) { nstep = -1; break; }
    if (s.conv_flag == 1) {
      if (nstep < nstep_max) {
        if (heads_out) memcpy(heads_out + (size_t)nstep * m->n, s.head_new, (size_t)m->n * 8);
        if (times_out) times_out[nstep] = (double)s.current_t;
        if (dts_out) dts_out[nstep] = s.delt;
        if (newton_out) newton_out[nstep] = s.out_iter;
        if (inner_out) inner_out[nstep] = s.in_iter_total;
      }
      nstep++;
      if (s.current_t >= cfg->end_time) break; /* check_lastts */
    }
  }
  orc_state_free(&s);
  return nstep;
}
