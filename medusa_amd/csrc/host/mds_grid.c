/*
 * mds_grid.c -- grid and geometry of one problem (host side of the product).
 *
 * Follows the reference's geometry initialisation (geometry.c:4-62):
 *   Cartesian  metrics/metric_cartesian.c:47-105 (moments), :134-187 (volume, area),
 *              unit metric and zero connection (:189-295)
 *   spherical  metrics/metric_spherical_general.c:558-603 (coords, moments),
 *              :708-820 (volume, area), :837-929 (connection), :941-1051
 *              (gcov, gcon, scale), 1D columns (x1 in [-1,1], dx2 = 2 pi),
 *              POLYTH theta map with the io.c:5014-5015 defaults,
 *              radial map r=|x| (:149-152), r=exp(x) (metric_spherical_logr.c)
 *              or r=rtrans sinh(|x|/rtrans) (:155-161); 3D grids uniform or
 *              dendritic (utils.c:181-590)
 *   Romberg    utils.c:14-62
 * and adds the FP32-bulk reconstruction coefficients (MDS_RC_*), evaluated in
 * long double from the same rounded FP64 volume moments.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/mds_hydro.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif
#define NG MDS_NG

/* ---- Romberg, utils.c:14-62 ---- */
typedef double (*integrand)(double, const void *);

static double trap_stage(double x0, double x1, int stage, integrand f, const void *ctx)
{
  unsigned long long nseg = 1 << stage;
  double h = (x1 - x0) / nseg;
  if (stage == 0) return (0.5 * h) * (f(x0, ctx) + f(x1, ctx));
  double acc = 0.0;
  for (unsigned long long q = 0; q < (unsigned long long)(1 << (stage - 1)); q++)
    acc += f(x0 + (2 * q + 1) * h, ctx);
  return h * acc;
}

static double romberg(double x0, double x1, integrand f, const void *ctx)
{
  double row[16];
  row[0] = trap_stage(x0, x1, 0, f, ctx);
  for (int st = 1; st <= 15; st++) {
    double prev = row[0];
    row[0] = (0.5 * prev) + trap_stage(x0, x1, st, f, ctx);
    for (int m = 1; m <= st; m++) {
      unsigned long long w = 1 << (2 * m);
      double keep = row[m];
      row[m] = (1.0 / (w - 1)) * ((w * row[m - 1]) - prev);
      prev = keep;
    }
    if (fabs((row[st] - row[st - 1]) / row[st]) < 1.e-15 || (st > 2 && fabs(row[st]) < 1.e-15))
      return row[st];
  }
  return row[15];
}

/* ---- spherical maps ---- */
typedef struct { int rcoord; double pnorm, pxt, palpha, rtrans; } maps;

static double rx(const maps *m, double x)
{
  if (m->rcoord == MDS_RCOORD_LOG) return exp(x);
  if (m->rcoord == MDS_RCOORD_SINH) return m->rtrans * sinh(fabs(x) / m->rtrans);
  return fabs(x);
}
static double drx(const maps *m, double x)
{
  if (m->rcoord == MDS_RCOORD_LOG) return exp(x);
  if (m->rcoord == MDS_RCOORD_SINH) return cosh(fabs(x) / m->rtrans);
  return 1.0;
}
static double d2rx(const maps *m, double x)
{
  if (m->rcoord == MDS_RCOORD_LOG) return exp(x);
  if (m->rcoord == MDS_RCOORD_SINH) return sinh(fabs(x) / m->rtrans) / m->rtrans;
  return 0.0;
}
static void maps_init(maps *m, const mds_grid *g)
{
  m->rcoord = g->rcoord;
  m->rtrans = g->rtrans;
  m->pxt = 1.2;
  m->palpha = 4;
  m->pnorm = ((0.5 * M_PI) * 1.0) / (1.0 + ((1.0 / (m->palpha + 1.0)) * (1.0 / pow(m->pxt, m->palpha))));
}
static double fold(double x) { if (x < -1.0) x = -2.0 - x; if (x > 1.0) x = 2.0 - x; return x; }
static double thx(const maps *m, double x)
{
  x = fold(x);
  return ((m->pnorm * x) * (1.0 + (pow(x / m->pxt, m->palpha) / (m->palpha + 1.0)))) + (0.5 * M_PI);
}
static double dthx(const maps *m, double x) { x = fold(x); return m->pnorm * (1.0 + pow(x / m->pxt, m->palpha)); }
static double d2thx(const maps *m, double x)
{
  x = fold(x);
  return ((m->pnorm * m->palpha) * pow(x / m->pxt, m->palpha - 1.0)) / m->pxt;
}
static double i_r2dr(double x, const void *c) { const maps *m = c; return pow(rx(m, x), 2) * drx(m, x); }
static double i_xr1(double x, const void *c) { const maps *m = c; return (x * pow(rx(m, x), 2)) * drx(m, x); }
static double i_xr2(double x, const void *c) { const maps *m = c; return ((x * x) * pow(rx(m, x), 2)) * drx(m, x); }
static double i_r2d2r(double x, const void *c) { const maps *m = c; return pow(rx(m, x), 2) * d2rx(m, x); }
static double i_r3(double x, const void *c) { const maps *m = c; return pow(rx(m, x), 3); }
static double i_r2dr3(double x, const void *c) { const maps *m = c; return pow(rx(m, x), 2) * pow(drx(m, x), 3); }
static double i_rdr2(double x, const void *c) { const maps *m = c; return rx(m, x) * pow(drx(m, x), 2); }
static double i_sdth(double x, const void *c) { const maps *m = c; return sin(thx(m, x)) * dthx(m, x); }
static double i_sdth3(double x, const void *c) { const maps *m = c; return sin(thx(m, x)) * pow(dthx(m, x), 3); }
static double i_s3dth(double x, const void *c) { const maps *m = c; return pow(sin(thx(m, x)), 3) * dthx(m, x); }
static double i_sd2th(double x, const void *c) { const maps *m = c; return sin(thx(m, x)) * d2thx(m, x); }
static double i_cdth2(double x, const void *c) { const maps *m = c; return cos(thx(m, x)) * pow(dthx(m, x), 2); }

/* FP32-bulk reconstruction coefficients of zone q (pencil index with NG
 * offset) from the rounded moments, in long double. */
static void recon_coeffs(const double *al, const double *be, const double *ga,
                         double startx, double dx, int nz, double *rc)
{
  for (int q = 0; q < nz; q++) {
    double *o = &rc[q * MDS_RC_COUNT];
    for (int k = 0; k < MDS_RC_COUNT; k++) o[k] = 0.0;
    long double xlo = (long double)(startx + (q - NG) * dx);
    long double xhi = (long double)(startx + (q + 1 - NG) * dx);
    long double h = xhi - xlo;
    long double a0 = al[q], b0 = be[q], g0 = ga[q];
    long double m1 = b0 / g0, m2 = a0 / g0;            /* <x>, <x^2> of the zone */
    long double mu1 = (m1 - xlo) / h;
    long double mu2 = ((m2 - 2.0L * xlo * m1) + xlo * xlo) / (h * h);
    long double nur = ((m2 - 2.0L * xhi * m1) + xhi * xhi) / (h * h);
    o[MDS_RC_MU1] = (double)mu1;
    o[MDS_RC_KINV] = (double)(1.0L / (mu2 - mu1));
    o[MDS_RC_INUR] = (double)(1.0L / nur);
    o[MDS_RC_IMU2] = (double)(1.0L / mu2);
    if (q == 0 || q == nz - 1) continue;   /* no neighbours: flat in the first fit */
    long double am = al[q - 1], ap = al[q + 1], bm = be[q - 1], bp = be[q + 1], gm = ga[q - 1], gp = ga[q + 1];
    long double det = am * b0 * gp + a0 * bp * gm + ap * bm * g0 - am * bp * g0 - a0 * bm * gp - ap * b0 * gm;
    long double den = 1.0L / det;
    /* Cramer coefficients of A and B in dqm and dqp (dqc = dqm + dqp), reconstruct.c:157-159 */
    long double Am = (bp * gm * g0 - gp * b0 * gm) * den;
    long double Ap = (gp * bm * g0 - gp * b0 * gm) * den;
    long double Bm = -(ap * gm * g0 - gp * a0 * gm) * den;
    long double Bp = -(gp * am * g0 - gp * a0 * gm) * den;
    long double X2L = xlo * xlo - m2, X1L = xlo - m1;
    long double X2R = xhi * xhi - m2, X1R = xhi - m1;
    o[MDS_RC_CLM] = (double)(Am * X2L + Bm * X1L);
    o[MDS_RC_CLP] = (double)(Ap * X2L + Bp * X1L);
    o[MDS_RC_CRM] = (double)(Am * X2R + Bm * X1R);
    o[MDS_RC_CRP] = (double)(Ap * X2R + Bp * X1R);
  }
}

static int alloc_dir(mds_grid *g, int d)
{
  int nz = g->n[d] + 2 * NG;
  g->alpha[d] = calloc(nz, sizeof(double));
  g->beta[d] = calloc(nz, sizeof(double));
  g->Gamma[d] = calloc(nz, sizeof(double));
  g->rcoef[d] = calloc((size_t)nz * MDS_RC_COUNT, sizeof(double));
  return (g->alpha[d] && g->beta[d] && g->Gamma[d] && g->rcoef[d]) ? 0 : MDS_ENOMEM;
}

static int build_cartesian(mds_grid *g)
{
  for (int d = 0; d < 3; d++) {
    if (alloc_dir(g, d)) return MDS_ENOMEM;
    for (int i = -NG; i < g->n[d] + NG; i++) {
      double x0 = g->startx[d] + i * g->dx[d];
      double x1 = x0 + g->dx[d];
      g->alpha[d][i + NG] = ((x1 * x1 * x1) - (x0 * x0 * x0)) / 3.0;
      g->beta[d][i + NG] = ((x1 * x1) - (x0 * x0)) / 2.0;
      g->Gamma[d][i + NG] = g->dx[d];
    }
    recon_coeffs(g->alpha[d], g->beta[d], g->Gamma[d], g->startx[d], g->dx[d], g->n[d] + 2 * NG, g->rcoef[d]);
  }
  g->volume = 1.0;
  for (int d = 0; d < g->ndim; d++) g->volume *= g->dx[d];
  for (int loc = 0; loc < 3; loc++) {
    g->area[loc] = 0.0;
    if (loc >= g->ndim) continue;
    g->area[loc] = 1.0;
    for (int d = 0; d < g->ndim; d++) g->area[loc] *= g->dx[d];
    g->area[loc] /= g->dx[loc];
  }
  g->zone = NULL;
  return 0;
}

static int build_spherical1d(mds_grid *g)
{
  if (g->ndim != 1) return MDS_ENOTIMPL;
  maps m;
  maps_init(&m, g);
  const double dphi = 2.0 * M_PI, x1lo = -1.0, x1hi = 1.0;
  int n = g->n[0], nz = n + 2 * NG;
  for (int d = 0; d < 3; d++)
    if (alloc_dir(g, d)) return MDS_ENOMEM;
  /* inactive directions: unit moments so that nothing downstream divides by 0 */
  for (int d = 1; d < 3; d++)
    for (int q = 0; q < g->n[d] + 2 * NG; q++) g->Gamma[d][q] = 1.0;
  for (int i = -NG; i < n + NG; i++) {
    double x0 = g->startx[0] + i * g->dx[0];
    double x1 = g->startx[0] + (i + 1) * g->dx[0];
    g->Gamma[0][i + NG] = romberg(x0, x1, i_r2dr, &m);
    g->beta[0][i + NG] = romberg(x0, x1, i_xr1, &m);
    g->alpha[0][i + NG] = romberg(x0, x1, i_xr2, &m);
  }
  recon_coeffs(g->alpha[0], g->beta[0], g->Gamma[0], g->startx[0], g->dx[0], nz, g->rcoef[0]);

  g->zone = calloc((size_t)nz * MDS_ZG_COUNT, sizeof(double));
  if (!g->zone) return MDS_ENOMEM;
  double sth = romberg(x1lo, x1hi, i_sdth, &m);
  double vthf = sth;
  double th_s3 = romberg(x1lo, x1hi, i_sdth3, &m);
  double th_s3c = romberg(x1lo, x1hi, i_s3dth, &m);
  double th_sd2 = romberg(x1lo, x1hi, i_sd2th, &m);
  double th_cd2 = romberg(x1lo, x1hi, i_cdth2, &m);
  for (int i = -NG; i < n + NG; i++) {
    double *z = &g->zone[(size_t)(i + NG) * MDS_ZG_COUNT];
    double xl = g->startx[0] + i * g->dx[0];
    double xh = g->startx[0] + (i + 1) * g->dx[0];
    double r2dr = romberg(xl, xh, i_r2dr, &m);
    z[MDS_ZG_VOLUME] = ((r2dr * sth) * 1) * dphi;
    z[MDS_ZG_AREA0] = (i >= 0 && i <= n) ? (i_r2dr(xl, &m) * sth) * (1 * dphi) : 0.0;
    double vrf = r2dr;
    double r5 = 0.2 * (pow(rx(&m, xh), 5) - pow(rx(&m, xl), 5));
    double gc[3], gf[3];
    gc[0] = romberg(xl, xh, i_r2dr3, &m) / vrf;
    gc[1] = fabs(((r5 / vrf) * th_s3) / vthf);
    gc[2] = fabs(((r5 / vrf) * th_s3c) / vthf);
    gf[0] = pow(drx(&m, xl), 2);
    gf[1] = (pow(rx(&m, xl), 2) * th_s3) / vthf;
    gf[2] = (pow(rx(&m, xl), 2) * th_s3c) / vthf;
    for (int d = 0; d < 3; d++) {
      z[MDS_ZG_GCOV00 + d] = gc[d];
      z[MDS_ZG_GCOVF0 + d] = gf[d];
      z[MDS_ZG_GCON0 + d] = (fabs(gc[d]) > 1.0e-50) ? 1.0 / gc[d] : 0.0;
    }
    z[MDS_ZG_SCALEF] = sqrt(gf[0]);
    z[MDS_ZG_SCALE0] = sqrt(gc[0]);
    if (i >= 0 && i < n) {
      double vr = romberg(xl, xl + g->dx[0], i_r2dr, &m);
      double c101 = romberg(xl, xh, i_rdr2, &m) / vr;
      double c212 = th_cd2 / vthf;
      z[MDS_ZG_C000] = romberg(xl, xh, i_r2d2r, &m) / vr;
      z[MDS_ZG_C011] = -(romberg(xl, xh, i_r3, &m) / vr) * (th_s3 / vthf);
      z[MDS_ZG_C101] = c101;
      z[MDS_ZG_C110] = c101;
      z[MDS_ZG_C202] = c101;
      z[MDS_ZG_C220] = c101;
      z[MDS_ZG_C212] = c212;
      z[MDS_ZG_C221] = c212;
      z[MDS_ZG_C022] = ((-c101) * gc[2]) / gc[0];
      z[MDS_ZG_C122] = ((-c212) * gc[2]) / gc[1];
      z[MDS_ZG_C111] = th_sd2 / vthf;
    }
  }
  g->volume = 0.0;
  g->area[0] = g->area[1] = g->area[2] = 0.0;
  return 0;
}

static double i_xth1(double x, const void *c) { const maps *m = c; return (x * sin(thx(m, x))) * dthx(m, x); }
static double i_xth2(double x, const void *c) { const maps *m = c; return ((x * x) * sin(thx(m, x))) * dthx(m, x); }

static int lev_of(int f)
{
  int l = 0;
  while (f >>= 1) l++;
  return l;
}

/* Dendritic levels (utils.c:181-368, checks :540-579): nj doubles with
 * radius while the zone aspect r dtheta / dr exceeds 2^0.5 times it, from 2
 * at the innermost shell; nk starts at 4 in the polar row of every shell and
 * doubles row by row towards the equator on the same rule, never below the
 * shell inside; both are mirrored into the ghost zones.  A 4:1 jump between
 * shells (the reference exits) or a 1/2:1 jump (the reference's fix-up
 * leaves its ghost rows unchanged) is rejected with MDS_EINVAL. */
static int build_levels(mds_grid *g, const maps *m)
{
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const int S1 = n2 + 2 * NG;
  const double max_log2_aspect = 0.5;
  if (n2 < 2 || (n2 & (n2 - 1))) return MDS_EINVAL;
  if (g->ndim == 3 && (n3 < 4 || (n3 & (n3 - 1)))) return MDS_EINVAL;
  if (g->startx[1] != -1.0 || fabs(g->dx[1] * n2 - 2.0) > 1.0e-14) return MDS_EINVAL;
  int *nj = calloc(n1 + 2 * NG, sizeof(int));
  int *nk = calloc((size_t)(n1 + 2 * NG) * S1, sizeof(int));
  g->dj = calloc(n1 + 2 * NG, sizeof(int));
  g->dk = calloc((size_t)(n1 + 2 * NG) * S1, sizeof(int));
  if (!nj || !nk || !g->dj || !g->dk) {
    free(nj);
    free(nk);
    return MDS_ENOMEM;
  }
#define NJ(i) nj[(i) + NG]
#define NK(i, j) nk[(size_t)((i) + NG) * S1 + (j) + NG]
#define DK(i, j) g->dk[(size_t)((i) + NG) * S1 + (j) + NG]
  const double dth = thx(m, g->startx[1] + n2 * g->dx[1]) - thx(m, g->startx[1]);
  NJ(0) = 2;
  for (int i = 1; i < n1; i++) {
    const double r0 = rx(m, g->startx[0] + (i + 0.0) * g->dx[0]);
    const double rc = rx(m, g->startx[0] + (i + 0.5) * g->dx[0]);
    const double r1 = rx(m, g->startx[0] + (i + 1.0) * g->dx[0]);
    const double aspect = rc * dth / (r1 - r0);
    int v = NJ(i - 1);
    if (log2(aspect / v) > max_log2_aspect) v *= 2;
    NJ(i) = v < n2 ? v : n2;
  }
  for (int i = -NG; i < 0; i++) NJ(i) = NJ(-i - 1);
  for (int i = n1; i < n1 + NG; i++) NJ(i) = NJ(n1 - 1);
  for (int i = -NG; i < n1 + NG; i++) g->dj[i + NG] = n2 / NJ(i);
  const double dph = n3 * g->dx[2];
  /* compute_dk; a 2D grid keeps every row at n3 = 1 zone (utils.c:311-316) */
  for (int i = 0; i < n1 && g->ndim == 2; i++)
    for (int j = 0; j < NJ(i); j++) NK(i, j) = n3;
  for (int i = 0; i < n1 && g->ndim == 3; i++) {
    const double r0 = rx(m, g->startx[0] + (i + 0.0) * g->dx[0]);
    const double rc = rx(m, g->startx[0] + (i + 0.5) * g->dx[0]);
    const double r1 = rx(m, g->startx[0] + (i + 1.0) * g->dx[0]);
    const double dr = r1 - r0;
    NK(i, 0) = 4;
    for (int j = 1; j < NJ(i); j++) {
      if (j >= NJ(i) / 2) {
        NK(i, j) = NK(i, NJ(i) - 1 - j);
        continue;
      }
      const double thc = thx(m, g->startx[1] + (j + 0.5) * g->dj[i + NG] * g->dx[1]);
      const double aspect = rc * sin(thc) * dph / dr;
      int v = NK(i, j - 1);
      if (log2(aspect / v) > max_log2_aspect) v *= 2;
      if (i > 0) {
        const int inner = NJ(i - 1) == NJ(i) ? NK(i - 1, j) : NK(i - 1, j / 2);
        if (inner > v) v = inner;
      }
      NK(i, j) = v < n3 ? v : n3;
    }
  }
  for (int i = -NG; i < 0; i++)
    for (int j = 0; j < NJ(i); j++) NK(i, j) = NK(-i - 1, j);
  for (int i = n1; i < n1 + NG; i++)
    for (int j = 0; j < NJ(i); j++) NK(i, j) = NK(n1 - 1, j);
  for (int i = -NG; i < n1 + NG; i++) {
    for (int j = -NG; j < 0; j++) NK(i, j) = NK(i, -j - 1);
    for (int j = NJ(i); j < NJ(i) + NG; j++) NK(i, j) = NK(i, 2 * NJ(i) - 1 - j);
    for (int j = -NG; j < NJ(i) + NG; j++) DK(i, j) = n3 / NK(i, j);
  }
  int rc = 0;
  for (int i = 0; i < n1 && !rc; i++) {
    if (g->dj[i + NG] != 2 * g->dj[i + 1 + NG]) continue;
    for (int j = 0; j < NJ(i); j++) {
      const int a = DK(i, j) / DK(i + 1, 2 * j), b = DK(i, j) / DK(i + 1, 2 * j + 1);
      if (a == 4 || b == 4 || a == 0 || b == 0) rc = MDS_EINVAL;
    }
  }
  g->maxjlev = lev_of(g->dj[NG]);
  g->maxklev = 0;
  for (int i = -NG; i < n1 + NG; i++)
    for (int j = -NG; j < NJ(i) + NG; j++) {
      const int l = lev_of(DK(i, j));
      if (l > g->maxklev) g->maxklev = l;
    }
#undef NJ
#undef NK
#undef DK
  free(nj);
  free(nk);
  return rc;
}

static int gdj(const mds_grid *g, int i) { return g->dj ? g->dj[i + NG] : 1; }
static int gdk(const mds_grid *g, int i, int j)
{
  return g->dk ? g->dk[(size_t)(i + NG) * (g->n[1] + 2 * NG) + j + NG] : 1;
}

/* 3D spherical grid, uniform or dendritic: metric_spherical_general.c:592-669
 * (moments per level), :708-740 (volume), :745-833 (areas, zero on the
 * axis), :837-929 (connection with the pressure-balanced
 * Gamma^theta_thetatheta; across a k-level change dA is the sum of the two
 * finer faces or half the coarser one), :941-1051 (gcov, gcon, scale).
 * Records are per (i, level-local j): nothing depends on phi, and a row's
 * phi width enters through its dk. */
static int build_spherical3d(mds_grid *g)
{
  maps m;
  maps_init(&m, g);
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const int N1 = n1 + 2 * NG, N2 = n2 + 2 * NG, N3 = n3 + 2 * NG;
  g->maxjlev = g->maxklev = 0;
  if (g->dendritic) {
    int rc = build_levels(g, &m);
    if (rc) return rc;
  }
  for (int d = 0; d < 3; d++)
    if (alloc_dir(g, d)) return MDS_ENOMEM;
  for (int i = -NG; i < n1 + NG; i++) {
    double x0 = g->startx[0] + i * g->dx[0], x1 = g->startx[0] + (i + 1) * g->dx[0];
    g->Gamma[0][i + NG] = romberg(x0, x1, i_r2dr, &m);
    g->beta[0][i + NG] = romberg(x0, x1, i_xr1, &m);
    g->alpha[0][i + NG] = romberg(x0, x1, i_xr2, &m);
  }
  recon_coeffs(g->alpha[0], g->beta[0], g->Gamma[0], g->startx[0], g->dx[0], N1, g->rcoef[0]);
  const int nlev[2] = {g->maxjlev + 1, g->maxklev + 1}, NN[2] = {N2, N3};
  for (int t = 0; t < 2; t++) {
    const size_t sz = (size_t)nlev[t] * NN[t];
    g->lalpha[t] = calloc(sz, sizeof(double));
    g->lbeta[t] = calloc(sz, sizeof(double));
    g->lGamma[t] = calloc(sz, sizeof(double));
    g->lrcoef[t] = calloc(sz * MDS_RC_COUNT, sizeof(double));
    if (!g->lalpha[t] || !g->lbeta[t] || !g->lGamma[t] || !g->lrcoef[t]) return MDS_ENOMEM;
  }
  for (int lev = 0; lev < nlev[0]; lev++) {
    const int sc = 1 << lev;
    double *A = g->lalpha[0] + (size_t)lev * N2, *B = g->lbeta[0] + (size_t)lev * N2, *G = g->lGamma[0] + (size_t)lev * N2;
    for (int j = -NG; j < n2 / sc + NG; j++) {
      double x0 = g->startx[1] + j * sc * g->dx[1], x1 = g->startx[1] + (j + 1) * sc * g->dx[1];
      G[j + NG] = romberg(x0, x1, i_sdth, &m);
      B[j + NG] = romberg(x0, x1, i_xth1, &m);
      A[j + NG] = romberg(x0, x1, i_xth2, &m);
    }
    for (int j = n2 / sc + NG; j < n2 + NG; j++) G[j + NG] = 1.0;
    recon_coeffs(A, B, G, g->startx[1], sc * g->dx[1], n2 / sc + 2 * NG, g->lrcoef[0] + (size_t)lev * N2 * MDS_RC_COUNT);
  }
  for (int lev = 0; lev < nlev[1]; lev++) {
    const int sc = 1 << lev;
    double *A = g->lalpha[1] + (size_t)lev * N3, *B = g->lbeta[1] + (size_t)lev * N3, *G = g->lGamma[1] + (size_t)lev * N3;
    for (int k = -NG; k < n3 / sc + NG; k++) {
      double x0 = g->startx[2] + k * sc * g->dx[2], x1 = g->startx[2] + (k + 1) * sc * g->dx[2];
      G[k + NG] = sc * g->dx[2];
      B[k + NG] = ((x1 * x1) - (x0 * x0)) / 2.0;
      A[k + NG] = ((x1 * x1 * x1) - (x0 * x0 * x0)) / 3.0;
    }
    for (int k = n3 / sc + NG; k < n3 + NG; k++) G[k + NG] = 1.0;
    recon_coeffs(A, B, G, g->startx[2], sc * g->dx[2], n3 / sc + 2 * NG, g->lrcoef[1] + (size_t)lev * N3 * MDS_RC_COUNT);
  }
  for (int t = 0; t < 2; t++) {
    const int d = t + 1;
    memcpy(g->alpha[d], g->lalpha[t], NN[t] * sizeof(double));
    memcpy(g->beta[d], g->lbeta[t], NN[t] * sizeof(double));
    memcpy(g->Gamma[d], g->lGamma[t], NN[t] * sizeof(double));
    memcpy(g->rcoef[d], g->lrcoef[t], (size_t)NN[t] * MDS_RC_COUNT * sizeof(double));
  }

  g->zone = calloc((size_t)N1 * N2 * MDS_ZS_COUNT, sizeof(double));
  if (!g->zone) return MDS_ENOMEM;
  const double dph = g->dx[2];
  const int top_axis = fabs((g->startx[1] + n2 * g->dx[1]) - 1) < 1.0e-15;
  for (int i = -NG; i < n1 + NG; i++) {
    double xl = g->startx[0] + i * g->dx[0], xh = g->startx[0] + (i + 1) * g->dx[0];
    double r2dr = romberg(xl, xh, i_r2dr, &m);
    double vrf = r2dr;
    double r2dr3 = romberg(xl, xh, i_r2dr3, &m);
    double r5 = 0.2 * (pow(rx(&m, xh), 5) - pow(rx(&m, xl), 5));
    const int dj = gdj(g, i), nj = n2 / dj;
    const double *lB = g->lbeta[0] + (size_t)lev_of(dj) * N2, *lG = g->lGamma[0] + (size_t)lev_of(dj) * N2;
    for (int j = -NG; j < nj + NG; j++) {
      double tl = g->startx[1] + j * dj * g->dx[1], th = g->startx[1] + (j + 1) * dj * g->dx[1];
      double sth = romberg(tl, th, i_sdth, &m);
      double vthf = sth;
      double th_s3 = romberg(tl, th, i_sdth3, &m);
      double th_s3c = romberg(tl, th, i_s3dth, &m);
      const int dk = gdk(g, i, j);
      double *z = &g->zone[((size_t)(j + NG) * N1 + (i + NG)) * MDS_ZS_COUNT];
      z[MDS_ZS_VOLUME] = ((r2dr * sth) * dk) * dph;
      if (i >= 0 && i <= n1 && j >= 0 && j <= nj) {
        double dph_int = dk * dph;
        z[MDS_ZS_AREA0] = (i_r2dr(xl, &m) * sth) * dph_int;
        z[MDS_ZS_AREA1] = (r2dr * i_sdth(tl, &m)) * dph_int;
        z[MDS_ZS_AREA2] = r2dr * sth;
        if ((j == 0 && g->startx[1] == -1) || (j == nj && top_axis)) z[MDS_ZS_AREA1] = 0.0;
      }
      double gc[3], gf0[3], gf1[3];
      gc[0] = r2dr3 / vrf;
      gc[1] = fabs(((r5 / vrf) * th_s3) / vthf);
      gc[2] = fabs(((r5 / vrf) * th_s3c) / vthf);
      gf0[0] = pow(drx(&m, xl), 2);
      gf0[1] = (pow(rx(&m, xl), 2) * th_s3) / vthf;
      gf0[2] = (pow(rx(&m, xl), 2) * th_s3c) / vthf;
      gf1[0] = gc[0];
      gf1[1] = fabs((r5 / vrf) * pow(dthx(&m, tl), 2));
      gf1[2] = fabs((r5 / vrf) * pow(sin(thx(&m, tl)), 2));
      for (int d = 0; d < 3; d++) {
        z[MDS_ZS_GCOV0 + d] = gc[d];
        z[MDS_ZS_GCOVF0 + d] = gf0[d];
        z[MDS_ZS_GCOVF1 + d] = gf1[d];
        z[MDS_ZS_GCON + d] = (fabs(gc[d]) > 1.0e-50) ? 1.0 / gc[d] : 0.0;
        z[MDS_ZS_SCALE0 + d] = sqrt(gc[d]);
      }
      z[MDS_ZS_SCALEF0] = sqrt(gf0[0]);
      z[MDS_ZS_SCALEF1] = sqrt(gf1[1]);
      const double thc = thx(&m, lB[j + NG] / lG[j + NG]);
      z[MDS_ZS_STH] = sin(thc);
      z[MDS_ZS_CTH] = cos(thc);
    }
  }
  /* connection coefficients of interior zones (areas above are complete) */
  for (int i = 0; i < n1; i++) {
    double xl = g->startx[0] + i * g->dx[0], xh = g->startx[0] + (i + 1) * g->dx[0];
    double r2d2r = romberg(xl, xh, i_r2d2r, &m);
    double r3i = romberg(xl, xh, i_r3, &m);
    double rdr2 = romberg(xl, xh, i_rdr2, &m);
    double vr = romberg(xl, xl + g->dx[0], i_r2dr, &m);
    const int dj = gdj(g, i);
    for (int j = 0; j < n2 / dj; j++) {
      double tl = g->startx[1] + j * dj * g->dx[1], th = g->startx[1] + (j + 1) * dj * g->dx[1];
      double vthf = romberg(tl, th, i_sdth, &m);
      double th_s3 = romberg(tl, th, i_sdth3, &m);
      double th_cd2 = romberg(tl, th, i_cdth2, &m);
      double *z = &g->zone[((size_t)(j + NG) * N1 + (i + NG)) * MDS_ZS_COUNT];
      const double *zu = &g->zone[((size_t)(j + 1 + NG) * N1 + (i + NG)) * MDS_ZS_COUNT];
      double c101 = rdr2 / vr;
      double c212 = th_cd2 / vthf;
      z[MDS_ZS_C000] = r2d2r / vr;
      z[MDS_ZS_C011] = -(r3i / vr) * (th_s3 / vthf);
      z[MDS_ZS_C101] = c101;
      z[MDS_ZS_C110] = c101;
      z[MDS_ZS_C202] = c101;
      z[MDS_ZS_C220] = c101;
      z[MDS_ZS_C212] = c212;
      z[MDS_ZS_C221] = c212;
      z[MDS_ZS_C022] = ((-c101) * z[MDS_ZS_GCOV0 + 2]) / z[MDS_ZS_GCOV0 + 0];
      z[MDS_ZS_C122] = ((-c212) * z[MDS_ZS_GCOV0 + 2]) / z[MDS_ZS_GCOV0 + 1];
      const int dk0 = gdk(g, i, j), dk1 = gdk(g, i, j + 1);
      double dA;
      if (dk1 < dk0) dA = zu[MDS_ZS_AREA1] + zu[MDS_ZS_AREA1];
      else if (dk1 > dk0) dA = 0.5 * zu[MDS_ZS_AREA1];
      else dA = zu[MDS_ZS_AREA1];
      dA -= z[MDS_ZS_AREA1];
      z[MDS_ZS_C111] = dA / z[MDS_ZS_VOLUME] - c212;
    }
  }
  g->volume = 0.0;
  g->area[0] = g->area[1] = g->area[2] = 0.0;
  return 0;
}

int mds_grid_build(mds_grid *g)
{
  if (!g || g->ndim < 1 || g->ndim > 3) return MDS_EINVAL;
  for (int d = 0; d < 3; d++) {
    if (d >= g->ndim) { g->n[d] = 1; }
    if (g->n[d] < 1) return MDS_EINVAL;
    g->alpha[d] = g->beta[d] = g->Gamma[d] = g->rcoef[d] = NULL;
  }
  g->zone = NULL;
  g->dj = g->dk = NULL;
  g->maxjlev = g->maxklev = 0;
  for (int t = 0; t < 2; t++) g->lalpha[t] = g->lbeta[t] = g->lGamma[t] = g->lrcoef[t] = NULL;
  if (g->geom == MDS_GEOM_SPHERICAL && g->rcoord == MDS_RCOORD_SINH && !(g->rtrans > 0.0)) return MDS_EINVAL;
  if (g->dendritic && !(g->geom == MDS_GEOM_SPHERICAL && g->ndim >= 2)) return MDS_EINVAL;
  int rc;
  if (g->geom == MDS_GEOM_SPHERICAL && g->ndim == 2) {
    /* 2D (r, theta): the 3D records of one phi zone spanning 2 pi, the
     * width the reference gives the absent direction of a spherical grid
     * (init_coords, dx[2] = 2 pi) */
    g->startx[2] = 0.0;
    g->dx[2] = 2.0 * M_PI;
  }
  if (g->geom == MDS_GEOM_SPHERICAL && g->ndim >= 2) {
    rc = build_spherical3d(g);
    if (!rc && g->ndim == 2) {
      /* a 2D grid has no phi faces: their area is zero, as in the reference */
      const size_t nf = (size_t)(g->n[0] + 2 * NG) * (g->n[1] + 2 * NG);
      for (size_t q = 0; q < nf; q++) g->zone[q * MDS_ZS_COUNT + MDS_ZS_AREA0 + 2] = 0.0;
    }
  }
  else if (g->geom == MDS_GEOM_SPHERICAL) rc = build_spherical1d(g);
  else rc = build_cartesian(g);
  if (rc) mds_grid_free(g);
  return rc;
}

int mds_grid_build_batch(mds_grid *g, int count)
{
  if (!g || count < 0) return MDS_EINVAL;
  int rc = MDS_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
  for (int q = 0; q < count; q++) {
    int r = mds_grid_build(&g[q]);
    if (r < rc) rc = r;
  }
  return rc;
}

void mds_grid_free(mds_grid *g)
{
  if (!g) return;
  for (int d = 0; d < 3; d++) {
    free(g->alpha[d]); free(g->beta[d]); free(g->Gamma[d]); free(g->rcoef[d]);
    g->alpha[d] = g->beta[d] = g->Gamma[d] = g->rcoef[d] = NULL;
  }
  free(g->zone);
  g->zone = NULL;
  free(g->dj);
  free(g->dk);
  g->dj = g->dk = NULL;
  for (int t = 0; t < 2; t++) {
    free(g->lalpha[t]); free(g->lbeta[t]); free(g->lGamma[t]); free(g->lrcoef[t]);
    g->lalpha[t] = g->lbeta[t] = g->lGamma[t] = g->lrcoef[t] = NULL;
  }
}

/* Radial shell geometry of the GR effective monopole (gravity.c:275-375),
 * evaluated on the host with the same libm calls and operation order as
 * the CPU restatement, so the device starts from the same bits.  out holds
 * four arrays of n[0] + 1 doubles:
 *   [0]  redge[i]  = r(|startx + i dx|)                      (face radius)
 *   [1]  r34[i]    = (4 pi) redge[i]^3                       (face pressure term)
 *   [2]  drdx[i]   = dr/dx(startx + i dx)                    (face gravity factor)
 *   [3]  vsh[i]    = 4 pi/3 (redge[i+1]^3 - redge[i]^3)      (shell volume, i < n[0])
 * Returns MDS_EINVAL for a grid that is not spherical. */
int mds_gr_shell_geometry(const mds_grid *g, double *out)
{
  if (!g || !out || g->geom != MDS_GEOM_SPHERICAL) return MDS_EINVAL;
  maps m;
  maps_init(&m, g);
  const int n1 = g->n[0], N = n1 + 1;
  double *re = out, *r34 = out + N, *dr = out + 2 * N, *vs = out + 3 * N;
  for (int i = 0; i <= n1; i++) {
    re[i] = rx(&m, fabs(i * g->dx[0] + g->startx[0]));
    r34[i] = (4.0 * M_PI) * pow(re[i], 3);
    dr[i] = drx(&m, g->startx[0] + i * g->dx[0]);
  }
  for (int i = 0; i < n1; i++) vs[i] = ((4.0 * M_PI) / 3.0) * (pow(re[i + 1], 3) - pow(re[i], 3));
  vs[n1] = 0.0;
  return MDS_OK;
}

/* ------------------------------------------------------------------ */
/* Multipole gravity tables (gravity.c:499-1110) */
#include "../common/mds_mp.h"
#define MP_LMAXC 15

static double mp_factorial_ratio(int n, int m)
{
  double fact = 1.0;
  while (n > m) { fact *= n; n--; }
  while (m > n) { fact /= m; m--; }
  return fact;
}

/* associated Legendre P[l,m](cos th), Condon-Shortley phase (gravity.c:579-607) */
static void mp_plms(double *plm, double th, int lmax, int mmax)
{
  const double cth = cos(th);
  for (int m = 0; m <= mmax; m++) {
    const double somx2 = sqrt((1.0 - cth) * (1.0 + cth));
    double fact = 1.0, pmm = 1.0;
    for (int i = 1; i <= m; i++) {
      pmm = ((-pmm) * fact) * somx2;
      fact = fact + 2.0;
    }
    plm[MDS_MP_LM(m, m)] = pmm;
    if (m != lmax) {
      plm[MDS_MP_LM(m + 1, m)] = (cth * ((2.0 * m) + 1.0)) * pmm;
      for (int l = m + 2; l <= lmax; l++)
        plm[MDS_MP_LM(l, m)] = (((cth * ((2.0 * l) - 1.0)) * plm[MDS_MP_LM(l - 1, m)]) -
                                ((l + m - 1) * plm[MDS_MP_LM(l - 2, m)])) / (l - m);
    }
  }
}

static double mp_norm(int l, int m)
{
  const double norm_fact = (m == 0) ? 1.0 : 1.41421356237309504880;
  return norm_fact * sqrt((((2.0 * l) + 1.0) / (4.0 * M_PI)) * mp_factorial_ratio(l - m, l + m));
}

/* real harmonics at (th, phi) (gravity.c:518-577) */
static void mp_ylms(double *ylm, double th, double phi, int lmax, int mmax)
{
  mp_plms(ylm, th, lmax, mmax);
  for (int m = 0; m <= mmax; m++) {
    const double smphi = sin(m * phi), cmphi = cos(m * phi);
    for (int l = m; l <= lmax; l++) {
      ylm[MDS_MP_LM(l, m)] *= mp_norm(l, m);
      if (m > 0) {
        ylm[MDS_MP_LM(l, -m)] = ylm[MDS_MP_LM(l, m)] * smphi;
        ylm[MDS_MP_LM(l, m)] *= cmphi;
      }
    }
  }
}

/* zone integrals of Y[l,m] sin(th) (gravity.c:609-697) */
static void mp_ylm_ints(double *slm, double thmin, double thmax, double phimin, double phimax, int lmax, int mmax)
{
  double plmp[MP_LMAXC * MP_LMAXC], plmm[MP_LMAXC * MP_LMAXC];
  const double cmin = cos(thmin), cmax = cos(thmax);
  mp_plms(plmm, thmin, lmax, mmax);
  mp_plms(plmp, thmax, lmax, mmax);
  for (int m = 0; m <= mmax; m++) {
    if (m > 1) {
      slm[MDS_MP_LM(m, m)] = (((m * (2 * m - 1) * (2 * m - 3)) * slm[MDS_MP_LM(m - 2, m - 2)] -
                               (cmax * plmp[MDS_MP_LM(m, m)])) + (cmin * plmm[MDS_MP_LM(m, m)])) / (m + 1.0);
    } else if (m == 0) {
      slm[MDS_MP_LM(0, 0)] = cmin - cmax;
    } else {
      slm[MDS_MP_LM(1, 1)] = 0.5 * ((((cmax * sqrt(1 - (cmax * cmax))) - thmax) - (cmin * sqrt(1 - (cmin * cmin)))) + thmin);
    }
    if (m != lmax) {
      slm[MDS_MP_LM(m + 1, m)] = ((((1 - (cmax * cmax)) * plmp[MDS_MP_LM(m, m)]) -
                                   ((1 - (cmin * cmin)) * plmm[MDS_MP_LM(m, m)])) * (2 * m + 1)) / (m + 2.0);
      for (int l = m + 2; l <= lmax; l++)
        slm[MDS_MP_LM(l, m)] = ((((l - 2) * (l + m - 1)) * slm[MDS_MP_LM(l - 2, m)]) +
                                ((2 * l - 1) * (((1 - (cmax * cmax)) * plmp[MDS_MP_LM(l - 1, m)]) -
                                                ((1 - (cmin * cmin)) * plmm[MDS_MP_LM(l - 1, m)])))) /
                               ((l + 1) * (l - m));
    }
  }
  for (int m = 0; m <= mmax; m++) {
    double smphi, cmphi;
    if (m == 0) {
      smphi = phimax - phimin;
      cmphi = 0.0;
    } else {
      smphi = (sin(m * phimax) - sin(m * phimin)) / m;
      cmphi = (cos(m * phimin) - cos(m * phimax)) / m;
    }
    for (int l = m; l <= lmax; l++) {
      slm[MDS_MP_LM(l, m)] *= mp_norm(l, m);
      if (m > 0) {
        slm[MDS_MP_LM(l, -m)] = slm[MDS_MP_LM(l, m)] * cmphi;
        slm[MDS_MP_LM(l, m)] *= smphi;
      } else {
        slm[MDS_MP_LM(l, 0)] *= smphi;
      }
    }
  }
}

void mds_mp_tab_free(mds_mp_tab *t)
{
  if (!t) return;
  free(t->S); free(t->Soff); free(t->Y); free(t->rc); free(t->rd); free(t->Cf); free(t->Df);
  free(t->rm); free(t->wx); free(t->wy); free(t->wz);
  memset(t, 0, sizeof(*t));
}

int mds_mp_tab_build(const mds_grid *g, int lmax, mds_mp_tab *t)
{
  if (!g || !t || g->geom != MDS_GEOM_SPHERICAL || g->ndim < 2 || lmax < 1 || lmax > MDS_MP_LMAX) return MDS_EINVAL;
  memset(t, 0, sizeof(*t));
  maps m;
  maps_init(&m, g);
  const double GN = 6.67384e-8;
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2], L1 = lmax + 1;
  t->lmax = lmax;
  t->mmax = g->ndim == 3 ? lmax : 0;
  t->nlm = L1 * L1;
  t->n1 = n1; t->n2 = n2; t->n3 = n3; t->ndim = g->ndim;
  t->maxjlev = g->dj ? g->maxjlev : 0;
  t->maxklev = g->dk ? g->maxklev : 0;
  const int NJ = t->maxjlev + 1, NK = t->maxklev + 1, nlm = t->nlm;
  t->Soff = calloc((size_t)NJ * NK, sizeof(long long));
  long long tot = 0;
  for (int lj = 0; lj < NJ; lj++)
    for (int lk = 0; lk < NK; lk++) {
      t->Soff[lj * NK + lk] = tot;
      tot += (long long)(n2 >> lj) * (n3 >> lk);
    }
  t->nS = tot * nlm;
  t->S = calloc((size_t)t->nS, sizeof(double));
  t->Y = calloc((size_t)(n2 + 1) * (n3 + 1) * nlm, sizeof(double));
  t->rc = calloc((size_t)n1 * L1, sizeof(double));
  t->rd = calloc((size_t)n1 * L1, sizeof(double));
  t->Cf = calloc((size_t)(n1 + 1) * L1, sizeof(double));
  t->Df = calloc((size_t)(n1 + 1) * L1, sizeof(double));
  t->rm = calloc((size_t)n1 + 1, sizeof(double));
  t->wx = calloc((size_t)2 * n1, sizeof(double));
  t->wy = calloc((size_t)2 * NJ * n2, sizeof(double));
  t->wz = calloc((size_t)2 * NK * n3, sizeof(double));
  if (!t->Soff || !t->S || !t->Y || !t->rc || !t->rd || !t->Cf || !t->Df || !t->rm || !t->wx || !t->wy || !t->wz) {
    mds_mp_tab_free(t);
    return MDS_ENOMEM;
  }
  double buf[MP_LMAXC * MP_LMAXC];
  /* zone integrals (gravity.c:739-756): theta edges of the row's level, phi
   * edges of the column's level; 2D grids integrate phi over [0, 2 pi] */
  for (int lj = 0; lj < NJ; lj++)
    for (int lk = 0; lk < NK; lk++)
      for (int j = 0; j < (n2 >> lj); j++)
        for (int k = 0; k < (n3 >> lk); k++) {
          const double thmin = thx(&m, g->startx[1] + (j * (1 << lj)) * g->dx[1]);
          const double thmax = thx(&m, g->startx[1] + ((j + 1) * (1 << lj)) * g->dx[1]);
          double phimin = 0.0, phimax = 2.0 * M_PI;
          if (g->ndim == 3) {
            phimin = g->startx[2] + (k * (1 << lk)) * g->dx[2];
            phimax = g->startx[2] + ((k + 1) * (1 << lk)) * g->dx[2];
          }
          mp_ylm_ints(buf, thmin, thmax, phimin, phimax, lmax, t->mmax);
          memcpy(t->S + (t->Soff[lj * NK + lk] + (long long)j * (n3 >> lk) + k) * nlm, buf, sizeof(double) * nlm);
        }
  /* corner harmonics (gravity.c:869-873) */
  for (int J = 0; J <= n2; J++)
    for (int K = 0; K <= n3; K++) {
      const double thm = thx(&m, g->startx[1] + J * g->dx[1]);
      const double phm = g->startx[2] + K * g->dx[2];
      mp_ylms(buf, thm, phm, lmax, t->mmax);
      memcpy(t->Y + ((size_t)J * (n3 + 1) + K) * nlm, buf, sizeof(double) * nlm);
    }
  /* radial weights of the moments (gravity.c:758-766) */
  for (int i = 0; i < n1; i++) {
    const double rmin = rx(&m, g->startx[0] + i * g->dx[0]), rmax = rx(&m, g->startx[0] + (i + 1) * g->dx[0]);
    for (int l = 0; l <= lmax; l++) {
      t->rc[i * L1 + l] = (pow(rmax, l + 3) * (1.0 - pow(rmin / rmax, l + 3))) / (l + 3);
      double rd;
      if (l == 0) rd = 0.5 * ((rmax * rmax) - (rmin * rmin));
      else if (l == 1) rd = rmax - rmin;
      else if (l == 2) rd = (rmin > 1.0e-16) ? log(rmax / rmin) : 0.0;
      else rd = (rmin > 1.0e-16) ? (pow(rmin, 2 - l) - pow(rmax, 2 - l)) / (l - 2) : 0.0;
      t->rd[i * L1 + l] = rd;
    }
    /* radial centroid weights (gravity.c:932-939) */
    const double rc = rx(&m, g->beta[0][i + NG] / g->Gamma[0][i + NG]);
    t->wx[2 * i] = (rmax - rc) / (rmax - rmin);
    t->wx[2 * i + 1] = (rc - rmin) / (rmax - rmin);
  }
  /* corner factors (gravity.c:877-904) */
  for (int ii = 0; ii <= n1; ii++) {
    const double rm = rx(&m, g->startx[0] + ii * g->dx[0]);
    t->rm[ii] = rm;
    for (int l = 1; l <= lmax; l++) {
      t->Cf[ii * L1 + l] = (((-4.0 * M_PI) * GN) / (2 * l + 1)) / pow(rm, l + 1);
      t->Df[ii * L1 + l] = (((-4.0 * M_PI) * GN) / (2 * l + 1)) * pow(rm, l);
    }
  }
  /* theta and phi centroid weights per level (gravity.c:944-963) */
  for (int lj = 0; lj < NJ; lj++)
    for (int j = 0; j < (n2 >> lj); j++) {
      const double thm = thx(&m, g->startx[1] + (j * (1 << lj)) * g->dx[1]);
      const double thp = thx(&m, g->startx[1] + ((j + 1) * (1 << lj)) * g->dx[1]);
      const double *be = g->lbeta[0] ? g->lbeta[0] + (size_t)lj * (n2 + 2 * NG) : g->beta[1];
      const double *ga = g->lGamma[0] ? g->lGamma[0] + (size_t)lj * (n2 + 2 * NG) : g->Gamma[1];
      const double thc = thx(&m, be[j + NG] / ga[j + NG]);
      t->wy[2 * (lj * n2 + j)] = (thp - thc) / (thp - thm);
      t->wy[2 * (lj * n2 + j) + 1] = (thc - thm) / (thp - thm);
    }
  for (int lk = 0; lk < NK; lk++)
    for (int k = 0; k < (n3 >> lk); k++) {
      const double phm = g->startx[2] + (k * (1 << lk)) * g->dx[2];
      const double php = g->startx[2] + ((k + 1) * (1 << lk)) * g->dx[2];
      const double phc = 0.5 * (phm + php);
      t->wz[2 * (lk * n3 + k)] = (php - phc) / (php - phm);
      t->wz[2 * (lk * n3 + k) + 1] = (phc - phm) / (php - phm);
    }
  return MDS_OK;
}
