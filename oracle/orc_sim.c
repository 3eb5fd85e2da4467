/*
 * orc_sim.c -- grid, geometry, boundaries, flux sweeps, RK2 push and the main
 * loop of the CPU oracle (test infrastructure only; see oracle.h).
 *
 * Restates:
 *   metrics/metric_cartesian.c:47-295       Cartesian moments / volume / area / metric
 *   metrics/metric_spherical_general.c:149-303, 558-1051  spherical 1D geometry
 *   metrics/metric_spherical_logr.c:3-38    log-r radial map r = exp(x)
 *   utils.c:32-62                           Romberg
 *   bc.c:705-1421                           ghost fills (OUTFLOW/REFLECT/EXTRAP/PERIODIC)
 *   eos/gamma/eos_gamma.c:5-57              gamma-law EOS
 *   rad_fluid.c:195-214, gravity.c:46-64    geometric and fixed-gravity sources
 *   step.c:482-567, 659-764, 1088-1430      flux sweeps on uniform (non-dendritic) grids
 *   step.c:327-478                          push (RK stage update)
 *   step.c:28-129                           step()
 *   step.c:163-203, main.c:206-252          estimate_dt and the dt policy
 *   fluid.c:75-89, 189-210                  shock flags / transverse_shock
 *
 * Layout: AoS zone records p[zone][nvars] in natural order, i fastest, with
 * NG ghost zones on both sides of each active direction.
 *
 * Shock-flag semantics.  The reference tags radial shock faces and reads the
 * tags for transverse faces inside the same parallel loop (step.c:711-755,
 * 1196-1418), so its result depends on iteration order.  The oracle (and the
 * device path) use the order-free reading: all radial faces of a stage first,
 * then the transverse faces.  Flags start at 0 everywhere (malloc_rank1 zeroes,
 * utils.c:171), set_shock_flag(0) resets interior zones to -1 once per step
 * (fluid.c:82-89, step.c:72), ghost zones keep 0, and 0 means "shocked".
 * Hence transverse faces whose lower zone is a ghost always use HLLE.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define OMIN(a, b) ((a) < (b) ? (a) : (b))
#define OMAX(a, b) ((a) > (b) ? (a) : (b))
#define OSQR(x) ((x) * (x))
#define OCUBE(x) ((x) * (x) * (x))
#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif

#define NEOS 5
#define E_PRESS 0
#define E_CS 1
#define E_TEMP 2
#define E_GAMMA 4
#define GNEWT 6.67384e-8     /* constants.h:3 */

struct orc_sim {
  orc_config c;
  int g[3];          /* ghost depth per direction (NG if active, else 0) */
  int N[3];          /* padded extent per direction */
  long nall;
  int nvars, nhydro, ninterp;
  double *p, *ph, *eos, *src;
  signed char *shock;
  orc_zone_geom *geom;
  double *mom[3][3];   /* [alpha,beta,Gamma][dir] length n[d]+2NG, index 0 = zone -NG */
  double *flux[3];     /* [zone][nvars] flux through the lower face in direction d */
  double t, dt, dtold, min_dt;
  int istep, firstc;
  long nshock, nhlle;
  /* dendritic levels (utils.c:181-590); NULL on uniform grids (every level 1) */
  int *dj;             /* [i + NG], i in [-NG, n1 + NG) */
  int *dk;             /* [(i + NG) * (n2 + 2 NG) + j + NG], j in [-NG, n2/dj + NG) */
  int maxjlev, maxklev;
  /* theta (0) and phi (1) moments per level l at l * (n + 2 NG), index 0 = zone -NG
   * (alpha1s/beta1s/Gamma1s, alpha2s/..., metric_spherical_general.c:592-669) */
  double *lmom[3][2];
  unsigned char *live; /* zones the dendritic sweeps read: interior and ghosts */
  /* GR effective monopole (use_monopole == 2): the TOV factor Gamma at the
   * radial faces, kept between calls (the reference's static gtov, set once
   * from the first state); the enclosed TOV mass, the lapses of the last
   * call and its iteration count */
  double *gtov, *gr_mr, *gr_lapse, *gr_lapse_edge;
  int gtov_set, gr_iter, gr_fail;
};

/* DJS/DKS and the level-local zone counts (decs.h:569-590) */
static int djs(const orc_sim *s, int i) { return s->dj ? s->dj[i + ORC_NG] : 1; }
static int dks(const orc_sim *s, int i, int j)
{
  return s->dk ? s->dk[(long)(i + ORC_NG) * (s->c.n[1] + 2 * ORC_NG) + j + ORC_NG] : 1;
}
static int njs(const orc_sim *s, int i) { return s->c.n[1] / djs(s, i); }
static int nks(const orc_sim *s, int i, int j) { return s->c.n[2] / dks(s, i, j); }
static int lev_of(int f)
{
  int l = 0;
  while (f >>= 1) l++;
  return l;
}

static long zid(const orc_sim *s, int i, int j, int k)
{
  return (long)(i + s->g[0]) + (long)s->N[0] * ((long)(j + s->g[1]) + (long)s->N[1] * (long)(k + s->g[2]));
}

/* ------------------------------------------------------------------ */
/* Romberg integration, utils.c:14-62 */
static double rtrap(double x0, double x1, int rstage, double (*f)(double, const void *), const void *ctx)
{
  unsigned long long nseg = 1 << rstage;
  double dx = (x1 - x0) / nseg;
  double sum = 0.0;
  if (rstage == 0) return (0.5 * dx) * (f(x0, ctx) + f(x1, ctx));
  for (unsigned long long i = 0; i < (unsigned long long)(1 << (rstage - 1)); i++)
    sum += f(x0 + (2 * i + 1) * dx, ctx);
  return dx * sum;
}

double orc_romb(double x0, double x1, double (*f)(double, const void *), const void *ctx)
{
  double rvec[16];
  double R0, R1;
  rvec[0] = rtrap(x0, x1, 0, f, ctx);
  for (int rstage = 1; rstage <= 15; rstage++) {
    R0 = rvec[0];
    rvec[0] = (0.5 * R0) + rtrap(x0, x1, rstage, f, ctx);
    for (int m = 1; m <= rstage; m++) {
      unsigned long long fm = 1 << (2 * m);
      R1 = rvec[m];
      rvec[m] = (1.0 / (fm - 1)) * ((fm * rvec[m - 1]) - R0);
      R0 = R1;
    }
    if (fabs((rvec[rstage] - rvec[rstage - 1]) / rvec[rstage]) < 1.e-15
        || (rstage > 2 && fabs(rvec[rstage]) < 1.e-15))
      return rvec[rstage];
  }
  return rvec[15];
}

/* ------------------------------------------------------------------ */
/* spherical maps: r(x) (uniform: metric_spherical_general.c:149-152; log:
 * metric_spherical_logr.c) and the POLYTH theta map with the io.c defaults
 * poly_xt=1.2, poly_alpha=4 (io.c:5014-5015, metric_spherical_general.c:222-243). */
typedef struct { int rcoord; double poly_norm, poly_xt, poly_alpha, rtrans; } sphmap;

/* sinh map: metric_spherical_general.c:159-161, 186-188, 206-208 */
static double r_of_x(const sphmap *m, double x)
{
  if (m->rcoord == ORC_RCOORD_LOG) return exp(x);
  if (m->rcoord == ORC_RCOORD_SINH) return m->rtrans * sinh(fabs(x) / m->rtrans);
  return fabs(x);
}
static double dr_dx(const sphmap *m, double x)
{
  if (m->rcoord == ORC_RCOORD_LOG) return exp(x);
  if (m->rcoord == ORC_RCOORD_SINH) return cosh(fabs(x) / m->rtrans);
  return 1.0;
}
static double d2r_dx2(const sphmap *m, double x)
{
  if (m->rcoord == ORC_RCOORD_LOG) return exp(x);
  if (m->rcoord == ORC_RCOORD_SINH) return sinh(fabs(x) / m->rtrans) / m->rtrans;
  return 0.0;
}
static double th_of_x(const sphmap *m, double x)
{
  if (x < -1.0) x = -2.0 - x;
  if (x > 1.0) x = 2.0 - x;
  return ((m->poly_norm * x) * (1.0 + (pow(x / m->poly_xt, m->poly_alpha) / (m->poly_alpha + 1.0)))) + (0.5 * M_PI);
}
static double dth_dx(const sphmap *m, double x)
{
  if (x < -1.0) x = -2.0 - x;
  if (x > 1.0) x = 2.0 - x;
  return m->poly_norm * (1.0 + pow(x / m->poly_xt, m->poly_alpha));
}
static double d2th_dx2(const sphmap *m, double x)
{
  if (x < -1.0) x = -2.0 - x;
  if (x > 1.0) x = 2.0 - x;
  return ((m->poly_norm * m->poly_alpha) * pow(x / m->poly_xt, m->poly_alpha - 1.0)) / m->poly_xt;
}
/* integrands, metric_spherical_general.c:290-460 */
static double f_r2dr_dx(double x, const void *c) { const sphmap *m = c; return pow(r_of_x(m, x), 2) * dr_dx(m, x); }
static double f_xr1g(double x, const void *c) { const sphmap *m = c; return (x * pow(r_of_x(m, x), 2)) * dr_dx(m, x); }
static double f_xr2g(double x, const void *c) { const sphmap *m = c; return ((x * x) * pow(r_of_x(m, x), 2)) * dr_dx(m, x); }
static double f_r2d2r_dx2(double x, const void *c) { const sphmap *m = c; return pow(r_of_x(m, x), 2) * d2r_dx2(m, x); }
static double f_r3(double x, const void *c) { const sphmap *m = c; return pow(r_of_x(m, x), 3); }
static double f_r2dr_dx3(double x, const void *c) { const sphmap *m = c; return pow(r_of_x(m, x), 2) * pow(dr_dx(m, x), 3); }
static double f_rdr_dx2(double x, const void *c) { const sphmap *m = c; return r_of_x(m, x) * pow(dr_dx(m, x), 2); }
static double f_sthdth_dx(double x, const void *c) { const sphmap *m = c; return sin(th_of_x(m, x)) * dth_dx(m, x); }
static double f_sthdth_dx3(double x, const void *c) { const sphmap *m = c; return sin(th_of_x(m, x)) * pow(dth_dx(m, x), 3); }
static double f_sth3dth_dx(double x, const void *c) { const sphmap *m = c; return pow(sin(th_of_x(m, x)), 3) * dth_dx(m, x); }
static double f_sthd2th_dx2(double x, const void *c) { const sphmap *m = c; return sin(th_of_x(m, x)) * d2th_dx2(m, x); }
static double f_cthdth_dx2(double x, const void *c) { const sphmap *m = c; return cos(th_of_x(m, x)) * pow(dth_dx(m, x), 2); }
static double f_xth1g(double x, const void *c) { const sphmap *m = c; return (x * sin(th_of_x(m, x))) * dth_dx(m, x); }
static double f_xth2g(double x, const void *c) { const sphmap *m = c; return ((x * x) * sin(th_of_x(m, x))) * dth_dx(m, x); }

static void sphmap_init(sphmap *m, int rcoord, double rtrans)
{
  m->rcoord = rcoord;
  m->rtrans = rtrans;
  m->poly_xt = 1.2;
  m->poly_alpha = 4;
  m->poly_norm = ((0.5 * M_PI) * 1.0) / (1.0 + ((1.0 / (m->poly_alpha + 1.0)) * (1.0 / pow(m->poly_xt, m->poly_alpha))));
}

static void build_geometry_cartesian(orc_sim *s)
{
  const orc_config *c = &s->c;
  /* metric_cartesian.c:47-105 moments */
  for (int d = 0; d < 3; d++) {
    int n = c->n[d] + 2 * ORC_NG;
    for (int q = 0; q < 3; q++) s->mom[q][d] = calloc(n, sizeof(double));
    for (int i = -ORC_NG; i < c->n[d] + ORC_NG; i++) {
      double x0 = c->startx[d] + i * c->dx[d];
      double x1 = x0 + c->dx[d];
      s->mom[0][d][i + ORC_NG] = (OCUBE(x1) - OCUBE(x0)) / 3.0;
      s->mom[1][d][i + ORC_NG] = (OSQR(x1) - OSQR(x0)) / 2.0;
      s->mom[2][d][i + ORC_NG] = c->dx[d];
    }
  }
  /* metric_cartesian.c:134-295: uniform volume/area, unit metric, zero connection */
  double vol = 1.0;
  for (int d = 0; d < c->ndim; d++) vol *= c->dx[d];
  double area[3];
  for (int loc = 0; loc < c->ndim; loc++) {
    area[loc] = 1.0;
    for (int d = 0; d < c->ndim; d++) area[loc] *= c->dx[d];
    area[loc] /= c->dx[loc];
  }
  for (long z = 0; z < s->nall; z++) {
    orc_zone_geom *g = &s->geom[z];
    memset(g, 0, sizeof(*g));
    g->volume = vol;
    for (int loc = 0; loc < c->ndim; loc++) g->area[loc] = area[loc];
    for (int loc = 0; loc < 4; loc++)
      for (int d = 0; d < 3; d++) { g->gcov[loc][d] = 1.0; g->scale[loc][d] = 1.0; }
    for (int d = 0; d < 3; d++) g->gcon[d] = 1.0;
  }
}

/* Dendritic levels: compute_dj (utils.c:181-233), compute_dk (:240-368) and
 * the 2:1 checks of init_dendritic (:540-579).  Returns 0, or -1 when the
 * reference's check finds a 4:1 jump (it exits there) or when its fix-up of
 * a 1/2:1 jump would change a level (the fix-up leaves the ghost rows
 * unchanged, so such grids are rejected here). */
static int build_levels(orc_sim *s, const sphmap *m)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], n2 = c->n[1], n3 = c->n[2], NG = ORC_NG;
  const int S1 = n2 + 2 * NG;
  const double MAXA = 0.5;                      /* MAX_LOG2_ASPECT, utils.c:12 */
  int *nj = calloc(n1 + 2 * NG, sizeof(int));
  int *nk = calloc((size_t)(n1 + 2 * NG) * S1, sizeof(int));
  s->dj = calloc(n1 + 2 * NG, sizeof(int));
  s->dk = calloc((size_t)(n1 + 2 * NG) * S1, sizeof(int));
#define NJ(i) nj[(i) + NG]
#define NK(i, j) nk[(long)((i) + NG) * S1 + (j) + NG]
#define DK(i, j) s->dk[(long)((i) + NG) * S1 + (j) + NG]
  /* compute_dj */
  NJ(0) = 2;
  const double dth = th_of_x(m, c->startx[1] + n2 * c->dx[1]) - th_of_x(m, c->startx[1]);
  for (int i = 1; i < n1; i++) {
    double r0 = r_of_x(m, c->startx[0] + (i + 0.0) * c->dx[0]);
    double rc = r_of_x(m, c->startx[0] + (i + 0.5) * c->dx[0]);
    double r1 = r_of_x(m, c->startx[0] + (i + 1.0) * c->dx[0]);
    double ideal = rc * dth / (r1 - r0);
    NJ(i) = NJ(i - 1);
    if (log2(ideal / NJ(i)) > MAXA) NJ(i) *= 2;
    NJ(i) = OMIN(OMAX(NJ(i), NJ(i - 1)), n2);
  }
  for (int i = -NG; i < 0; i++) NJ(i) = NJ(-i - 1);
  for (int i = n1; i < n1 + NG; i++) NJ(i) = NJ(n1 - 1);
  for (int i = -NG; i < n1 + NG; i++) s->dj[i + NG] = n2 / NJ(i);
  /* compute_dk; a 2D grid keeps every row at n3 = 1 zone (utils.c:311-316) */
  const int nk_min = 4, nk_max = n3;
  const double dph = n3 * c->dx[2];
  for (int i = 0; i < n1 && c->ndim == 2; i++)
    for (int j = 0; j < NJ(i); j++) NK(i, j) = n3;
  for (int i = 0; i < n1 && c->ndim == 3; i++) {
    double r0 = r_of_x(m, c->startx[0] + (i + 0.0) * c->dx[0]);
    double rc = r_of_x(m, c->startx[0] + (i + 0.5) * c->dx[0]);
    double r1 = r_of_x(m, c->startx[0] + (i + 1.0) * c->dx[0]);
    double dr = r1 - r0;
    NK(i, 0) = nk_min;
    for (int j = 1; j < NJ(i); j++) {
      double thc = th_of_x(m, c->startx[1] + (j + 0.5) * s->dj[i + NG] * c->dx[1]);
      double ideal = rc * sin(thc) * dph / dr;
      if (j < NJ(i) / 2) {
        NK(i, j) = NK(i, j - 1);
        if (log2(ideal / NK(i, j)) > MAXA) NK(i, j) *= 2;
        if (i == 0) NK(i, j) = OMIN(NK(i, j), nk_max);
        else if (NJ(i - 1) == NJ(i)) NK(i, j) = OMIN(OMAX(NK(i, j), NK(i - 1, j)), nk_max);
        else NK(i, j) = OMIN(OMAX(NK(i, j), NK(i - 1, j / 2)), nk_max);
      } else {
        NK(i, j) = NK(i, NJ(i) - 1 - j);
      }
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
  int bad = 0;
  for (int i0 = 0; i0 < n1; i0++) {
    const int i1 = i0 + 1;
    if (s->dj[i0 + NG] / s->dj[i1 + NG] != 2) continue;
    for (int jj = 0; jj < NJ(i0); jj++) {
      const int j10 = jj * s->dj[i0 + NG] / s->dj[i1 + NG], j11 = j10 + 1;
      const int a = DK(i0, jj) / DK(i1, j10), b = DK(i0, jj) / DK(i1, j11);
      if (a == 4 || b == 4 || a == 0 || b == 0) bad = 1;
    }
  }
  s->maxjlev = lev_of(s->dj[NG]);
  s->maxklev = 0;
  for (int i = -NG; i < n1 + NG; i++)
    for (int j = -NG; j < NJ(i) + NG; j++) s->maxklev = OMAX(s->maxklev, lev_of(DK(i, j)));
#undef NJ
#undef NK
#undef DK
  free(nj);
  free(nk);
  return bad ? -1 : 0;
}

/* 2D/3D spherical geometry, metric_spherical_general.c:592-1051: moments
 * per level (:592-669), volumes (:708-740), areas on the faces the reference
 * loops over (i <= n1, j <= nj) with the axis faces forced to zero
 * (:745-833), gcov/gcon/scale (:941-1051, ghosts included, the theta
 * integrals reflect through th_of_x), and the connection coefficients of
 * interior zones with the pressure-balanced Gamma^theta_thetatheta
 * dA/dV - Gamma^phi_thetaphi, dA taken across a dendritic k-level change as
 * the sum of the two finer faces or half the coarser one (:837-929).
 * Returns -1 if the dendritic levels are rejected (build_levels). */
static int build_geometry_spherical(orc_sim *s)
{
  const orc_config *c = &s->c;
  const int NG = ORC_NG;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  const int n1 = c->n[0], n2 = c->n[1], n3 = c->n[2];
  const double dph = c->ndim == 3 ? c->dx[2] : 2.0 * M_PI;
  int rc = 0;
  s->maxjlev = s->maxklev = 0;
  if (c->dendritic) rc = build_levels(s, &m);
  for (int q = 0; q < 3; q++) {
    s->mom[q][0] = calloc(n1 + 2 * NG, sizeof(double));
    s->lmom[q][0] = calloc((size_t)(s->maxjlev + 1) * (n2 + 2 * NG), sizeof(double));
    s->lmom[q][1] = calloc((size_t)(s->maxklev + 1) * (n3 + 2 * NG), sizeof(double));
    s->mom[q][1] = s->lmom[q][0];
    s->mom[q][2] = s->lmom[q][1];
  }
  /* init_interp_vol :592-669 */
  for (int i = -NG; i < n1 + NG; i++) {
    double x0 = c->startx[0] + i * c->dx[0], x1 = c->startx[0] + (i + 1) * c->dx[0];
    s->mom[2][0][i + NG] = orc_romb(x0, x1, f_r2dr_dx, &m);
    s->mom[1][0][i + NG] = orc_romb(x0, x1, f_xr1g, &m);
    s->mom[0][0][i + NG] = orc_romb(x0, x1, f_xr2g, &m);
  }
  for (int lev = 0; lev <= s->maxjlev; lev++) {
    const int sc = 1 << lev;
    double *G = s->lmom[2][0] + (size_t)lev * (n2 + 2 * NG), *B = s->lmom[1][0] + (size_t)lev * (n2 + 2 * NG);
    double *A = s->lmom[0][0] + (size_t)lev * (n2 + 2 * NG);
    for (int j = -NG; j < n2 / sc + NG; j++) {
      double x0 = c->startx[1] + j * sc * c->dx[1], x1 = c->startx[1] + (j + 1) * sc * c->dx[1];
      G[j + NG] = orc_romb(x0, x1, f_sthdth_dx, &m);
      B[j + NG] = orc_romb(x0, x1, f_xth1g, &m);
      A[j + NG] = orc_romb(x0, x1, f_xth2g, &m);
    }
  }
  for (int lev = 0; lev <= s->maxklev; lev++) {
    const int sc = 1 << lev;
    double *G = s->lmom[2][1] + (size_t)lev * (n3 + 2 * NG), *B = s->lmom[1][1] + (size_t)lev * (n3 + 2 * NG);
    double *A = s->lmom[0][1] + (size_t)lev * (n3 + 2 * NG);
    for (int k = -NG; k < n3 / sc + NG; k++) {
      double x0 = c->startx[2] + k * sc * c->dx[2], x1 = c->startx[2] + (k + 1) * sc * c->dx[2];
      G[k + NG] = sc * c->dx[2];
      B[k + NG] = (OSQR(x1) - OSQR(x0)) / 2.0;
      A[k + NG] = (OCUBE(x1) - OCUBE(x0)) / 3.0;
    }
  }
  const int kg = c->ndim == 3 ? NG : 0;
  const int top_axis = fabs((c->startx[1] + n2 * c->dx[1]) - 1) < 1.0e-15;
  for (int i = -NG; i < n1 + NG; i++) {
    double x0lo = c->startx[0] + i * c->dx[0], x0hi = c->startx[0] + (i + 1) * c->dx[0];
    double r2dr_int = orc_romb(x0lo, x0hi, f_r2dr_dx, &m);
    double vrf = r2dr_int;
    double rint_r2dr_dx3 = orc_romb(x0lo, x0hi, f_r2dr_dx3, &m);
    double r5 = 0.2 * (pow(r_of_x(&m, x0hi), 5) - pow(r_of_x(&m, x0lo), 5));
    const int nj = njs(s, i);
    for (int j = -NG; j < nj + NG; j++) {
      double x1lo = c->startx[1] + j * djs(s, i) * c->dx[1], x1hi = c->startx[1] + (j + 1) * djs(s, i) * c->dx[1];
      double sthdth_int = orc_romb(x1lo, x1hi, f_sthdth_dx, &m);
      double vthf = sthdth_int;
      double thint_sthdth_dx3 = orc_romb(x1lo, x1hi, f_sthdth_dx3, &m);
      double thint_sth3dth_dx = orc_romb(x1lo, x1hi, f_sth3dth_dx, &m);
      const int dkz = c->ndim == 3 ? dks(s, i, j) : 1;
      orc_zone_geom g;
      memset(&g, 0, sizeof(g));
      g.volume = ((r2dr_int * sthdth_int) * dkz) * dph;                        /* :708-740 */
      if (i >= 0 && i <= n1 && j >= 0 && j <= nj) {                            /* :745-806 */
        double dph_int = dkz * dph;
        g.area[0] = (f_r2dr_dx(x0lo, &m) * sthdth_int) * dph_int;
        g.area[1] = (r2dr_int * f_sthdth_dx(x1lo, &m)) * dph_int;
        if (c->ndim == 3) g.area[2] = r2dr_int * sthdth_int;
        if ((j == 0 && c->startx[1] == -1) || (j == nj && top_axis)) g.area[1] = 0.0;   /* :808-833 */
      }
      /* init_gcov :941-1010 */
      g.gcov[0][0] = rint_r2dr_dx3 / vrf;
      g.gcov[0][1] = fabs(((r5 / vrf) * thint_sthdth_dx3) / vthf);
      g.gcov[0][2] = fabs(((r5 / vrf) * thint_sth3dth_dx) / vthf);
      g.gcov[1][0] = pow(dr_dx(&m, x0lo), 2);
      g.gcov[1][1] = (pow(r_of_x(&m, x0lo), 2) * thint_sthdth_dx3) / vthf;
      g.gcov[1][2] = (pow(r_of_x(&m, x0lo), 2) * thint_sth3dth_dx) / vthf;
      g.gcov[2][0] = g.gcov[0][0];
      g.gcov[2][1] = fabs((r5 / vrf) * pow(dth_dx(&m, x1lo), 2));
      g.gcov[2][2] = fabs((r5 / vrf) * pow(sin(th_of_x(&m, x1lo)), 2));
      if (c->ndim == 3)
        for (int d = 0; d < 3; d++) g.gcov[3][d] = g.gcov[0][d];
      for (int d = 0; d < 3; d++)                                              /* :1016-1032 */
        g.gcon[d] = (fabs(g.gcov[0][d]) > 1.0e-50) ? 1.0 / g.gcov[0][d] : 0.0;
      for (int loc = 0; loc <= c->ndim; loc++)                                 /* :1037-1051 */
        for (int d = 0; d < 3; d++) g.scale[loc][d] = sqrt(g.gcov[loc][d]);
      const int nk = c->ndim == 3 ? nks(s, i, j) : 1;
      for (int k = -kg; k < nk + kg; k++) s->geom[zid(s, i, j, k)] = g;
    }
  }
  /* init_conn :837-929, interior zones (after volumes and areas) */
  for (int i = 0; i < n1; i++) {
    double x0lo = c->startx[0] + i * c->dx[0], x0hi = c->startx[0] + (i + 1) * c->dx[0];
    double rint_r2d2r_dx2 = orc_romb(x0lo, x0hi, f_r2d2r_dx2, &m);
    double rint_r3 = orc_romb(x0lo, x0hi, f_r3, &m);
    double rint_rdr_dx2 = orc_romb(x0lo, x0hi, f_rdr_dx2, &m);
    double vrf = orc_romb(x0lo, x0lo + c->dx[0], f_r2dr_dx, &m);      /* Vr(x0lo) */
    for (int j = 0; j < njs(s, i); j++) {
      double x1lo = c->startx[1] + j * djs(s, i) * c->dx[1], x1hi = c->startx[1] + (j + 1) * djs(s, i) * c->dx[1];
      double vthf = orc_romb(x1lo, x1hi, f_sthdth_dx, &m);
      double thint_sthdth_dx3 = orc_romb(x1lo, x1hi, f_sthdth_dx3, &m);
      double thint_cthdth_dx2 = orc_romb(x1lo, x1hi, f_cthdth_dx2, &m);
      const int nk = c->ndim == 3 ? nks(s, i, j) : 1;
      for (int k = 0; k < nk; k++) {
        orc_zone_geom *g = &s->geom[zid(s, i, j, k)];
        memset(g->conn, 0, sizeof(g->conn));
        g->conn[0][0][0] = rint_r2d2r_dx2 / vrf;
        g->conn[0][1][1] = -(rint_r3 / vrf) * (thint_sthdth_dx3 / vthf);
        g->conn[1][0][1] = rint_rdr_dx2 / vrf;
        g->conn[1][1][0] = g->conn[1][0][1];
        g->conn[2][0][2] = g->conn[1][0][1];
        g->conn[2][2][0] = g->conn[2][0][2];
        g->conn[2][1][2] = thint_cthdth_dx2 / vthf;
        g->conn[2][2][1] = g->conn[2][1][2];
        g->conn[0][2][2] = ((-g->conn[2][2][0]) * g->gcov[0][2]) / g->gcov[0][0];
        g->conn[1][2][2] = ((-g->conn[2][2][1]) * g->gcov[0][2]) / g->gcov[0][1];
        double dA;
        const int dk0 = c->ndim == 3 ? dks(s, i, j) : 1, dk1 = c->ndim == 3 ? dks(s, i, j + 1) : 1;
        if (dk1 < dk0)
          dA = s->geom[zid(s, i, j + 1, 2 * k)].area[1] + s->geom[zid(s, i, j + 1, 2 * k + 1)].area[1];
        else if (dk1 > dk0)
          dA = 0.5 * s->geom[zid(s, i, j + 1, k / 2)].area[1];
        else
          dA = s->geom[zid(s, i, j + 1, k)].area[1];
        dA -= g->area[1];
        double dV = g->volume;
        g->conn[1][1][1] = dA / dV - g->conn[2][1][2];
      }
    }
  }
  return rc;
}

/* 1D spherical geometry, metric_spherical_general.c:558-1051 with NDIM=1
 * (x1 spans [-1,1], dx[2]=2*pi; init_coords :558-590). */
static void build_geometry_spherical1d(orc_sim *s)
{
  const orc_config *c = &s->c;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  const double dx2 = 2.0 * M_PI;
  const double x1lo = -1.0, x1hi = 1.0;
  int n = c->n[0];
  for (int q = 0; q < 3; q++) {
    s->mom[q][0] = calloc(n + 2 * ORC_NG, sizeof(double));
    s->mom[q][1] = calloc(1 + 2 * ORC_NG, sizeof(double));
    s->mom[q][2] = calloc(1 + 2 * ORC_NG, sizeof(double));
  }
  /* init_interp_vol :592-603 */
  for (int i = -ORC_NG; i < n + ORC_NG; i++) {
    double x0 = c->startx[0] + i * c->dx[0];
    double x1 = c->startx[0] + (i + 1) * c->dx[0];
    s->mom[2][0][i + ORC_NG] = orc_romb(x0, x1, f_r2dr_dx, &m);
    s->mom[1][0][i + ORC_NG] = orc_romb(x0, x1, f_xr1g, &m);
    s->mom[0][0][i + ORC_NG] = orc_romb(x0, x1, f_xr2g, &m);
  }
  double sthdth_int = orc_romb(x1lo, x1hi, f_sthdth_dx, &m);
  double vthf = sthdth_int;
  double thint_sthdth_dx3 = orc_romb(x1lo, x1hi, f_sthdth_dx3, &m);
  double thint_sth3dth_dx = orc_romb(x1lo, x1hi, f_sth3dth_dx, &m);
  double thint_sthd2th_dx2 = orc_romb(x1lo, x1hi, f_sthd2th_dx2, &m);
  double thint_cthdth_dx2 = orc_romb(x1lo, x1hi, f_cthdth_dx2, &m);
  for (int i = -ORC_NG; i < n + ORC_NG; i++) {
    orc_zone_geom *g = &s->geom[zid(s, i, 0, 0)];
    memset(g, 0, sizeof(*g));
    double x0lo = c->startx[0] + i * c->dx[0];
    double x0hi = c->startx[0] + (i + 1) * c->dx[0];
    double r2dr_int = orc_romb(x0lo, x0hi, f_r2dr_dx, &m);
    /* init_volume :708-740 */
    g->volume = ((r2dr_int * sthdth_int) * 1) * dx2;
    /* init_area :745-820 (faces istart..istop only) */
    if (i >= 0 && i <= n) g->area[0] = (f_r2dr_dx(x0lo, &m) * sthdth_int) * (1 * dx2);
    /* init_gcov :941-1010 */
    double vrf = r2dr_int;
    double rint_r2dr_dx3 = orc_romb(x0lo, x0hi, f_r2dr_dx3, &m);
    double r5 = 0.2 * (pow(r_of_x(&m, x0hi), 5) - pow(r_of_x(&m, x0lo), 5));
    g->gcov[0][0] = rint_r2dr_dx3 / vrf;
    g->gcov[0][1] = fabs(((r5 / vrf) * thint_sthdth_dx3) / vthf);
    g->gcov[0][2] = fabs(((r5 / vrf) * thint_sth3dth_dx) / vthf);
    g->gcov[1][0] = pow(dr_dx(&m, x0lo), 2);
    g->gcov[1][1] = (pow(r_of_x(&m, x0lo), 2) * thint_sthdth_dx3) / vthf;
    g->gcov[1][2] = (pow(r_of_x(&m, x0lo), 2) * thint_sth3dth_dx) / vthf;
    /* init_gcon :1016-1032 */
    for (int d = 0; d < 3; d++)
      g->gcon[d] = (fabs(g->gcov[0][d]) > 1.0e-50) ? 1.0 / g->gcov[0][d] : 0.0;
    /* init_scale :1037-1051 (loc 0..NDIM) */
    for (int loc = 0; loc <= 1; loc++)
      for (int d = 0; d < 3; d++) g->scale[loc][d] = sqrt(g->gcov[loc][d]);
    /* init_conn :837-929 (interior zones) */
    if (i >= 0 && i < n) {
      double rint_r2d2r_dx2 = orc_romb(x0lo, x0hi, f_r2d2r_dx2, &m);
      double rint_r3 = orc_romb(x0lo, x0hi, f_r3, &m);
      double rint_rdr_dx2 = orc_romb(x0lo, x0hi, f_rdr_dx2, &m);
      double vrf2 = orc_romb(x0lo, x0lo + c->dx[0], f_r2dr_dx, &m);   /* Vr(x0lo) */
      g->conn[0][0][0] = rint_r2d2r_dx2 / vrf2;
      g->conn[0][1][1] = -(rint_r3 / vrf2) * (thint_sthdth_dx3 / vthf);
      g->conn[1][0][1] = rint_rdr_dx2 / vrf2;
      g->conn[1][1][0] = g->conn[1][0][1];
      g->conn[2][0][2] = g->conn[1][0][1];
      g->conn[2][2][0] = g->conn[2][0][2];
      g->conn[2][1][2] = thint_cthdth_dx2 / vthf;
      g->conn[2][2][1] = g->conn[2][1][2];
      g->conn[0][2][2] = ((-g->conn[2][2][0]) * g->gcov[0][2]) / g->gcov[0][0];
      g->conn[1][2][2] = ((-g->conn[2][2][1]) * g->gcov[0][2]) / g->gcov[0][1];
      g->conn[1][1][1] = thint_sthd2th_dx2 / vthf;
    }
  }
}

orc_sim *orc_create(const orc_config *cfg)
{
  for (int d = 0; d < cfg->ndim; d++)
    if ((cfg->bc_lo[d] == ORC_BC_PROB || cfg->bc_hi[d] == ORC_BC_PROB) && !cfg->prob_bounds && !cfg->prob_bounds_t)
      return NULL;
  orc_sim *s = calloc(1, sizeof(*s));
  s->c = *cfg;
  for (int d = 0; d < 3; d++) {
    if (d >= cfg->ndim) s->c.n[d] = 1;
    /* a 2D dendritic grid keeps phi ghosts around its one phi zone: the
     * dendritic pencils (kslopes0/1) read one zone either side in phi */
    const int dend2 = d == 2 && cfg->dendritic && cfg->geom == ORC_GEOM_SPHERICAL;
    s->g[d] = (d < cfg->ndim || dend2) ? ORC_NG : 0;
    s->N[d] = s->c.n[d] + 2 * s->g[d];
  }
  s->nall = (long)s->N[0] * s->N[1] * s->N[2];
  s->nvars = cfg->nvars;
  s->nhydro = cfg->nvars;
  s->ninterp = cfg->nvars + 2;
  s->p = calloc(s->nall * s->nvars, sizeof(double));
  s->ph = calloc(s->nall * s->nvars, sizeof(double));
  s->src = calloc(s->nall * s->nvars, sizeof(double));
  s->eos = calloc(s->nall * NEOS, sizeof(double));
  s->shock = calloc(s->nall, 1);
  s->geom = calloc(s->nall, sizeof(orc_zone_geom));
  for (int d = 0; d < 3; d++) s->flux[d] = calloc(s->nall * s->nvars, sizeof(double));
  if (cfg->dendritic && !(cfg->geom == ORC_GEOM_SPHERICAL && cfg->ndim >= 2)) {
    orc_destroy(s);
    return NULL;
  }
  if (cfg->geom == ORC_GEOM_SPHERICAL && cfg->ndim == 1) build_geometry_spherical1d(s);
  else if (cfg->geom == ORC_GEOM_SPHERICAL) {
    if (build_geometry_spherical(s)) {
      orc_destroy(s);
      return NULL;
    }
  } else build_geometry_cartesian(s);
  if (s->dj) {
    /* zones read by the dendritic sweeps: each level's interior and NG ghosts */
    s->live = calloc(s->nall, 1);
    for (int i = -ORC_NG; i < s->c.n[0] + ORC_NG; i++)
      for (int j = -ORC_NG; j < njs(s, i) + ORC_NG; j++)
        for (int k = -ORC_NG; k < nks(s, i, j) + ORC_NG; k++) s->live[zid(s, i, j, k)] = 1;
  }
  if (cfg->use_monopole == 2) {
    s->gtov = calloc(s->c.n[0] + 1, sizeof(double));
    s->gr_mr = calloc(s->c.n[0] + 1, sizeof(double));
    s->gr_lapse = calloc(s->c.n[0], sizeof(double));
    s->gr_lapse_edge = calloc(s->c.n[0] + 1, sizeof(double));
  }
  s->t = 0.0;
  s->istep = 0;
  s->firstc = 1;
  s->dtold = 1.0e100;
  s->min_dt = 1.0e100;
  return s;
}

void orc_destroy(orc_sim *s)
{
  if (!s) return;
  free(s->p); free(s->ph); free(s->src); free(s->eos); free(s->shock); free(s->geom);
  for (int d = 0; d < 3; d++) free(s->flux[d]);
  for (int q = 0; q < 3; q++) {
    if (s->lmom[q][0] || s->lmom[q][1]) {
      free(s->mom[q][0]);
      free(s->lmom[q][0]);
      free(s->lmom[q][1]);
    } else {
      for (int d = 0; d < 3; d++) free(s->mom[q][d]);
    }
  }
  free(s->dj); free(s->dk); free(s->live);
  free(s->gtov); free(s->gr_mr); free(s->gr_lapse); free(s->gr_lapse_edge);
  free(s);
}

/* zone (i,j,k) of the level-local index space is a real zone */
static int zone_valid(const orc_sim *s, int i, int j, int k)
{
  return j < njs(s, i) && k < nks(s, i, j);
}

/* Dendritic grids keep the (i, j, k) storage of the uniform grid with j and
 * k level-local; positions beyond a level's nj / nk are not zones. */
void orc_set_prims(orc_sim *s, const double *prims)
{
  const int *n = s->c.n;
  long ncell = (long)n[0] * n[1] * n[2];
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        if (!zone_valid(s, i, j, k)) continue;
        long z = zid(s, i, j, k);
        long q = i + (long)n[0] * (j + (long)n[1] * k);
        for (int v = 0; v < s->nvars; v++) s->p[z * s->nvars + v] = prims[v * ncell + q];
      }
}

void orc_get_prims(const orc_sim *s, double *prims)
{
  const int *n = s->c.n;
  long ncell = (long)n[0] * n[1] * n[2];
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        long z = zid(s, i, j, k);
        long q = i + (long)n[0] * (j + (long)n[1] * k);
        if (!zone_valid(s, i, j, k)) {
          for (int v = 0; v < s->nvars; v++) prims[v * ncell + q] = 0.0;
          continue;
        }
        for (int v = 0; v < s->nvars; v++) prims[v * ncell + q] = s->p[z * s->nvars + v];
      }
}

int orc_zone_count_all(const orc_sim *s) { return (int)s->nall; }
const orc_zone_geom *orc_geometry(const orc_sim *s) { return s->geom; }
const double *orc_moment(const orc_sim *s, int which, int dir) { return s->mom[which][dir]; }
double orc_time(const orc_sim *s) { return s->t; }
double orc_last_dt(const orc_sim *s) { return s->dt; }
int orc_istep(const orc_sim *s) { return s->istep; }
long orc_shock_face_count(const orc_sim *s) { return s->nshock; }
long orc_hlle_face_count(const orc_sim *s) { return s->nhlle; }

/* ------------------------------------------------------------------ */
/* bc.c:705-1421 for uniform (non-dendritic) grids.  Direction d fills its own
 * ghost slabs over interior zones of the other directions only. */

/* one PROB ghost zone from the problem's hook: the time-dependent form sees
 * the global t as the reference's fill sets it (bc.c:717-721) */
static void prob_hook(const orc_sim *s, int i, int j, int k, double *dst)
{
  const orc_config *c = &s->c;
  if (c->prob_bounds_t) c->prob_bounds_t(i, j, k, s->t, dst, c->prob_user);
  else c->prob_bounds(i, j, k, dst, c->prob_user);
}

static void fill_ghost(orc_sim *s, double *P, int d, int side)
{
  const orc_config *c = &s->c;
  int nd = c->n[d];
  int code = side ? c->bc_hi[d] : c->bc_lo[d];
  int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
  int nv = s->nvars;
  for (int b = 0; b < c->n[o2]; b++)
    for (int a = 0; a < c->n[o1]; a++)
      for (int gidx = 0; gidx < ORC_NG; gidx++) {
        int ig = side ? nd + gidx : -ORC_NG + gidx;   /* ghost index along d */
        int src = 0, flip = 0, extrap = 0, pole = 0;
        int origin = 0, prob = 0, disk = 0;
        switch (code) {
          case ORC_BC_SPHERICAL_ORIGIN:
            /* theta boundary across the pole: mirror in j, phi + pi, flip v2 and v3
             * (bc.c:1094-1121, 1235-1262); inner radial boundary through the
             * origin: mirror in i, theta -> pi - theta, phi + pi, flip v1 and v3
             * (bc.c:774-801) */
            src = side ? 2 * nd - ig - 1 : -1 - ig;
            pole = (d == 1);
            origin = (d == 0 && !side);
            break;
          case ORC_BC_OUTFLOW:
            src = side ? nd - 1 : 0;
            break;
          case ORC_BC_REFLECT:
            if (d == 2) src = side ? nd - 1 : 0;     /* bc.c:1285-1300, 1360-1377: copy, not mirror */
            else src = side ? 2 * nd - ig - 1 : -1 - ig;
            flip = 1;
            break;
          case ORC_BC_PERIODIC:
            src = side ? ig - nd : ig + nd;
            break;
          case ORC_BC_EXTRAP:
            extrap = 1;
            break;
          case ORC_BC_PROB:
            prob = 1;   /* bc.c:125-140: prob_bounds(i, j, k, p) per ghost zone */
            break;
          case ORC_BC_DISK:
            /* bc.c:863-878, inner radius only: p(0) + i (p(1) - p(0)), then an
             * outward v1 set to 0 */
            if (d == 0 && !side) { extrap = 1; disk = 1; }
            else src = side ? nd - 1 : 0;
            break;
          default:
            src = side ? nd - 1 : 0;
        }
        int idx[3], sidx[3];
        idx[d] = ig; idx[o1] = a; idx[o2] = b;
        sidx[d] = src; sidx[o1] = a; sidx[o2] = b;
        if (pole) sidx[2] = (idx[2] + c->n[2] / 2) % c->n[2];
        if (origin) {
          sidx[1] = c->n[1] - 1 - idx[1];
          sidx[2] = (idx[2] + c->n[2] / 2) % c->n[2];
        }
        double *dst = &P[zid(s, idx[0], idx[1], idx[2]) * nv];
        if (prob) {
          prob_hook(s, idx[0], idx[1], idx[2], dst);
        } else if (extrap) {
          int i0, i1;
          double w;
          if (side) { i0 = nd - 1; i1 = nd - 2; w = (double)(ig - nd + 1); }
          else { i0 = 0; i1 = 1; w = (double)(ig - 0); }
          int a0[3], a1[3];
          a0[d] = i0; a0[o1] = a; a0[o2] = b;
          a1[d] = i1; a1[o1] = a; a1[o2] = b;
          const double *q0 = &P[zid(s, a0[0], a0[1], a0[2]) * nv];
          const double *q1 = &P[zid(s, a1[0], a1[1], a1[2]) * nv];
          for (int v = 0; v < nv; v++)
            dst[v] = side ? q0[v] + w * (q0[v] - q1[v]) : q0[v] + w * (q1[v] - q0[v]);
          if (disk && dst[ORC_U1] > 0.0) dst[ORC_U1] = 0.0;
        } else {
          const double *sp = &P[zid(s, sidx[0], sidx[1], sidx[2]) * nv];
          for (int v = 0; v < nv; v++) dst[v] = sp[v];
          if (flip) dst[ORC_U1 + d] *= -1.0;
          if (pole) { dst[ORC_U2] *= -1.0; dst[ORC_U3] *= -1.0; }
          if (origin) { dst[ORC_U1] *= -1.0; dst[ORC_U3] *= -1.0; }
        }
      }
}

static void reset_boundaries(orc_sim *s, double *P)
{
  for (int d = 0; d < s->c.ndim; d++) {
    fill_ghost(s, P, d, 0);
    fill_ghost(s, P, d, 1);
  }
}

/* eos_gamma.c:45-57, or the table (eos_stuff.c:507-650, walk started from
 * the temperature of the previous fill), over every zone (interior and
 * ghosts, eos.c:4-32) */
static void update_eos(orc_sim *s, const double *P)
{
  double gam = s->c.gam;
  for (long z = 0; z < s->nall; z++) {
    if (s->live && !s->live[z]) continue;
    const double *p = &P[z * s->nvars];
    double *e = &s->eos[z * NEOS];
    if (s->c.eos_tab) {
      double out[4];
      orc_eos_walk(s->c.eos_tab, p[ORC_RHO], p[ORC_UU], p[ORC_YE], e[E_TEMP], out);
      e[E_PRESS] = out[0];
      e[E_CS] = out[1];
      e[E_TEMP] = out[2];
      e[E_GAMMA] = out[3];
      continue;
    }
    double press = (gam - 1.0) * p[ORC_UU];
    e[E_PRESS] = press;
    e[E_CS] = sqrt((gam * press) / p[ORC_RHO]);
    e[E_GAMMA] = gam;
  }
}

void orc_update_eos_all(orc_sim *s) { update_eos(s, s->p); }

/* rad_fluid.c:195-214 then gravity.c:46-64 (interior zones) */
static void sources(orc_sim *s, const double *P)
{
  const orc_config *c = &s->c;
  for (int k = 0; k < c->n[2]; k++)
    for (int j = 0; j < c->n[1]; j++)
      for (int i = 0; i < c->n[0]; i++) {
        if (!zone_valid(s, i, j, k)) continue;
        long z = zid(s, i, j, k);
        const double *p = &P[z * s->nvars];
        double *src = &s->src[z * s->nvars];
        const orc_zone_geom *g = &s->geom[z];
        orc_geom_src(p, s->eos[z * NEOS + E_PRESS], src, g, s->nvars, c->geom);
        if (c->use_fixed_grav) {
          double vsq = 0.0;   /* geom_dot, geometry.c:65-73 */
          for (int d = 0; d < 3; d++) vsq += (p[ORC_U1 + d] * c->g0[d]) * g->gcov[0][d];
          src[ORC_ETOT] += p[ORC_RHO] * vsq;
          for (int d = 0; d < c->ndim; d++) src[ORC_U1 + d] += (p[ORC_RHO] * c->g0[d]) * g->gcov[0][d];
        }
      }
}

static void interp_pencil(orc_sim *s, double pen[][2 * ORC_NG], double pl[][2 * ORC_NG],
                          double pr[][2 * ORC_NG], const double *al, const double *be,
                          const double *ga, const double *x)
{
  for (int vv = 0; vv < s->ninterp; vv++) {
    switch (s->c.interp_order) {
      case 3: orc_para_recon(pen[vv], pl[vv], pr[vv], al, be, ga, x, -1, 0); break;
      case 2: orc_lin_recon(pen[vv], pl[vv], pr[vv], be, ga, x, -1, 0); break;
      default:
        for (int q = -1; q <= 0; q++) pl[vv][q + ORC_NG] = pr[vv][q + ORC_NG] = pen[vv][q + ORC_NG];
    }
  }
}

/* Polar-axis treatment of the phi faces in the last theta row
 * (step.c:1299-1397): first-order states, with the velocities of zones k-1
 * and k carried through Cartesian components (rotation at the zone centre)
 * into the frame of the face at phi_k.  Only the face's two states are
 * needed: pright of zone k-1 and pleft of zone k. */
static void polar_states(orc_sim *s, const double *P, int i, int j, int k, double pen[][2 * ORC_NG],
                         double pleft[][2 * ORC_NG], double pright[][2 * ORC_NG])
{
  const orc_config *c = &s->c;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  const int nv = s->nvars;
  const size_t lo = (size_t)lev_of(djs(s, i)) * (c->n[1] + 2 * ORC_NG) + j + ORC_NG;
  const double th = th_of_x(&m, s->lmom[1][0][lo] / s->lmom[2][0][lo]);
  const double sth = sin(th), cth = cos(th);
  const int dk = dks(s, i, j);
  for (int kt = k - 1; kt <= k; kt++) {
    long z = zid(s, i, j, kt);
    const orc_zone_geom *g = &s->geom[z];
    double ph = c->startx[2] + (kt + 0.5) * dk * c->dx[2];
    double sph = sin(ph), cph = cos(ph);
    double lam[3][3] = {{sth * cph, cth * cph, -sph}, {sth * sph, cth * sph, cph}, {cth, -sth, 0.0}};
    double v[3];
    for (int l = 0; l < 3; l++) {
      v[l] = 0.0;
      for (int mm = 0; mm < 3; mm++) v[l] += lam[l][mm] * P[z * nv + ORC_U1 + mm] * g->scale[0][mm];
    }
    /* the face at phi_k: the lower face of zone k, the upper face of zone k-1 */
    ph = c->startx[2] + k * dk * c->dx[2];
    sph = sin(ph);
    cph = cos(ph);
    double lf[3][3] = {{sth * cph, cth * cph, -sph}, {sth * sph, cth * sph, cph}, {cth, -sth, 0.0}};
    const double *scf = s->geom[zid(s, i, j, k)].scale[3];
    const int q = ORC_NG + kt - k;
    double (*dst)[2 * ORC_NG] = (kt == k) ? pleft : pright;
    for (int vv = 0; vv < s->ninterp; vv++) dst[vv][q] = pen[vv][q];
    for (int l = 0; l < 3; l++) {
      dst[ORC_U1 + l][q] = 0.0;
      for (int mm = 0; mm < 3; mm++) dst[ORC_U1 + l][q] += lf[mm][l] * v[mm] / scf[l];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Dendritic 3D spherical grids */

/* bc.c:705-1421 on a dendritic grid: radial REFLECT/OUTFLOW or through the
 * origin (inner radius, SPHERICAL_ORIGIN), theta across
 * the pole (SPHERICAL_ORIGIN), phi PERIODIC, each at the level of the ghost
 * zone's own row.  Filled in the order radial, theta, phi; the theta fill
 * also covers the radial ghost zones and the phi fill every row including
 * the ghost rows, so that the corner ghosts the prolongation stencils reach
 * hold the composition of the fills (the reference's loops leave corners
 * unset). */
static void fill_ghost_dend(orc_sim *s, double *P)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], NG = ORC_NG, nv = s->nvars;
  for (int side = 0; side < 2; side++) {
    const int code = side ? c->bc_hi[0] : c->bc_lo[0];
    for (int gi = 0; gi < NG; gi++) {
      const int i = side ? n1 + gi : -NG + gi;
      for (int j = 0; j < njs(s, i); j++)
        for (int k = 0; k < nks(s, i, j); k++) {
          int ip, jp = j, kp = k;
          const int origin = code == ORC_BC_SPHERICAL_ORIGIN && !side;
          double *dst = &P[zid(s, i, j, k) * nv];
          if (code == ORC_BC_PROB) {
            prob_hook(s, i, j, k, dst);
            continue;
          }
          if (code == ORC_BC_EXTRAP) {
            /* bc.c EXTRAP: the edge zone plus (distance) x (edge - next) at
             * the ghost's own (j, k), as the reference indexes it */
            const int i0 = side ? n1 - 1 : 0, i1 = side ? n1 - 2 : 1;
            const double w = side ? (double)(i - n1 + 1) : (double)i;
            const double *q0 = &P[zid(s, i0, j, k) * nv], *q1 = &P[zid(s, i1, j, k) * nv];
            for (int v = 0; v < nv; v++) dst[v] = side ? q0[v] + w * (q0[v] - q1[v]) : q0[v] + w * (q1[v] - q0[v]);
            continue;
          }
          if (code == ORC_BC_REFLECT) {
            ip = side ? 2 * n1 - 1 - i : -1 - i;
            jp = j * djs(s, i) / djs(s, ip);
            kp = k * dks(s, i, j) / dks(s, ip, jp);
          } else if (origin) {
            /* bc.c:774-801: i and j reflected across the origin, k + nk/2 */
            ip = -1 - i;
            jp = njs(s, ip) - 1 - j;
            kp = (k + nks(s, ip, jp) / 2) % nks(s, ip, jp);
          } else {
            ip = side ? n1 - 1 : 0;
          }
          const double *src = &P[zid(s, ip, jp, kp) * nv];
          for (int v = 0; v < nv; v++) dst[v] = src[v];
          if (code == ORC_BC_REFLECT) dst[ORC_U1] *= -1.0;
          if (origin) { dst[ORC_U1] *= -1.0; dst[ORC_U3] *= -1.0; }
        }
    }
  }
  for (int i = -NG; i < n1 + NG; i++) {
    const int nj = njs(s, i);
    for (int side = 0; side < 2; side++)
      for (int gj = 0; gj < NG; gj++) {
        const int j = side ? nj + gj : -NG + gj;
        const int jp = side ? 2 * nj - 1 - j : -1 - j;
        const int nkp = nks(s, i, jp);
        for (int k = 0; k < nks(s, i, j); k++) {
          const int kp = (k + nkp / 2) % nkp;
          double *dst = &P[zid(s, i, j, k) * nv];
          const double *src = &P[zid(s, i, jp, kp) * nv];
          for (int v = 0; v < nv; v++) dst[v] = src[v];
          dst[ORC_U2] *= -1.0;
          dst[ORC_U3] *= -1.0;
        }
      }
  }
  for (int i = -NG; i < n1 + NG; i++)
    for (int j = -NG; j < njs(s, i) + NG; j++) {
      const int nk = nks(s, i, j);
      for (int gk = 0; gk < NG; gk++) {
        const int klo = -NG + gk, khi = nk + gk;
        double *dl = &P[zid(s, i, j, klo) * nv], *dh = &P[zid(s, i, j, khi) * nv];
        const double *sl = &P[zid(s, i, j, klo + nk) * nv], *sh = &P[zid(s, i, j, khi - nk) * nv];
        for (int v = 0; v < nv; v++) {
          dl[v] = sl[v];
          dh[v] = sh[v];
        }
      }
    }
}

#define OSIGN(a) ((a) < 0.0 ? -1.0 : 1.0)
#define OMINMOD(a, b) (0.5 * (OSIGN(a) + OSIGN(b)) * OMIN(fabs(a), fabs(b)))   /* decs.h:629-630 */

static const double *lev_arr(const orc_sim *s, int which, int dir, int lev)
{
  const int n = dir == 0 ? s->c.n[1] : s->c.n[2];
  return s->lmom[which][dir] + (size_t)lev * (n + 2 * ORC_NG) + ORC_NG;
}

/* zone values of a pencil row: nvars primitives, then P and Gamma_1 */
static void zone_row(const orc_sim *s, const double *P, int i, int j, int k, double *pp)
{
  const long z = zid(s, i, j, k);
  for (int v = 0; v < s->nvars; v++) pp[v] = P[z * s->nvars + v];
  pp[s->nvars] = s->eos[z * NEOS + E_PRESS];
  pp[s->nvars + 1] = s->eos[z * NEOS + E_GAMMA];
}

/* kslopes0, step.c:783-840: limited k-slopes of row (ii,jj) evaluated at the
 * phi centroid of zone kp of row (ip,jp) */
static void kslopes0(const orc_sim *s, const double *P, int ip, int jp, int kp, int ii, int jj, double *pp)
{
  const int nv = s->nvars;
  const int lp = lev_of(dks(s, ip, jp)), l = lev_of(dks(s, ii, jj));
  const double *Bp = lev_arr(s, 1, 1, lp), *Gp = lev_arr(s, 2, 1, lp);
  const double *B = lev_arr(s, 1, 1, l), *G = lev_arr(s, 2, 1, l);
  const double x2p = Bp[kp] / Gp[kp];
  const int kk = kp * dks(s, ip, jp) / dks(s, ii, jj);
  const double x20 = B[kk - 1] / G[kk - 1], x21 = B[kk] / G[kk], x22 = B[kk + 1] / G[kk + 1];
  const double idx2m = 1.0 / (x21 - x20), idx2p = 1.0 / (x22 - x21), idx2c = 1.0 / (x22 - x20);
  double a0[ORC_NSIZE], a1[ORC_NSIZE], a2[ORC_NSIZE];
  zone_row(s, P, ii, jj, kk - 1, a0);
  zone_row(s, P, ii, jj, kk, a1);
  zone_row(s, P, ii, jj, kk + 1, a2);
  for (int v = 0; v < nv + 2; v++) {
    const double sm = (a1[v] - a0[v]) * idx2m, sp = (a2[v] - a1[v]) * idx2p, sc = (a2[v] - a0[v]) * idx2c;
    const double sg = OMINMOD(2.0 * OMINMOD(sm, sp), sc);
    pp[v] = a1[v] + sg * (x2p - x21);
  }
}

/* prolongate0, step.c:905-951: the value of radial level ii at the (theta,
 * phi) centroid of zone (i,j,k): k-slopes on rows jj-1..jj+1, then a limited
 * theta slope */
static void prolongate0(const orc_sim *s, const double *P, int i, int j, int k, int ii, double *pp)
{
  const int nv = s->nvars;
  const int jj = j * djs(s, i) / djs(s, ii);
  const int lp = lev_of(djs(s, i)), l = lev_of(djs(s, ii));
  const double *Bp = lev_arr(s, 1, 0, lp), *Gp = lev_arr(s, 2, 0, lp);
  const double *B = lev_arr(s, 1, 0, l), *G = lev_arr(s, 2, 0, l);
  const double x1p = Bp[j] / Gp[j];
  double p0[ORC_NSIZE], p1[ORC_NSIZE], p2[ORC_NSIZE];
  kslopes0(s, P, i, j, k, ii, jj - 1, p0);
  kslopes0(s, P, i, j, k, ii, jj, p1);
  kslopes0(s, P, i, j, k, ii, jj + 1, p2);
  const double x10 = B[jj - 1] / G[jj - 1], x11 = B[jj] / G[jj], x12 = B[jj + 1] / G[jj + 1];
  const double idx1m = 1.0 / (x11 - x10), idx1p = 1.0 / (x12 - x11), idx1c = 1.0 / (x12 - x10);
  for (int v = 0; v < nv + 2; v++) {
    const double sm = (p1[v] - p0[v]) * idx1m, sp = (p2[v] - p1[v]) * idx1p, sc = (p2[v] - p0[v]) * idx1c;
    const double sg = OMINMOD(2.0 * OMINMOD(sm, sp), sc);
    pp[v] = p1[v] + sg * (x1p - x11);
  }
}

/* restrict0, step.c:1012-1049: volume-weighted mean of the finer zones of
 * level ii under zone (i,j,k).  The last pencil entry averages the primitive
 * at the Gamma_1 index of the eos record (p[...][GAMMA], step.c:1048), i.e.
 * v^3, as the reference does. */
static void restrict0(const orc_sim *s, const double *P, int i, int j, int k, int ii, double *pp)
{
  const int nv = s->nvars;
  const int njp = djs(s, i) / djs(s, ii), jp = j * njp;
  double vol_sum = 0.0;
  for (int jj = jp; jj < jp + njp; jj++) {
    const int nkp = dks(s, i, j) / dks(s, ii, jj), kp = k * nkp;
    for (int kk = kp; kk < kp + nkp; kk++) vol_sum += s->geom[zid(s, ii, jj, kk)].volume;
  }
  vol_sum = 1.0 / vol_sum;
  for (int v = 0; v < nv + 2; v++) pp[v] = 0.0;
  for (int jj = jp; jj < jp + njp; jj++) {
    const int nkp = dks(s, i, j) / dks(s, ii, jj), kp = k * nkp;
    for (int kk = kp; kk < kp + nkp; kk++) {
      const long z = zid(s, ii, jj, kk);
      const double vol_frac = s->geom[z].volume * vol_sum;
      for (int v = 0; v < nv; v++) pp[v] += P[z * nv + v] * vol_frac;
      pp[nv] += s->eos[z * NEOS + E_PRESS] * vol_frac;
      pp[nv + 1] += P[z * nv + E_GAMMA] * vol_frac;
    }
  }
}

/* kslopes1, step.c:842-896: row jj evaluated at the phi centroid of the
 * theta face's zone kp, the face at the finer k-level of rows j-1 and j */
static void kslopes1(const orc_sim *s, const double *P, int i, int j, int kp, int jj, double *pp)
{
  const int nv = s->nvars;
  const int dkf = OMIN(dks(s, i, j), dks(s, i, j - 1));
  const int lp = lev_of(dkf), l = lev_of(dks(s, i, jj));
  const double *Bp = lev_arr(s, 1, 1, lp), *Gp = lev_arr(s, 2, 1, lp);
  const double *B = lev_arr(s, 1, 1, l), *G = lev_arr(s, 2, 1, l);
  const double x2p = Bp[kp] / Gp[kp];
  const int kk = kp * dkf / dks(s, i, jj);
  const double x20 = B[kk - 1] / G[kk - 1], x21 = B[kk] / G[kk], x22 = B[kk + 1] / G[kk + 1];
  const double idx2m = 1.0 / (x21 - x20), idx2p = 1.0 / (x22 - x21), idx2c = 1.0 / (x22 - x20);
  double a0[ORC_NSIZE], a1[ORC_NSIZE], a2[ORC_NSIZE];
  zone_row(s, P, i, jj, kk - 1, a0);
  zone_row(s, P, i, jj, kk, a1);
  zone_row(s, P, i, jj, kk + 1, a2);
  for (int v = 0; v < nv + 2; v++) {
    const double sm = (a1[v] - a0[v]) * idx2m, sp = (a2[v] - a1[v]) * idx2p, sc = (a2[v] - a0[v]) * idx2c;
    const double sg = OMINMOD(2.0 * OMINMOD(sm, sp), sc);
    pp[v] = a1[v] + sg * (x2p - x21);
  }
}

/* prolongate1 (step.c:962-1010) and restrict1 (:1051-1085) */
static void theta_row(const orc_sim *s, const double *P, int i, int j, int kp, int jj, double *pp)
{
  const int nv = s->nvars;
  const int dkf = OMIN(dks(s, i, j), dks(s, i, j - 1)), dkr = dks(s, i, jj);
  if (dkr < dkf) {
    const int nkp = dkf / dkr, k0 = kp * nkp;
    double vol_sum = 0.0;
    for (int kk = k0; kk < k0 + nkp; kk++) vol_sum += s->geom[zid(s, i, jj, kk)].volume;
    vol_sum = 1.0 / vol_sum;
    for (int v = 0; v < nv + 2; v++) pp[v] = 0.0;
    for (int kk = k0; kk < k0 + nkp; kk++) {
      const long z = zid(s, i, jj, kk);
      const double vol_frac = s->geom[z].volume * vol_sum;
      for (int v = 0; v < nv; v++) pp[v] += P[z * nv + v] * vol_frac;
      pp[nv] += s->eos[z * NEOS + E_PRESS] * vol_frac;
      pp[nv + 1] += s->eos[z * NEOS + E_GAMMA] * vol_frac;
    }
  } else if (dkr == dkf) {
    zone_row(s, P, i, jj, kp * dkf / dkr, pp);
  } else {
    kslopes1(s, P, i, j, kp, jj, pp);
  }
}

/* transverse_shock, fluid.c:189-210 */
static int transverse_shock_dend(const orc_sim *s, int i, int j, int k, int dir)
{
  if (dir == 1) {
    const int nkm = dks(s, i, j) / dks(s, i, j - 1), km = k * nkm;
    int orflag = (s->shock[zid(s, i, j, k)] == 0);
    for (int kk = km; kk < km + nkm; kk++) orflag = (orflag || s->shock[zid(s, i, j - 1, kk)] == 0);
    return orflag;
  }
  return s->shock[zid(s, i, j, k)] == 0 || s->shock[zid(s, i, j, k - 1)] == 0;
}

static void riemann_store(orc_sim *s, int d, int isT, double *pl, double *pr, long zg, long zf, double width,
                          int *isshock)
{
  double flux[ORC_NSIZE], vr[3], ss;
  const int nv = s->nvars;
  const orc_zone_geom *g = &s->geom[zg];
  *isshock = 0;
  if (isT) {
    orc_riemann_hlle(s->nhydro, nv, pl, pr, g, d, flux, &ss, vr);
    s->nhlle++;
  } else {
    *isshock = orc_riemann_hllc(s->nhydro, nv, pl, pr, g, d, flux, &ss, vr);
  }
  for (int vv = 0; vv < s->nhydro; vv++) s->flux[d][zf * nv + vv] = flux[vv];
  s->min_dt = OMIN(s->min_dt, width / fabs(ss));
}

/* prim_to_fluxes3 on a dendritic grid, step.c:1159-1199 (radial faces),
 * 1201-1265 (theta faces at the finer k-level of the two rows), 1266-1421
 * (phi faces, polar-axis states in the last theta row); on a 2D grid
 * (prim_to_fluxes2, step.c:659-764) the radial and theta faces only, with
 * one phi zone per row, where kslopes0 is flat and the pencils reduce to the
 * 2D prolongate0 / restrict0 (step.c:579-656) */
static void sweep_dend(orc_sim *s, const double *P)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], nv = s->nvars, NG = ORC_NG;
  double pen[ORC_NSIZE][2 * NG], pleft[ORC_NSIZE][2 * NG], pright[ORC_NSIZE][2 * NG];
  double al[2 * NG], be[2 * NG], ga[2 * NG], x[2 * NG], pp[ORC_NSIZE], pl[ORC_NSIZE], pr[ORC_NSIZE];
  for (int i = 0; i <= n1; i++)
    for (int j = 0; j < njs(s, i); j++)
      for (int k = 0; k < nks(s, i, j); k++) {
        for (int q = 0; q < 2 * NG; q++) {
          const int ii = i - NG + q;
          al[q] = s->mom[0][0][ii + NG];
          be[q] = s->mom[1][0][ii + NG];
          ga[q] = s->mom[2][0][ii + NG];
          x[q] = c->startx[0] + ii * c->dx[0];
          if (djs(s, ii) < djs(s, i)) restrict0(s, P, i, j, k, ii, pp);
          else prolongate0(s, P, i, j, k, ii, pp);
          for (int vv = 0; vv < nv + 2; vv++) pen[vv][q] = pp[vv];
        }
        interp_pencil(s, pen, pleft, pright, al, be, ga, x);
        for (int vv = 0; vv < nv + 2; vv++) {
          pl[vv] = pright[vv][NG - 1];
          pr[vv] = pleft[vv][NG];
        }
        int isshock;
        const long z = zid(s, i, j, k);
        riemann_store(s, 0, 0, pl, pr, z, z, c->dx[0], &isshock);
        if (isshock) {
          s->nshock++;
          s->shock[z] = 0;
          int jm = j * djs(s, i) / djs(s, i - 1), km = k * dks(s, i, j) / dks(s, i - 1, jm);
          s->shock[zid(s, i - 1, jm, km)] = 0;
          jm = j * djs(s, i) / djs(s, i - 2);
          km = k * dks(s, i, j) / dks(s, i - 2, jm);
          s->shock[zid(s, i - 2, jm, km)] = 0;
        }
      }
  for (int i = 0; i < n1; i++) {
    const int lj = lev_of(djs(s, i));
    const double *A = lev_arr(s, 0, 0, lj), *B = lev_arr(s, 1, 0, lj), *G = lev_arr(s, 2, 0, lj);
    for (int j = 0; j <= njs(s, i); j++)
      for (int k = 0; k < nks(s, i, j); k++) {
        const int dkj = dks(s, i, j), dkm = dks(s, i, j - 1), dkf = OMIN(dkj, dkm);
        const int nkp = OMAX(1, dkj / dkf), kp = k * OMAX(1, dkj / dkm);
        for (int kk = kp; kk < kp + nkp; kk++) {
          for (int q = 0; q < 2 * NG; q++) {
            const int jj = j - NG + q;
            al[q] = A[jj];
            be[q] = B[jj];
            ga[q] = G[jj];
            x[q] = c->startx[1] + jj * djs(s, i) * c->dx[1];
            theta_row(s, P, i, j, kk, jj, pp);
            for (int vv = 0; vv < nv + 2; vv++) pen[vv][q] = pp[vv];
          }
          interp_pencil(s, pen, pleft, pright, al, be, ga, x);
          for (int vv = 0; vv < nv + 2; vv++) {
            pl[vv] = pright[vv][NG - 1];
            pr[vv] = pleft[vv][NG];
          }
          int isshock;
          riemann_store(s, 1, transverse_shock_dend(s, i, j, k, 1), pl, pr, zid(s, i, j, k), zid(s, i, j, kk),
                        djs(s, i) * c->dx[1], &isshock);
        }
      }
  }
  for (int i = 0; i < n1 && c->ndim == 3; i++)
    for (int j = 0; j < njs(s, i); j++) {
      const int dk = dks(s, i, j), lk = lev_of(dk);
      const double *A = lev_arr(s, 0, 1, lk), *B = lev_arr(s, 1, 1, lk), *G = lev_arr(s, 2, 1, lk);
      for (int k = 0; k <= nks(s, i, j); k++) {
        for (int q = 0; q < 2 * NG; q++) {
          const int kt = k - NG + q;
          x[q] = c->startx[2] + kt * dk * c->dx[2];
          ga[q] = G[kt];
          be[q] = B[kt];
          al[q] = A[kt];
          zone_row(s, P, i, j, kt, pp);
          for (int vv = 0; vv < nv + 2; vv++) pen[vv][q] = pp[vv];
        }
        interp_pencil(s, pen, pleft, pright, al, be, ga, x);
        if (j == njs(s, i) - 1) polar_states(s, P, i, j, k, pen, pleft, pright);
        for (int vv = 0; vv < nv + 2; vv++) {
          pl[vv] = pright[vv][NG - 1];
          pr[vv] = pleft[vv][NG];
        }
        int isshock;
        const long z = zid(s, i, j, k);
        riemann_store(s, 2, transverse_shock_dend(s, i, j, k, 2), pl, pr, z, z, dk * c->dx[2], &isshock);
      }
    }
}

/* push on a dendritic grid (step.c:374-470): radial fan-in over the finer
 * outer zones, theta faces with the half-area and two-face rules, phi faces
 * (3D only: #if NDIM>2, step.c:435-438) */
static void push_dend(orc_sim *s, double dtpush, const double *p, const double *ph, double *pf, int stage)
{
  const orc_config *c = &s->c;
  const int nv = s->nvars;
  double pfact, phfact, ffact;
  if (stage == 0) { pfact = 1; phfact = 0; ffact = 1; }
  else { pfact = 0.5; phfact = 0.5; ffact = 0.5; }
  double min_dt_src = 1.0e100;
  const double *F0 = s->flux[0], *F1 = s->flux[1], *F2 = s->flux[2];
  for (int ii = 0; ii < c->n[0]; ii++)
    for (int jj = 0; jj < njs(s, ii); jj++)
      for (int kk = 0; kk < nks(s, ii, jj); kk++) {
        const long z = zid(s, ii, jj, kk);
        const orc_zone_geom *gm = &s->geom[z];
        double u[ORC_NSIZE], uh[ORC_NSIZE], divF[ORC_NSIZE];
        orc_p_to_u(&p[z * nv], u, gm->gcov[0], s->nhydro);
        orc_p_to_u(&ph[z * nv], uh, gm->gcov[0], s->nhydro);
        double local_dt = orc_dt_source_control(&p[z * nv], u, &s->src[z * nv]);
        min_dt_src = OMIN(min_dt_src, local_dt);
        for (int vv = 0; vv < nv; vv++) divF[vv] = 0.0;
        const int jp = jj * djs(s, ii) / djs(s, ii + 1);
        const int njp = (djs(s, ii + 1) < djs(s, ii)) ? djs(s, ii) / djs(s, ii + 1) : 1;
        for (int j = jp; j < jp + njp; j++) {
          const int nkp = dks(s, ii, jj) / dks(s, ii + 1, j), kp = kk * nkp;
          for (int k = kp; k < kp + nkp; k++) {
            const long zf = zid(s, ii + 1, j, k);
            for (int vv = 0; vv < nv; vv++) divF[vv] += F0[zf * nv + vv];
          }
        }
        for (int vv = 0; vv < nv; vv++) divF[vv] -= F0[z * nv + vv];
        const int dk0 = dks(s, ii, jj), dkp = dks(s, ii, jj + 1), dkm = dks(s, ii, jj - 1);
        if (dkp > dk0) {
          const long zf = zid(s, ii, jj + 1, kk);
          for (int vv = 0; vv < nv; vv++) divF[vv] += 0.5 * F1[zf * nv + vv];
        } else if (dkp < dk0) {
          const long za = zid(s, ii, jj + 1, 2 * kk), zb = zid(s, ii, jj + 1, 2 * kk + 1);
          for (int vv = 0; vv < nv; vv++) divF[vv] += F1[za * nv + vv] + F1[zb * nv + vv];
        } else {
          const long zf = zid(s, ii, jj + 1, kk);
          for (int vv = 0; vv < nv; vv++) divF[vv] += F1[zf * nv + vv];
        }
        if (dkm < dk0) {
          const long za = zid(s, ii, jj, 2 * kk), zb = zid(s, ii, jj, 2 * kk + 1);
          for (int vv = 0; vv < nv; vv++) divF[vv] -= 0.5 * F1[za * nv + vv] + 0.5 * F1[zb * nv + vv];
        } else {
          for (int vv = 0; vv < nv; vv++) divF[vv] -= F1[z * nv + vv];
        }
        if (c->ndim == 3) {
          const long z2 = zid(s, ii, jj, kk + 1);
          for (int vv = 0; vv < nv; vv++) divF[vv] += F2[z2 * nv + vv] - F2[z * nv + vv];
        }
        for (int vv = 0; vv < nv; vv++) {
          double ddt = (divF[vv] / gm->volume) - s->src[z * nv + vv];
          u[vv] = ((pfact * u[vv]) + (phfact * uh[vv])) - ((ffact * dtpush) * ddt);
        }
        orc_u_to_p(u, &pf[z * nv], gm->gcon, s->nhydro);
      }
  s->min_dt = OMIN(s->min_dt, min_dt_src);
}

/* One direction of the flux sweep: step.c:500-562 (d=0), 736-760 (d=1),
 * 1266-1421 (d=2).  Face (i,j,k) is the lower face of zone (i,j,k). */
static void sweep_dir(orc_sim *s, const double *P, int d)
{
  const orc_config *c = &s->c;
  int nv = s->nvars;
  int lo[3] = {0, 0, 0}, hi[3] = {c->n[0], c->n[1], c->n[2]};
  hi[d] += 1;   /* faces 0..n[d] */
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        int f[3] = {i, j, k};
        double pen[ORC_NSIZE][2 * ORC_NG], pleft[ORC_NSIZE][2 * ORC_NG], pright[ORC_NSIZE][2 * ORC_NG];
        double al[2 * ORC_NG], be[2 * ORC_NG], ga[2 * ORC_NG], x[2 * ORC_NG];
        for (int q = 0; q < 2 * ORC_NG; q++) {
          int zz[3] = {i, j, k};
          zz[d] = f[d] - ORC_NG + q;
          al[q] = s->mom[0][d][zz[d] + ORC_NG];
          be[q] = s->mom[1][d][zz[d] + ORC_NG];
          ga[q] = s->mom[2][d][zz[d] + ORC_NG];
          x[q] = c->startx[d] + zz[d] * c->dx[d];
          long z = zid(s, zz[0], zz[1], zz[2]);
          for (int vv = 0; vv < nv; vv++) pen[vv][q] = P[z * nv + vv];
          pen[nv][q] = s->eos[z * NEOS + E_PRESS];
          pen[nv + 1][q] = s->eos[z * NEOS + E_GAMMA];
        }
        interp_pencil(s, pen, pleft, pright, al, be, ga, x);
        if (d == 2 && c->geom == ORC_GEOM_SPHERICAL && j == c->n[1] - 1)
          polar_states(s, P, i, j, k, pen, pleft, pright);
        double pl[ORC_NSIZE], pr[ORC_NSIZE], flux[ORC_NSIZE], vr[3], ss;
        for (int vv = 0; vv < s->ninterp; vv++) {
          pl[vv] = pright[vv][ORC_NG - 1];
          pr[vv] = pleft[vv][ORC_NG];
        }
        long zf = zid(s, i, j, k);
        const orc_zone_geom *g = &s->geom[zf];
        if (d == 0) {
          int isshock = orc_riemann_hllc(s->nhydro, nv, pl, pr, g, 0, flux, &ss, vr);
          if (isshock) {
            s->nshock++;
            for (int q = 0; q < 3; q++) s->shock[zid(s, i - q, j, k)] = 0;
          }
        } else {
          int zlow[3] = {i, j, k};
          zlow[d] -= 1;
          int tshock = (s->shock[zf] == 0) || (s->shock[zid(s, zlow[0], zlow[1], zlow[2])] == 0);
          if (tshock) {
            orc_riemann_hlle(s->nhydro, nv, pl, pr, g, d, flux, &ss, vr);
            s->nhlle++;
          } else {
            orc_riemann_hllc(s->nhydro, nv, pl, pr, g, d, flux, &ss, vr);
          }
        }
        for (int vv = 0; vv < s->nhydro; vv++) s->flux[d][zf * nv + vv] = flux[vv];
        s->min_dt = OMIN(s->min_dt, c->dx[d] / fabs(ss));
      }
}

/* step.c:327-478 on uniform grids (DJS=DKS=1). */
static void push(orc_sim *s, double dtpush, const double *p, const double *ph, double *pf, int stage)
{
  const orc_config *c = &s->c;
  int nv = s->nvars;
  double pfact, phfact, ffact;
  if (stage == 0) { pfact = 1; phfact = 0; ffact = 1; }
  else { pfact = 0.5; phfact = 0.5; ffact = 0.5; }
  double min_dt_src = 1.0e100;
  for (int k = 0; k < c->n[2]; k++)
    for (int j = 0; j < c->n[1]; j++)
      for (int i = 0; i < c->n[0]; i++) {
        long z = zid(s, i, j, k);
        const orc_zone_geom *gm = &s->geom[z];
        double u[ORC_NSIZE], uh[ORC_NSIZE], divF[ORC_NSIZE];
        orc_p_to_u(&p[z * nv], u, gm->gcov[0], s->nhydro);
        orc_p_to_u(&ph[z * nv], uh, gm->gcov[0], s->nhydro);
        double local_dt = orc_dt_source_control(&p[z * nv], u, &s->src[z * nv]);
        min_dt_src = OMIN(min_dt_src, local_dt);
        for (int vv = 0; vv < nv; vv++) divF[vv] = 0.0;
        long zp0 = zid(s, i + 1, j, k);
        for (int vv = 0; vv < nv; vv++) divF[vv] += s->flux[0][zp0 * nv + vv];
        for (int vv = 0; vv < nv; vv++) divF[vv] -= s->flux[0][z * nv + vv];
        if (c->ndim > 1) {
          long zp1 = zid(s, i, j + 1, k);
          for (int vv = 0; vv < nv; vv++) divF[vv] += s->flux[1][zp1 * nv + vv];
          for (int vv = 0; vv < nv; vv++) divF[vv] -= s->flux[1][z * nv + vv];
        }
        if (c->ndim > 2) {
          long zp2 = zid(s, i, j, k + 1);
          for (int vv = 0; vv < nv; vv++) divF[vv] += s->flux[2][zp2 * nv + vv] - s->flux[2][z * nv + vv];
        }
        for (int vv = 0; vv < nv; vv++) {
          double ddt = (divF[vv] / gm->volume) - s->src[z * nv + vv];
          u[vv] = ((pfact * u[vv]) + (phfact * uh[vv])) - ((ffact * dtpush) * ddt);
        }
        orc_u_to_p(u, &pf[z * nv], gm->gcon, s->nhydro);
      }
  s->min_dt = OMIN(s->min_dt, min_dt_src);
}

/* Newtonian monopole gravity on radial shells (gravity.c:145-251 start and
 * 254-498 finish without GR_MONOPOLE, 1D: one zone per shell): shell masses
 * from the stage state, enclosed mass at the faces, g = -G M / r^2 dr/dx at
 * the faces, linear in x across the zone evaluated at the volume centroid
 * beta/Gamma; momentum source rho g, energy source (linear mass flux per
 * area between the faces at the centroid) * g.  Needs this stage's fluxes. */
/* face values of g (gravity.c:420-425): grav[0] = 0, grav[ii] = -G mr[ii] / r^2 dr/dx */
static void monopole_edges(const orc_sim *s, const double *P, double *grav)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], nv = s->nvars;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  double mr = 0.0;
  grav[0] = 0.0;
  for (int ii = 1; ii <= n1; ii++) {
    double ms = 0.0;
    for (int k = 0; k < c->n[2]; k++)
      for (int j = 0; j < c->n[1]; j++) {
        if (!zone_valid(s, ii - 1, j, k)) continue;
        long z = zid(s, ii - 1, j, k);
        ms += s->geom[z].volume * P[z * nv + ORC_RHO];
      }
    mr = mr + ms;
    double xe = c->startx[0] + ii * c->dx[0];
    grav[ii] = -GNEWT * mr / OSQR(r_of_x(&m, xe)) * dr_dx(&m, xe);
  }
}

/* GR effective monopole (gravity.c:142-251 start and 266-375 finish with
 * GR_MONOPOLE, no radiation, so the radiation pressure, energy and v.F
 * shell terms are zero): volume averages of rho, P, u and v^1 over each
 * shell (divided by the exact shell volume 4 pi/3 d(r^3)); the TOV mass and
 * factor Gamma = sqrt(1 + v^2/c^2 - 2 G m / (r c^2)) at the faces,
 * iterated until the outer mass changes by less than 1e-10 (at most 20
 * passes; Gamma starts from the first call's Newtonian estimate and is kept
 * between calls); the face gravity
 *   g = -G (m + 4 pi r^3 P / c^2) / (r Gamma)^2 (rho + (u + P)/c^2) / rho dr/dx
 * with face averages of neighbouring shells (the outer face uses the last
 * shell); and the lapses exp(phi / c^2) of the potential integrated inwards
 * from -G M / r at the outer face.  Returns 0, or 1 when 20 passes do not
 * converge (the reference exits there). */
#define CLIGHT 2.99792458e10    /* constants.h:10 */
static int monopole_edges_gr(orc_sim *s, const double *P, double *grav)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], nv = s->nvars;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  double *rhoavg = calloc(n1, sizeof(double)), *pgasavg = calloc(n1, sizeof(double));
  double *pradavg = calloc(n1, sizeof(double)), *uavg = calloc(n1, sizeof(double));
  double *vavg = calloc(n1, sizeof(double)), *eradavg = calloc(n1, sizeof(double));
  double *vdF = calloc(n1, sizeof(double)), *redge = calloc(n1 + 1, sizeof(double));
  double *gtov = s->gtov, *mr = s->gr_mr;
  for (int ii = 0; ii <= n1; ii++) redge[ii] = r_of_x(&m, fabs(ii * c->dx[0] + c->startx[0]));
  for (int ii = 0; ii < n1; ii++)
    for (int k = 0; k < c->n[2]; k++)
      for (int j = 0; j < c->n[1]; j++) {
        if (!zone_valid(s, ii, j, k)) continue;
        const long z = zid(s, ii, j, k);
        const double vol = s->geom[z].volume;
        rhoavg[ii] += vol * P[z * nv + ORC_RHO];
        pgasavg[ii] += vol * s->eos[z * NEOS + E_PRESS];
        pradavg[ii] += vol * 0.0;
        uavg[ii] += vol * P[z * nv + ORC_UU];
        vavg[ii] += vol * P[z * nv + ORC_U1];
        eradavg[ii] += vol * 0.0;
        vdF[ii] += vol * 0.0;
      }
  for (int ii = 0; ii < n1; ii++) {
    const double Vshell = 1.0 / (((4.0 * M_PI) / 3.0) * (pow(redge[ii + 1], 3) - pow(redge[ii], 3)));
    rhoavg[ii] *= Vshell;
    pgasavg[ii] *= Vshell;
    pradavg[ii] *= Vshell;
    uavg[ii] *= Vshell;
    vavg[ii] *= Vshell;
    eradavg[ii] *= Vshell;
    vdF[ii] *= Vshell;
  }
  const double c2 = OSQR(CLIGHT);
  mr[0] = 0.0;
  grav[0] = 0.0;
  double mstar = 0.0, mlast = 0.0;
  if (!s->gtov_set) {
    gtov[0] = 1.0;
    for (int ii = 1; ii <= n1; ii++) {
      const double Vshell = ((4.0 * M_PI) / 3.0) * (pow(redge[ii], 3) - pow(redge[ii - 1], 3));
      mstar += Vshell * rhoavg[ii - 1];
      gtov[ii] = sqrt(1.0 - ((2.0 * GNEWT) * mstar) / (c2 * redge[ii]));
    }
    s->gtov_set = 1;
  }
  int iter = 0, fail = 0;
  const double *be = s->mom[1][0] + ORC_NG, *ga = s->mom[2][0] + ORC_NG;
  while (1) {
    mstar = 0.0;
    for (int ii = 1; ii <= n1; ii++) {
      const double xl = c->startx[0] + ii * c->dx[0];
      double A = (gtov[ii] - gtov[ii - 1]) / c->dx[0];
      double B = gtov[ii - 1] - A * xl;
      double gtov_avg = (A * be[ii]) / ga[ii] + B;
      const double Vshell = ((4.0 * M_PI) / 3.0) * (pow(redge[ii], 3) - pow(redge[ii - 1], 3));
      const double vedge = (ii < n1) ? 0.5 * (vavg[ii - 1] + vavg[ii]) : vavg[ii - 1];
      const double w = (rhoavg[ii - 1] + uavg[ii - 1] / c2) + eradavg[ii - 1];
      gtov[ii] = sqrt((1.0 + pow(vedge / CLIGHT, 2)) -
                      ((2.0 * GNEWT) * (mr[ii - 1] + Vshell * ((w * gtov_avg) + vdF[ii - 1]))) / (redge[ii] * c2));
      A = (gtov[ii] - gtov[ii - 1]) / c->dx[0];
      B = gtov[ii - 1] - A * xl;
      gtov_avg = (A * be[ii]) / ga[ii] + B;
      mr[ii] = mr[ii - 1] + Vshell * ((w * gtov_avg) + vdF[ii - 1]);
      mstar += Vshell * rhoavg[ii - 1];
    }
    if (iter > 0 && fabs(mlast - mr[n1]) / mr[n1] < 1.0e-10) break;
    mlast = mr[n1];
    iter++;
    if (iter > 20) { fail = 1; break; }
  }
  s->gr_iter = iter;
  s->gr_fail = fail;
  for (int ii = 1; ii < n1; ii++) {
    const double pgasedge = 0.5 * (pgasavg[ii - 1] + pgasavg[ii]);
    const double pradedge = 0.5 * (pradavg[ii - 1] + pradavg[ii]);
    const double rhoedge = 0.5 * (rhoavg[ii - 1] + rhoavg[ii]);
    const double uedge = 0.5 * (uavg[ii - 1] + uavg[ii]);
    grav[ii] = ((((-GNEWT * (mr[ii] + (((4.0 * M_PI) * pow(redge[ii], 3)) * (pgasedge + pradedge)) / c2)) /
                  pow(redge[ii] * gtov[ii], 2)) * (rhoedge + (uedge + pgasedge) / c2)) / rhoedge) *
               dr_dx(&m, c->startx[0] + ii * c->dx[0]);
  }
  grav[n1] = ((((-GNEWT * (mr[n1] + (((4.0 * M_PI) * pow(redge[n1], 3)) * (pgasavg[n1 - 1] + pradavg[n1 - 1])) / c2)) /
                pow(redge[n1] * gtov[n1], 2)) * (rhoavg[n1 - 1] + (uavg[n1 - 1] + pgasavg[n1 - 1]) / c2)) /
              rhoavg[n1 - 1]) * dr_dx(&m, c->startx[0] + n1 * c->dx[0]);
  /* the potential, integrated inwards from the outer face, and the lapses */
  double *le = s->gr_lapse_edge, *lc = s->gr_lapse;
  le[n1] = (-GNEWT * mstar) / redge[n1];
  for (int ii = n1 - 1; ii >= 0; ii--) le[ii] = le[ii + 1] + (0.5 * (grav[ii] + grav[ii + 1])) * c->dx[0];
  for (int ii = 0; ii < n1; ii++) {
    const double xl = c->startx[0] + ii * c->dx[0];
    const double A = (le[ii + 1] - le[ii]) / c->dx[0];
    const double B = le[ii] - A * xl;
    lc[ii] = (A * be[ii]) / ga[ii] + B;
  }
  for (int ii = 0; ii < n1; ii++) {
    lc[ii] = exp(lc[ii] / c2);
    le[ii] = exp(le[ii] / c2);
  }
  le[n1] = exp(le[n1] / c2);
  free(rhoavg); free(pgasavg); free(pradavg); free(uavg); free(vavg); free(eradavg); free(vdF); free(redge);
  return fail;
}

void orc_monopole_edges(const orc_sim *s, double *grav) { monopole_edges(s, s->p, grav); }

static void update_eos(orc_sim *s, const double *P);
int orc_gr_edges(orc_sim *s, double *grav)
{
  if (!s->gtov) return -1;
  update_eos(s, s->p);
  return monopole_edges_gr(s, s->p, grav);
}

int orc_gr_state(const orc_sim *s, double *gtov, double *mr, double *lapse, double *lapse_edge, int *iter)
{
  if (!s->gtov) return -1;
  const int n1 = s->c.n[0];
  for (int ii = 0; ii <= n1; ii++) {
    if (gtov) gtov[ii] = s->gtov[ii];
    if (mr) mr[ii] = s->gr_mr[ii];
    if (lapse_edge) lapse_edge[ii] = s->gr_lapse_edge[ii];
    if (lapse && ii < n1) lapse[ii] = s->gr_lapse[ii];
  }
  if (iter) *iter = s->gr_iter;
  return s->gr_fail;
}

static void monopole(orc_sim *s, const double *P)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], nv = s->nvars;
  double *grav = calloc(n1 + 1, sizeof(double));
  if (c->use_monopole == 2) monopole_edges_gr(s, P, grav);
  else monopole_edges(s, P, grav);
  for (int ii = 0; ii < n1; ii++) {
    double xl = c->startx[0] + ii * c->dx[0];
    double A = (grav[ii + 1] - grav[ii]) / c->dx[0];
    double B = grav[ii] - A * xl;
    const double be = s->mom[1][0][ii + ORC_NG], ga = s->mom[2][0][ii + ORC_NG];
    double gz = (A * be) / ga + B;
    for (int k = 0; k < c->n[2]; k++)
      for (int j = 0; j < c->n[1]; j++) {
        if (!zone_valid(s, ii, j, k)) continue;
        long z = zid(s, ii, j, k);
        double vr = 0.0, area = 0.0;
        const int njp = djs(s, ii) / djs(s, ii + 1), jp = j * njp;
        for (int jj = jp; jj < jp + njp; jj++) {
          const int nkp = dks(s, ii, j) / dks(s, ii + 1, jj), kp = k * nkp;
          for (int kk = kp; kk < kp + nkp; kk++) {
            const long zp = zid(s, ii + 1, jj, kk);
            vr += s->flux[0][zp * nv + ORC_RHO];
            area += s->geom[zp].area[0] + 1.0e-16;
          }
        }
        vr /= area;
        s->src[z * nv + ORC_U1] += P[z * nv + ORC_RHO] * gz;
        double vl = s->flux[0][z * nv + ORC_RHO] / (s->geom[z].area[0] + 1.0e-16);
        double Av = (vr - vl) / c->dx[0];
        double Bv = vl - Av * xl;
        s->src[z * nv + ORC_ETOT] += ((Av * be) / ga + Bv) * gz;
      }
  }
  free(grav);
}

/* ------------------------------------------------------------------ */
/* Multipole gravity (gravity.c:499-1110, SPHERICAL_MULTIPOLE_GRAV): the
 * l = 1..lmax terms of the potential, added after the monopole
 * (gravity.c:1137-1140, 1162-1165).  Index of (l, m): l (l + 1) + m. */
#define ORC_LM(l, m) ((l) * ((l) + 1) + (m))
#define ORC_MP_LMAXC 15                 /* MULTIPOLE_LMAX, decs.h:38 */
#define ORC_SQRT2 1.41421356237309504880

/* n! / m! (gravity.c:499-516) */
static double mp_factorial_ratio(int n, int m)
{
  double fact = 1.0;
  while (n > m) { fact *= n; n--; }
  while (m > n) { fact /= m; m--; }
  return fact;
}

/* associated Legendre P[l,m](cos th), m >= 0, with the Condon-Shortley
 * phase: P[m,m] from (2m-1)!! (1-x^2)^(m/2), P[m+1,m], then the three-term
 * recurrence in l (gravity.c:579-607) */
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
    plm[ORC_LM(m, m)] = pmm;
    if (m != lmax) {
      plm[ORC_LM(m + 1, m)] = (cth * ((2.0 * m) + 1.0)) * pmm;
      for (int l = m + 2; l <= lmax; l++)
        plm[ORC_LM(l, m)] = (((cth * ((2.0 * l) - 1.0)) * plm[ORC_LM(l - 1, m)]) - ((l + m - 1) * plm[ORC_LM(l - 2, m)])) /
                            (l - m);
    }
  }
}

/* the real harmonics: sqrt((2l+1)/(4pi) (l-|m|)!/(l+|m|)!) P[l,|m|], times
 * sqrt(2) cos(m phi) for m > 0 and sqrt(2) sin(|m| phi) for m < 0
 * (gravity.c:518-577) */
void orc_ylms(double *ylm, double th, double phi, int lmax, int mmax)
{
  mp_plms(ylm, th, lmax, mmax);
  for (int m = 0; m <= mmax; m++) {
    const double smphi = sin(m * phi), cmphi = cos(m * phi);
    const double norm_fact = (m == 0) ? 1 : ORC_SQRT2;
    for (int l = m; l <= lmax; l++) {
      const double norm = norm_fact * sqrt((((2.0 * l) + 1.0) / (4.0 * M_PI)) * mp_factorial_ratio(l - m, l + m));
      ylm[ORC_LM(l, m)] *= norm;
      if (m > 0) {
        ylm[ORC_LM(l, -m)] = ylm[ORC_LM(l, m)] * smphi;
        ylm[ORC_LM(l, m)] *= cmphi;
      }
    }
  }
}

/* the integral of Y[l,m] sin(th) over a zone (gravity.c:609-697): the theta
 * integrals of P[l,m] by their recurrences, the phi integrals of cos / sin */
void orc_ylm_ints(double *slm, double thmin, double thmax, double phimin, double phimax, int lmax, int mmax)
{
  double plmp[ORC_MP_LMAXC * ORC_MP_LMAXC], plmm[ORC_MP_LMAXC * ORC_MP_LMAXC];
  const double cmin = cos(thmin), cmax = cos(thmax);
  mp_plms(plmm, thmin, lmax, mmax);
  mp_plms(plmp, thmax, lmax, mmax);
  for (int m = 0; m <= mmax; m++) {
    if (m > 1) {
      slm[ORC_LM(m, m)] = (((m * (2 * m - 1) * (2 * m - 3)) * slm[ORC_LM(m - 2, m - 2)] - (cmax * plmp[ORC_LM(m, m)])) +
                           (cmin * plmm[ORC_LM(m, m)])) / (m + 1.0);
    } else if (m == 0) {
      slm[ORC_LM(0, 0)] = cmin - cmax;
    } else {
      slm[ORC_LM(1, 1)] = 0.5 * ((((cmax * sqrt(1 - (cmax * cmax))) - thmax) - (cmin * sqrt(1 - (cmin * cmin)))) + thmin);
    }
    if (m != lmax) {
      slm[ORC_LM(m + 1, m)] = ((((1 - (cmax * cmax)) * plmp[ORC_LM(m, m)]) - ((1 - (cmin * cmin)) * plmm[ORC_LM(m, m)])) *
                               (2 * m + 1)) / (m + 2.0);
      for (int l = m + 2; l <= lmax; l++)
        slm[ORC_LM(l, m)] = ((((l - 2) * (l + m - 1)) * slm[ORC_LM(l - 2, m)]) +
                             ((2 * l - 1) * (((1 - (cmax * cmax)) * plmp[ORC_LM(l - 1, m)]) -
                                             ((1 - (cmin * cmin)) * plmm[ORC_LM(l - 1, m)])))) /
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
    const double norm_fact = (m == 0) ? 1.0 : ORC_SQRT2;
    for (int l = m; l <= lmax; l++) {
      const double norm = norm_fact * sqrt((((2.0 * l) + 1.0) / (4.0 * M_PI)) * mp_factorial_ratio(l - m, l + m));
      slm[ORC_LM(l, m)] *= norm;
      if (m > 0) {
        slm[ORC_LM(l, -m)] = slm[ORC_LM(l, m)] * cmphi;
        slm[ORC_LM(l, m)] *= smphi;
      } else {
        slm[ORC_LM(l, 0)] *= smphi;
      }
    }
  }
}

/* corner (ii, jj, kk) of the potential array */
static long mp_corner(const orc_sim *s, int ii, int jj, int kk)
{
  return ((long)ii * (s->c.n[1] + 1) + jj) * (long)(s->c.n[2] + 1) + kk;
}

/* The moments C[l,m](r) (mass inside, r^(l+3) weights) and D[l,m](r) (mass
 * outside, r^(2-l)) of the state, summed over shells in the reference's
 * order (gravity.c:720-793 start, 822-839 cumulative sums), then the
 * potential at every zone corner (gravity.c:848-906).  phi: the corner
 * array (mp_corner). */
static void multipole_phi(const orc_sim *s, const double *P, double *phi)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], nv = s->nvars;
  const int lmax = c->multipole_lmax, mmax = c->ndim == 3 ? lmax : 0;
  const int L1 = lmax + 1, nlm = L1 * L1;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  double *Clm = calloc((size_t)n1 * nlm, sizeof(double)), *Dlm = calloc((size_t)n1 * nlm, sizeof(double));
  double ylm[ORC_MP_LMAXC * ORC_MP_LMAXC];
  for (int kk = 0; kk < c->n[2]; kk++)
    for (int jj = 0; jj < c->n[1]; jj++)
      for (int ii = 0; ii < n1; ii++) {
        if (!zone_valid(s, ii, jj, kk)) continue;
        const double rmin = r_of_x(&m, c->startx[0] + ii * c->dx[0]);
        const double rmax = r_of_x(&m, c->startx[0] + (ii + 1) * c->dx[0]);
        const double thmin = th_of_x(&m, c->startx[1] + (jj * djs(s, ii)) * c->dx[1]);
        const double thmax = th_of_x(&m, c->startx[1] + ((jj + 1) * djs(s, ii)) * c->dx[1]);
        double phimin = 0.0, phimax = 2.0 * M_PI;
        if (c->ndim == 3) {
          phimin = c->startx[2] + (kk * dks(s, ii, jj)) * c->dx[2];
          phimax = c->startx[2] + ((kk + 1) * dks(s, ii, jj)) * c->dx[2];
        }
        orc_ylm_ints(ylm, thmin, thmax, phimin, phimax, lmax, mmax);
        const double rho = P[zid(s, ii, jj, kk) * nv + ORC_RHO];
        for (int l = 0; l <= lmax; l++) {
          const double rc = (pow(rmax, l + 3) * (1.0 - pow(rmin / rmax, l + 3))) / (l + 3);
          double rd;
          if (l == 0) rd = 0.5 * (OSQR(rmax) - OSQR(rmin));
          else if (l == 1) rd = rmax - rmin;
          else if (l == 2) rd = (rmin > 1.0e-16) ? log(rmax / rmin) : 0.0;
          else rd = (rmin > 1.0e-16) ? (pow(rmin, 2 - l) - pow(rmax, 2 - l)) / (l - 2) : 0.0;
          const int mm = l < mmax ? l : mmax;
          for (int q = -mm; q <= mm; q++) {
            Clm[(size_t)ii * nlm + ORC_LM(l, q)] += (rho * ylm[ORC_LM(l, q)]) * rc;
            Dlm[(size_t)ii * nlm + ORC_LM(l, q)] += (rho * ylm[ORC_LM(l, q)]) * rd;
          }
        }
      }
  for (int ii = 1; ii < n1; ii++)
    for (int l = 0; l <= lmax; l++) {
      const int mm = l < mmax ? l : mmax;
      for (int q = -mm; q <= mm; q++) Clm[(size_t)ii * nlm + ORC_LM(l, q)] += Clm[(size_t)(ii - 1) * nlm + ORC_LM(l, q)];
    }
  for (int ii = n1 - 2; ii >= 0; ii--)
    for (int l = 0; l <= lmax; l++) {
      const int mm = l < mmax ? l : mmax;
      for (int q = -mm; q <= mm; q++) Dlm[(size_t)ii * nlm + ORC_LM(l, q)] += Dlm[(size_t)(ii + 1) * nlm + ORC_LM(l, q)];
    }
  for (int ii = 0; ii <= n1; ii++) {
    const int njc = c->ndim > 1 ? njs(s, ii) : 0;
    for (int jj = 0; jj <= njc; jj++) {
      const int nkc = c->ndim == 3 ? nks(s, ii, jj) : 0;
      for (int kk = 0; kk <= nkc; kk++) {
        const double rm = r_of_x(&m, c->startx[0] + ii * c->dx[0]);
        const double thm = th_of_x(&m, c->startx[1] + (jj * djs(s, ii)) * c->dx[1]);
        const double phm = c->startx[2] + (kk * dks(s, ii, jj)) * c->dx[2];
        orc_ylms(ylm, thm, phm, lmax, mmax);
        double v = 0.0;
        if (rm < 1.0e-16) {
        } else if (ii == n1) {
          for (int l = 1; l <= lmax; l++) {
            const double Clfac = (((-4.0 * M_PI) * GNEWT) / (2 * l + 1)) / pow(rm, l + 1);
            const int mm = l < mmax ? l : mmax;
            for (int q = -mm; q <= mm; q++) v += (Clfac * ylm[ORC_LM(l, q)]) * Clm[(size_t)(ii - 1) * nlm + ORC_LM(l, q)];
          }
        } else {
          for (int l = 1; l <= lmax; l++) {
            const double Clfac = (((-4.0 * M_PI) * GNEWT) / (2 * l + 1)) / pow(rm, l + 1);
            const double Dlfac = (((-4.0 * M_PI) * GNEWT) / (2 * l + 1)) * pow(rm, l);
            const int mm = l < mmax ? l : mmax;
            for (int q = -mm; q <= mm; q++) {
              /* the inner radius of a grid that does not reach the origin:
               * no mass inside it (the reference reads Clm[ii - 1] before
               * the array there, gravity.c:897) */
              if (ii > 0) v += (Clfac * ylm[ORC_LM(l, q)]) * Clm[(size_t)(ii - 1) * nlm + ORC_LM(l, q)];
              v += (Dlfac * ylm[ORC_LM(l, q)]) * Dlm[(size_t)ii * nlm + ORC_LM(l, q)];
            }
          }
        }
        phi[mp_corner(s, ii, jj, kk)] = v;
      }
    }
  }
  free(Clm);
  free(Dlm);
}

/* The sources of the multipole potential (gravity.c:919-1103): per zone the
 * gradient of the corner potential along each direction, as a weighted sum
 * over the eight corners (weights from the volume centroid, 1/2 in phi),
 * the momentum source -rho dPhi/dx and the energy source -<rho v> dPhi/dx
 * with <rho v> from this stage's mass fluxes per area, weighted alike. */
static void multipole(orc_sim *s, const double *P)
{
  const orc_config *c = &s->c;
  if (c->ndim == 1) return;
  const int n1 = c->n[0], nv = s->nvars;
  sphmap m;
  sphmap_init(&m, c->rcoord, c->rtrans);
  double *phi = calloc((size_t)(n1 + 1) * (c->n[1] + 1) * (c->n[2] + 1), sizeof(double));
  multipole_phi(s, P, phi);
  const double *F0 = s->flux[0], *F1 = s->flux[1], *F2 = s->flux[2];
  const double sgn[2] = {-1.0, 1.0};
  for (int ii = 0; ii < n1; ii++)
    for (int jj = 0; jj < njs(s, ii); jj++)
      for (int kk = 0; kk < nks(s, ii, jj); kk++) {
        double wx[2], wy[2] = {0.5, 0.5}, wz[2] = {0.5, 0.5};
        double dx1 = 1.0, dx2 = 1.0;
        const double rm = r_of_x(&m, c->startx[0] + ii * c->dx[0]);
        const double rp = r_of_x(&m, c->startx[0] + (ii + 1) * c->dx[0]);
        const double rc = r_of_x(&m, s->mom[1][0][ii + ORC_NG] / s->mom[2][0][ii + ORC_NG]);
        wx[0] = (rp - rc) / (rp - rm);
        wx[1] = (rc - rm) / (rp - rm);
        const double dx0 = c->dx[0];
        const int lev = lev_of(djs(s, ii));
        {
          const double thm = th_of_x(&m, c->startx[1] + (jj * djs(s, ii)) * c->dx[1]);
          const double thp = th_of_x(&m, c->startx[1] + ((jj + 1) * djs(s, ii)) * c->dx[1]);
          const double thc = th_of_x(&m, lev_arr(s, 1, 0, lev)[jj] / lev_arr(s, 2, 0, lev)[jj]);
          wy[0] = (thp - thc) / (thp - thm);
          wy[1] = (thc - thm) / (thp - thm);
          dx1 = djs(s, ii) * c->dx[1];
        }
        if (c->ndim == 3) {
          const double phm = c->startx[2] + (kk * dks(s, ii, jj)) * c->dx[2];
          const double php = c->startx[2] + ((kk + 1) * dks(s, ii, jj)) * c->dx[2];
          const double phc = 0.5 * (phm + php);
          wz[0] = (php - phc) / (php - phm);
          wz[1] = (phc - phm) / (php - phm);
          dx2 = dks(s, ii, jj) * c->dx[2];
        }
        const int nkk = c->ndim == 3 ? 2 : 1;          /* 2D: no phi corners (m = 0 only) */
        const long z = zid(s, ii, jj, kk);
        const double rho = P[z * nv + ORC_RHO];
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
        for (int a = 0; a <= 1; a++) {
          const int ip = ii + a;
          for (int b = 0; b <= 1; b++) {
            const int jp = ((jj + b) * djs(s, ii)) / djs(s, ip);
            for (int q = 0; q < nkk; q++) {
              const int kp = c->ndim == 3 ? ((kk + q) * dks(s, ii, jj)) / dks(s, ip, jp) : 0;
              const double ph = phi[mp_corner(s, ip, jp, kp)];
              const double w2 = c->ndim == 3 ? wz[q] : 1.0;
              d0 += ((ph * sgn[a]) * wy[b]) * w2;
              d1 += ((ph * wx[a]) * sgn[b]) * w2;
              if (c->ndim == 3) d2 += ((ph * wx[a]) * wy[b]) * sgn[q];
            }
          }
        }
        /* radial <rho v>: the lower face, and the faces of the next shell over this zone */
        const double dA0 = s->geom[z].area[0] + 1.0e-16;
        const double vx0 = F0[z * nv + ORC_RHO] / dA0;
        const int njp = djs(s, ii) / djs(s, ii + 1), jp0 = jj * njp;
        double dA1 = 0.0, vx1 = 0.0;
        for (int j = jp0; j < jp0 + njp; j++) {
          const int nkp = dks(s, ii, jj) / dks(s, ii + 1, j), kp0 = kk * nkp;
          for (int k = kp0; k < kp0 + nkp; k++) {
            const long zf = zid(s, ii + 1, j, k);
            dA1 += s->geom[zf].area[0] + 1.0e-16;
            vx1 += F0[zf * nv + ORC_RHO];
          }
        }
        vx1 /= dA1;
        const double vx = (vx0 * wx[0]) + (vx1 * wx[1]);
        s->src[z * nv + ORC_U1] -= (rho * d0) / dx0;
        s->src[z * nv + ORC_ETOT] -= (vx * d0) / dx0;
        /* theta <rho v>, with the phi level rules of the faces (gravity.c:1030-1058) */
        double a1 = 0.0, vy1 = 0.0, a0 = 0.0, vy0 = 0.0;
        const int dk0 = dks(s, ii, jj), dkp = dks(s, ii, jj + 1), dkm = dks(s, ii, jj - 1);
        if (dkp > dk0) {
          a1 += (0.5 * s->geom[zid(s, ii, jj + 1, kk / 2)].area[1]) + 1.0e-16;
          vy1 += F1[zid(s, ii, jj + 1, kk) * nv + ORC_RHO];
        } else if (dkp < dk0) {
          a1 += (s->geom[zid(s, ii, jj + 1, 2 * kk)].area[1] + s->geom[zid(s, ii, jj + 1, 2 * kk + 1)].area[1]) + 1.0e-16;
          vy1 += F1[zid(s, ii, jj + 1, 2 * kk) * nv + ORC_RHO] + F1[zid(s, ii, jj + 1, 2 * kk + 1) * nv + ORC_RHO];
        } else {
          a1 += s->geom[zid(s, ii, jj + 1, kk)].area[1] + 1.0e-16;
          vy1 += F1[zid(s, ii, jj + 1, kk) * nv + ORC_RHO];
        }
        vy1 /= a1;
        if (dkm < dk0) {
          a0 += s->geom[z].area[1] + 1.0e-16;
          vy0 += F1[zid(s, ii, jj, 2 * kk) * nv + ORC_RHO] + F1[zid(s, ii, jj, 2 * kk + 1) * nv + ORC_RHO];
        } else {
          a0 += s->geom[z].area[1] + 1.0e-16;
          vy0 += F1[z * nv + ORC_RHO];
        }
        vy0 /= a0;
        const double vy = (vy0 * wy[0]) + (vy1 * wy[1]);
        s->src[z * nv + ORC_U2] -= (rho * d1) / dx1;
        s->src[z * nv + ORC_ETOT] -= (vy * d1) / dx1;
        if (c->ndim == 3) {
          const long z2 = zid(s, ii, jj, kk + 1);
          const double vz0 = F2[z * nv + ORC_RHO] / (s->geom[z].area[2] + 1.0e-16);
          const double vz1 = F2[z2 * nv + ORC_RHO] / (s->geom[z2].area[2] + 1.0e-16);
          const double vz = (vz0 * wz[0]) + (vz1 * wz[1]);
          s->src[z * nv + ORC_U3] -= (rho * d2) / dx2;
          s->src[z * nv + ORC_ETOT] -= (vz * d2) / dx2;
        }
      }
  free(phi);
}

int orc_multipole_sources(orc_sim *s, double *src)
{
  if (s->c.multipole_lmax < 1 || s->c.ndim < 2) return -1;
  const long n = s->nall * s->nvars;
  for (long q = 0; q < n; q++) s->src[q] = 0.0;
  for (int d = 0; d < 3; d++)
    for (long q = 0; q < n; q++) s->flux[d][q] = 0.0;
  multipole(s, s->p);
  const orc_config *c = &s->c;
  long o = 0;
  for (int k = 0; k < c->n[2]; k++)
    for (int j = 0; j < c->n[1]; j++)
      for (int i = 0; i < c->n[0]; i++, o++)
        for (int v = 0; v < 3; v++) src[3 * o + v] = s->src[zid(s, i, j, k) * s->nvars + ORC_U1 + v];
  return 0;
}

int orc_multipole_phi(orc_sim *s, double *phi)
{
  if (s->c.multipole_lmax < 1 || s->c.ndim < 2) return -1;
  multipole_phi(s, s->p, phi);
  return 0;
}

/* The ghost fill that starts each stage, applied to the current state; with
 * orc_zone_prims it lets a test read any ghost zone. */
void orc_fill_ghosts(orc_sim *s)
{
  if (s->dj) fill_ghost_dend(s, s->p);
  else reset_boundaries(s, s->p);
}

int orc_zone_prims(const orc_sim *s, int i, int j, int k, double *out)
{
  const int idx[3] = {i, j, k};
  for (int d = 0; d < 3; d++)
    if (idx[d] < -s->g[d] || idx[d] >= s->c.n[d] + s->g[d]) return -1;
  const double *src = &s->p[zid(s, i, j, k) * s->nvars];
  for (int v = 0; v < s->nvars; v++) out[v] = src[v];
  return 0;
}

static void stage(orc_sim *s, int st)
{
  double *P = st == 0 ? s->p : s->ph;
  /* the second stage's fill runs at t + dt and takes dt off again after it
   * (bc.c:717-721, 1417: half_step_sync), so the global t of the reference
   * becomes ((t + dt) - dt) before the main loop adds dt */
  if (st == 1) s->t += s->dt;
  if (s->dj) fill_ghost_dend(s, P);
  else reset_boundaries(s, P);
  if (st == 1) s->t -= s->dt;
  update_eos(s, P);
  sources(s, P);
  if (s->dj) sweep_dend(s, P);
  else
    for (int d = 0; d < s->c.ndim; d++) sweep_dir(s, P, d);
  if (s->c.use_monopole) monopole(s, P);
  if (s->c.multipole_lmax > 0) multipole(s, P);
  if (s->dj) {
    if (st == 0) push_dend(s, s->dt, s->p, s->p, s->ph, 0);
    else push_dend(s, s->dt, s->p, s->ph, s->p, 1);
  } else if (st == 0) push(s, s->dt, s->p, s->p, s->ph, 0);
  else push(s, s->dt, s->p, s->ph, s->p, 1);
}

/* step.c:28-129 */
void orc_step(orc_sim *s)
{
  s->min_dt = 1.0e100;
  const orc_config *c = &s->c;
  if (c->ndim > 1) {   /* set_shock_flag(0): interior zones only (fluid.c:82-89) */
    for (int k = 0; k < c->n[2]; k++)
      for (int j = 0; j < c->n[1]; j++)
        for (int i = 0; i < c->n[0]; i++)
          if (zone_valid(s, i, j, k)) s->shock[zid(s, i, j, k)] = -1;
  }
  stage(s, 0);
  stage(s, 1);
}

/* step.c:163-203.  Uses the eos array as last filled (by the previous
 * step's second stage, or by the initial update_eos). */
double orc_estimate_dt(orc_sim *s)
{
  const orc_config *c = &s->c;
  if (s->firstc) { s->min_dt = 1.0e100; s->firstc = 0; }
  for (int k = 0; k < c->n[2]; k++)
    for (int j = 0; j < c->n[1]; j++)
      for (int i = 0; i < c->n[0]; i++) {
        if (!zone_valid(s, i, j, k)) continue;
        long z = zid(s, i, j, k);
        double local_dt = 0.0;
        for (int d = 0; d < c->ndim; d++) {
          double local_dx = c->dx[d];
          if (d == 1) local_dx *= djs(s, i);
          if (d == 2) local_dx *= dks(s, i, j);
          local_dt += (fabs(s->p[z * s->nvars + ORC_U1 + d]) + (s->eos[z * NEOS + E_CS] / s->geom[z].scale[0][d])) / local_dx;
          s->min_dt = OMIN(s->min_dt, 1.0 / local_dt);
        }
      }
  if (s->istep == 0) s->min_dt = OMIN(s->min_dt, c->initial_dt);
  s->min_dt = OMIN(s->min_dt, c->dtmax);
  s->min_dt *= c->cfl;
  return s->min_dt;
}

/* main.c:206-265 without dumps (dt_dump = dt_pdump = 1e100). */
int orc_run(orc_sim *s, double tmax, int max_steps)
{
  const orc_config *c = &s->c;
  int taken = 0;
  if (s->istep == 0 && s->firstc) update_eos(s, s->p);
  while (s->t < tmax && taken < max_steps) {
    s->dt = orc_estimate_dt(s);
    if (s->dt < c->dtmin) return -1;
    if (s->istep == 0) s->dt *= c->dt_init_safety;
    else s->dt = OMIN(s->dtold * c->dt_grow_safety, s->dt);
    s->dtold = s->dt;
    if (s->t + 1 * s->dt > tmax) s->dt = (1.000000001 * (tmax - s->t)) / 1;
    if (s->dt < c->dtmin) s->dt = c->dtmin;
    orc_step(s);
    s->t += s->dt;
    s->istep++;
    taken++;
  }
  return taken;
}

/* Neumaier-compensated sums of rho*V and E*V over interior zones. */
void orc_conserved_totals(const orc_sim *s, double out[2])
{
  const orc_config *c = &s->c;
  double sum[2] = {0, 0}, comp[2] = {0, 0};
  for (int k = 0; k < c->n[2]; k++)
    for (int j = 0; j < c->n[1]; j++)
      for (int i = 0; i < c->n[0]; i++) {
        if (!zone_valid(s, i, j, k)) continue;
        long z = zid(s, i, j, k);
        double u[ORC_NSIZE];
        orc_p_to_u(&s->p[z * s->nvars], u, s->geom[z].gcov[0], s->nhydro);
        for (int q = 0; q < 2; q++) {
          double v = u[q] * s->geom[z].volume;
          double t = sum[q] + v;
          if (fabs(sum[q]) >= fabs(v)) comp[q] += (sum[q] - t) + v;
          else comp[q] += (v - t) + sum[q];
          sum[q] = t;
        }
      }
  out[0] = sum[0] + comp[0];
  out[1] = sum[1] + comp[1];
}

/* dendritic level counts: nj[i] (n1 entries) and nk[i*n2 + j] for j < nj[i]
 * (0 beyond); uniform grids report n2 and n3.  Returns the maximum theta
 * and phi refinement levels as maxjlev * 256 + maxklev. */
int orc_dendritic_levels(const orc_sim *s, int *nj, int *nk)
{
  const int n1 = s->c.n[0], n2 = s->c.n[1];
  for (int i = 0; i < n1; i++) {
    nj[i] = njs(s, i);
    for (int j = 0; j < n2; j++) nk[(long)i * n2 + j] = j < njs(s, i) ? nks(s, i, j) : 0;
  }
  return s->maxjlev * 256 + s->maxklev;
}

/* theta (dir 0) or phi (dir 1) moments of refinement level lev, length n + 2 NG */
const double *orc_level_moment(const orc_sim *s, int which, int dir, int lev)
{
  const int maxl = dir == 0 ? s->maxjlev : s->maxklev;
  if (lev < 0 || lev > maxl || !s->lmom[which][dir]) return 0;
  return lev_arr(s, which, dir, lev) - ORC_NG;
}

/* ------------------------------------------------------------------ */
/* Entry points for pinning the pencil entries and the polar phi-face
 * states (tests only; the stepping code never calls them).  They read the
 * current state after orc_prepare_entries has applied the stage-start
 * ghost fill and EOS update (step.c:76-81). */

void orc_prepare_entries(orc_sim *s)
{
  orc_fill_ghosts(s);
  update_eos(s, s->p);
}

/* Entry of row ii of the radial pencil of the face below zone (i,j,k) of a
 * 3D spherical grid, as sweep_dend chooses it (step.c:1159-1175): restrict0
 * when row ii is finer in theta than row i, prolongate0 otherwise.  out:
 * nvars primitives, then P and Gamma_1.  Returns -1 outside the grid. */
int orc_dend_radial_entry(orc_sim *s, int i, int j, int k, int ii, double *out)
{
  const orc_config *c = &s->c;
  if (c->ndim != 3 || c->geom != ORC_GEOM_SPHERICAL) return -1;
  if (i < 0 || i > c->n[0] || ii < i - ORC_NG || ii > i + ORC_NG - 1) return -1;
  if (j < 0 || j >= njs(s, i) || k < 0 || k >= nks(s, i, j)) return -1;
  if (djs(s, ii) < djs(s, i)) restrict0(s, s->p, i, j, k, ii, out);
  else prolongate0(s, s->p, i, j, k, ii, out);
  return 0;
}

/* Entry of row jj of the theta pencil of the face below row j of shell i,
 * at phi index kk of the face's (finer) phi level (theta_row: prolongate1,
 * restrict1 or the zone itself, step.c:1201-1240). */
int orc_dend_theta_entry(orc_sim *s, int i, int j, int kk, int jj, double *out)
{
  const orc_config *c = &s->c;
  if (c->ndim != 3 || c->geom != ORC_GEOM_SPHERICAL) return -1;
  if (i < 0 || i >= c->n[0] || j < 0 || j > njs(s, i) || jj < j - ORC_NG || jj > j + ORC_NG - 1) return -1;
  const int dkf = OMIN(dks(s, i, j), dks(s, i, j - 1));
  if (kk < 0 || kk >= c->n[2] / dkf) return -1;
  theta_row(s, s->p, i, j, kk, jj, out);
  return 0;
}

/* The two states of the phi face below zone (i, j, k) in the last theta row
 * (step.c:1299-1397), after reconstruction and the polar-axis treatment:
 * left = the upper state of zone k-1, right = the lower state of zone k,
 * nvars primitives then P and Gamma_1 each. */
int orc_polar_face_states(orc_sim *s, int i, int j, int k, double *left, double *right)
{
  const orc_config *c = &s->c;
  if (c->ndim != 3 || c->geom != ORC_GEOM_SPHERICAL) return -1;
  if (i < 0 || i >= c->n[0] || j != njs(s, i) - 1 || k < 0 || k > nks(s, i, j)) return -1;
  const int nv = s->nvars, dk = dks(s, i, j), lk = lev_of(dk);
  const double *A = lev_arr(s, 0, 1, lk), *B = lev_arr(s, 1, 1, lk), *G = lev_arr(s, 2, 1, lk);
  double pen[ORC_NSIZE][2 * ORC_NG], pleft[ORC_NSIZE][2 * ORC_NG], pright[ORC_NSIZE][2 * ORC_NG];
  double al[2 * ORC_NG], be[2 * ORC_NG], ga[2 * ORC_NG], x[2 * ORC_NG], pp[ORC_NSIZE];
  for (int q = 0; q < 2 * ORC_NG; q++) {
    const int kt = k - ORC_NG + q;
    x[q] = c->startx[2] + kt * dk * c->dx[2];
    al[q] = A[kt];
    be[q] = B[kt];
    ga[q] = G[kt];
    zone_row(s, s->p, i, j, kt, pp);
    for (int vv = 0; vv < nv + 2; vv++) pen[vv][q] = pp[vv];
  }
  interp_pencil(s, pen, pleft, pright, al, be, ga, x);
  polar_states(s, s->p, i, j, k, pen, pleft, pright);
  for (int vv = 0; vv < nv + 2; vv++) {
    left[vv] = pright[vv][ORC_NG - 1];
    right[vv] = pleft[vv][ORC_NG];
  }
  return 0;
}
