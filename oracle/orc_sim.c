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
};

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
typedef struct { int rcoord; double poly_norm, poly_xt, poly_alpha; } sphmap;

static double r_of_x(const sphmap *m, double x) { return m->rcoord == ORC_RCOORD_LOG ? exp(x) : fabs(x); }
static double dr_dx(const sphmap *m, double x) { return m->rcoord == ORC_RCOORD_LOG ? exp(x) : 1.0; }
static double d2r_dx2(const sphmap *m, double x) { return m->rcoord == ORC_RCOORD_LOG ? exp(x) : 0.0; }
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

static void sphmap_init(sphmap *m, int rcoord)
{
  m->rcoord = rcoord;
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

/* 2D/3D spherical (r, theta, phi) geometry on a uniform (non-dendritic) grid,
 * metric_spherical_general.c:592-1051 with DJS = DKS = 1: x1 in [-1,1] through
 * the POLYTH theta map, phi = x2 (dx[2] = 2 pi in 2D, init_coords :558-590).
 * Volume, metric and scale for every zone (ghosts included; theta ghosts
 * reflect through th_of_x), areas on the faces the reference loops over
 * (i <= n1, j <= n2) with the axis faces forced to zero (:808-833), and the
 * connection coefficients of interior zones with the pressure-balanced
 * Gamma^theta_theta_theta = dA/dV - Gamma^phi_theta_phi (:905-915). */
static void build_geometry_spherical(orc_sim *s)
{
  const orc_config *c = &s->c;
  sphmap m;
  sphmap_init(&m, c->rcoord);
  const int n1 = c->n[0], n2 = c->n[1], n3 = c->n[2];
  const double dph = c->ndim == 3 ? c->dx[2] : 2.0 * M_PI;
  const int nn[3] = {n1, n2, n3};
  for (int d = 0; d < 3; d++)
    for (int q = 0; q < 3; q++) s->mom[q][d] = calloc(nn[d] + 2 * ORC_NG, sizeof(double));
  /* init_interp_vol :592-669 */
  for (int i = -ORC_NG; i < n1 + ORC_NG; i++) {
    double x0 = c->startx[0] + i * c->dx[0], x1 = c->startx[0] + (i + 1) * c->dx[0];
    s->mom[2][0][i + ORC_NG] = orc_romb(x0, x1, f_r2dr_dx, &m);
    s->mom[1][0][i + ORC_NG] = orc_romb(x0, x1, f_xr1g, &m);
    s->mom[0][0][i + ORC_NG] = orc_romb(x0, x1, f_xr2g, &m);
  }
  for (int j = -ORC_NG; j < n2 + ORC_NG; j++) {
    double x0 = c->startx[1] + j * 1 * c->dx[1], x1 = c->startx[1] + (j + 1) * 1 * c->dx[1];
    s->mom[2][1][j + ORC_NG] = orc_romb(x0, x1, f_sthdth_dx, &m);
    s->mom[1][1][j + ORC_NG] = orc_romb(x0, x1, f_xth1g, &m);
    s->mom[0][1][j + ORC_NG] = orc_romb(x0, x1, f_xth2g, &m);
  }
  for (int k = -ORC_NG; k < n3 + ORC_NG; k++) {
    double x0 = c->startx[2] + k * 1 * c->dx[2], x1 = c->startx[2] + (k + 1) * 1 * c->dx[2];
    s->mom[2][2][k + ORC_NG] = 1 * c->dx[2];
    s->mom[1][2][k + ORC_NG] = (OSQR(x1) - OSQR(x0)) / 2.0;
    s->mom[0][2][k + ORC_NG] = (OCUBE(x1) - OCUBE(x0)) / 3.0;
  }
  const int kg = c->ndim == 3 ? ORC_NG : 0;
  for (int i = -ORC_NG; i < n1 + ORC_NG; i++) {
    double x0lo = c->startx[0] + i * c->dx[0], x0hi = c->startx[0] + (i + 1) * c->dx[0];
    double r2dr_int = orc_romb(x0lo, x0hi, f_r2dr_dx, &m);
    double vrf = r2dr_int;
    double rint_r2dr_dx3 = orc_romb(x0lo, x0hi, f_r2dr_dx3, &m);
    double r5 = 0.2 * (pow(r_of_x(&m, x0hi), 5) - pow(r_of_x(&m, x0lo), 5));
    for (int j = -ORC_NG; j < n2 + ORC_NG; j++) {
      double x1lo = c->startx[1] + j * 1 * c->dx[1], x1hi = c->startx[1] + (j + 1) * 1 * c->dx[1];
      double sthdth_int = orc_romb(x1lo, x1hi, f_sthdth_dx, &m);
      double vthf = sthdth_int;
      double thint_sthdth_dx3 = orc_romb(x1lo, x1hi, f_sthdth_dx3, &m);
      double thint_sth3dth_dx = orc_romb(x1lo, x1hi, f_sth3dth_dx, &m);
      orc_zone_geom g;
      memset(&g, 0, sizeof(g));
      g.volume = ((r2dr_int * sthdth_int) * 1) * dph;                          /* :708-740 */
      if (i >= 0 && i <= n1 && j >= 0 && j <= n2) {                            /* :745-806 */
        double dph_int = 1 * dph;
        g.area[0] = (f_r2dr_dx(x0lo, &m) * sthdth_int) * dph_int;
        g.area[1] = (r2dr_int * f_sthdth_dx(x1lo, &m)) * dph_int;
        if (c->ndim == 3) g.area[2] = r2dr_int * sthdth_int;
        if ((j == 0 && c->startx[1] == -1) || (j == n2 && fabs((c->startx[1] + n2 * c->dx[1]) - 1) < 1.0e-15))
          g.area[1] = 0.0;                                                     /* :808-833 */
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
      for (int k = -kg; k < n3 + kg; k++) s->geom[zid(s, i, j, k)] = g;
    }
  }
  /* init_conn :837-929, interior zones (after volumes and areas) */
  for (int i = 0; i < n1; i++) {
    double x0lo = c->startx[0] + i * c->dx[0], x0hi = c->startx[0] + (i + 1) * c->dx[0];
    double rint_r2d2r_dx2 = orc_romb(x0lo, x0hi, f_r2d2r_dx2, &m);
    double rint_r3 = orc_romb(x0lo, x0hi, f_r3, &m);
    double rint_rdr_dx2 = orc_romb(x0lo, x0hi, f_rdr_dx2, &m);
    double vrf = orc_romb(x0lo, x0lo + c->dx[0], f_r2dr_dx, &m);      /* Vr(x0lo) */
    for (int j = 0; j < n2; j++) {
      double x1lo = c->startx[1] + j * 1 * c->dx[1], x1hi = c->startx[1] + (j + 1) * 1 * c->dx[1];
      double vthf = orc_romb(x1lo, x1hi, f_sthdth_dx, &m);
      double thint_sthdth_dx3 = orc_romb(x1lo, x1hi, f_sthdth_dx3, &m);
      double thint_cthdth_dx2 = orc_romb(x1lo, x1hi, f_cthdth_dx2, &m);
      for (int k = 0; k < n3; k++) {
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
        double dA = s->geom[zid(s, i, j + 1, k)].area[1];
        dA -= g->area[1];
        double dV = g->volume;
        g->conn[1][1][1] = dA / dV - g->conn[2][1][2];
      }
    }
  }
}

/* 1D spherical geometry, metric_spherical_general.c:558-1051 with NDIM=1
 * (x1 spans [-1,1], dx[2]=2*pi; init_coords :558-590). */
static void build_geometry_spherical1d(orc_sim *s)
{
  const orc_config *c = &s->c;
  sphmap m;
  m.rcoord = c->rcoord;
  m.poly_xt = 1.2;
  m.poly_alpha = 4;
  m.poly_norm = ((0.5 * M_PI) * 1.0) / (1.0 + ((1.0 / (m.poly_alpha + 1.0)) * (1.0 / pow(m.poly_xt, m.poly_alpha))));
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
  orc_sim *s = calloc(1, sizeof(*s));
  s->c = *cfg;
  for (int d = 0; d < 3; d++) {
    if (d >= cfg->ndim) s->c.n[d] = 1;
    s->g[d] = (d < cfg->ndim) ? ORC_NG : 0;
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
  if (cfg->geom == ORC_GEOM_SPHERICAL && cfg->ndim == 1) build_geometry_spherical1d(s);
  else if (cfg->geom == ORC_GEOM_SPHERICAL) build_geometry_spherical(s);
  else build_geometry_cartesian(s);
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
  for (int d = 0; d < 3; d++) {
    free(s->flux[d]);
    for (int q = 0; q < 3; q++) free(s->mom[q][d]);
  }
  free(s);
}

void orc_set_prims(orc_sim *s, const double *prims)
{
  const int *n = s->c.n;
  long ncell = (long)n[0] * n[1] * n[2];
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
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
        switch (code) {
          case ORC_BC_SPHERICAL_ORIGIN:
            /* theta boundary across the pole: mirror in j, phi + pi, flip v2 and v3
             * (bc.c:1094-1121, 1235-1262); other directions are not supported */
            src = side ? 2 * nd - ig - 1 : -1 - ig;
            pole = (d == 1);
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
          default:
            src = side ? nd - 1 : 0;
        }
        int idx[3], sidx[3];
        idx[d] = ig; idx[o1] = a; idx[o2] = b;
        sidx[d] = src; sidx[o1] = a; sidx[o2] = b;
        if (pole) sidx[2] = (idx[2] + c->n[2] / 2) % c->n[2];
        double *dst = &P[zid(s, idx[0], idx[1], idx[2]) * nv];
        if (extrap) {
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
        } else {
          const double *sp = &P[zid(s, sidx[0], sidx[1], sidx[2]) * nv];
          for (int v = 0; v < nv; v++) dst[v] = sp[v];
          if (flip) dst[ORC_U1 + d] *= -1.0;
          if (pole) { dst[ORC_U2] *= -1.0; dst[ORC_U3] *= -1.0; }
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
  sphmap_init(&m, c->rcoord);
  const int nv = s->nvars;
  const double th = th_of_x(&m, s->mom[1][1][j + ORC_NG] / s->mom[2][1][j + ORC_NG]);
  const double sth = sin(th), cth = cos(th);
  for (int kt = k - 1; kt <= k; kt++) {
    long z = zid(s, i, j, kt);
    const orc_zone_geom *g = &s->geom[z];
    double ph = c->startx[2] + (kt + 0.5) * 1 * c->dx[2];
    double sph = sin(ph), cph = cos(ph);
    double lam[3][3] = {{sth * cph, cth * cph, -sph}, {sth * sph, cth * sph, cph}, {cth, -sth, 0.0}};
    double v[3];
    for (int l = 0; l < 3; l++) {
      v[l] = 0.0;
      for (int mm = 0; mm < 3; mm++) v[l] += lam[l][mm] * P[z * nv + ORC_U1 + mm] * g->scale[0][mm];
    }
    /* the face at phi_k: the lower face of zone k, the upper face of zone k-1 */
    ph = c->startx[2] + k * 1 * c->dx[2];
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
  sphmap m = {c->rcoord, 0.0, 0.0, 0.0};
  double mr = 0.0;
  grav[0] = 0.0;
  for (int ii = 1; ii <= n1; ii++) {
    double ms = 0.0;
    for (int k = 0; k < c->n[2]; k++)
      for (int j = 0; j < c->n[1]; j++) {
        long z = zid(s, ii - 1, j, k);
        ms += s->geom[z].volume * P[z * nv + ORC_RHO];
      }
    mr = mr + ms;
    double xe = c->startx[0] + ii * c->dx[0];
    grav[ii] = -GNEWT * mr / OSQR(r_of_x(&m, xe)) * dr_dx(&m, xe);
  }
}

void orc_monopole_edges(const orc_sim *s, double *grav) { monopole_edges(s, s->p, grav); }

static void monopole(orc_sim *s, const double *P)
{
  const orc_config *c = &s->c;
  const int n1 = c->n[0], nv = s->nvars;
  double *grav = calloc(n1 + 1, sizeof(double));
  monopole_edges(s, P, grav);
  for (int ii = 0; ii < n1; ii++) {
    double xl = c->startx[0] + ii * c->dx[0];
    double A = (grav[ii + 1] - grav[ii]) / c->dx[0];
    double B = grav[ii] - A * xl;
    const double be = s->mom[1][0][ii + ORC_NG], ga = s->mom[2][0][ii + ORC_NG];
    double gz = (A * be) / ga + B;
    for (int k = 0; k < c->n[2]; k++)
      for (int j = 0; j < c->n[1]; j++) {
        long z = zid(s, ii, j, k), zp = zid(s, ii + 1, j, k);
        double vr = s->flux[0][zp * nv + ORC_RHO];
        double area = s->geom[zp].area[0] + 1.0e-16;
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

static void stage(orc_sim *s, int st)
{
  double *P = st == 0 ? s->p : s->ph;
  reset_boundaries(s, P);
  update_eos(s, P);
  sources(s, P);
  for (int d = 0; d < s->c.ndim; d++) sweep_dir(s, P, d);
  if (s->c.use_monopole) monopole(s, P);
  if (st == 0) push(s, s->dt, s->p, s->p, s->ph, 0);
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
        for (int i = 0; i < c->n[0]; i++) s->shock[zid(s, i, j, k)] = -1;
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
        long z = zid(s, i, j, k);
        double local_dt = 0.0;
        for (int d = 0; d < c->ndim; d++) {
          double local_dx = c->dx[d];
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
