/*
 * mds_problems.c -- problem initialisers of the benchmark configurations
 * (host C, the reference's init_problem() role, problems/init_*.c).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include "../../../include/mds_host.h"
#include "../../../include/mds_eos.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846264338327
#endif

unsigned long long mds_splitmix64(unsigned long long seed, unsigned long long k)
{
  unsigned long long z = seed + (k + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double mds_uniform(unsigned long long seed, unsigned long long k)
{
  return (double)(mds_splitmix64(seed, k) >> 11) * (1.0 / 9007199254740992.0);
}

static size_t nzones(const mds_grid *g)
{
  return (size_t)g->n[0] * (size_t)(g->ndim > 1 ? g->n[1] : 1) * (size_t)(g->ndim > 2 ? g->n[2] : 1);
}

int mds_init_sod(const mds_grid *g, double gam, double *prims)
{
  if (!g || !prims || g->ndim != 1) return MDS_EINVAL;
  size_t nz = nzones(g);
  const double rhol = 1.0, el = 1.0 / (gam - 1.0), ul = 0.0;
  const double rhor = 0.125, er = 0.1 / (gam - 1.0), ur = 0.0;
  const double dx = g->dx[0];
  for (int i = 0; i < g->n[0]; i++) {
    double xl = g->startx[0] + i * dx;
    double xr = xl + dx;
    double rho, e, u;
    if (xr < 0.5) {
      rho = rhol; e = el; u = ul;
    } else if (xl < 0.5) {   /* straddling zone: volume-fraction mix (init_sod.c:75-80) */
      rho = rhol * (0.5 - xl) / dx + rhor * (xr - 0.5) / dx;
      e = el * (0.5 - xl) / dx + er * (xr - 0.5) / dx;
      u = ul * (0.5 - xl) / dx + ur * (xr - 0.5) / dx;
    } else {
      rho = rhor; e = er; u = ur;
    }
    prims[0 * nz + i] = rho;
    prims[1 * nz + i] = e;
    prims[2 * nz + i] = u;
    prims[3 * nz + i] = 0.0;
    prims[4 * nz + i] = 0.0;
  }
  return MDS_OK;
}

int mds_init_kh(const mds_grid *g, double gam, double p0, double amp, double noise,
                unsigned long long seed, double *prims)
{
  if (!g || !prims || g->ndim != 2) return MDS_EINVAL;
  size_t nz = nzones(g);
  for (int j = 0; j < g->n[1]; j++) {
    double y = g->startx[1] + (j + 0.5) * g->dx[1];
    int inner = fabs(y - 0.5) < 0.25;
    for (int i = 0; i < g->n[0]; i++) {
      size_t q = (size_t)j * g->n[0] + i;
      double x = g->startx[0] + (i + 0.5) * g->dx[0];
      double r = 2.0 * mds_uniform(seed, q) - 1.0;
      prims[0 * nz + q] = inner ? 2.0 : 1.0;
      prims[1 * nz + q] = p0 / (gam - 1.0);
      prims[2 * nz + q] = inner ? 0.5 : -0.5;
      prims[3 * nz + q] = amp * sin(4.0 * M_PI * x) + noise * r;
      prims[4 * nz + q] = 0.0;
    }
  }
  return MDS_OK;
}

int mds_init_rt(const mds_grid *g, double gam, double p0, double gz, double rho_lo, double rho_hi,
                double amp, double noise, unsigned long long seed, double *prims)
{
  if (!g || !prims || g->ndim != 3) return MDS_EINVAL;
  size_t nz = nzones(g);
  for (int k = 0; k < g->n[2]; k++) {
    double z = g->startx[2] + (k + 0.5) * g->dx[2];
    double rho = (z > 0.0) ? rho_hi : rho_lo;
    double p = p0 + gz * rho * z;
    for (int j = 0; j < g->n[1]; j++) {
      double y = g->startx[1] + (j + 0.5) * g->dx[1];
      for (int i = 0; i < g->n[0]; i++) {
        double x = g->startx[0] + (i + 0.5) * g->dx[0];
        size_t q = ((size_t)k * g->n[1] + j) * g->n[0] + i;
        double r = 2.0 * mds_uniform(seed, q) - 1.0;
        prims[0 * nz + q] = rho;
        prims[1 * nz + q] = p / (gam - 1.0);
        prims[2 * nz + q] = 0.0;
        prims[3 * nz + q] = 0.0;
        prims[4 * nz + q] = amp * (1.0 + cos(4.0 * M_PI * x)) * (1.0 + cos(4.0 * M_PI * y))
                            * (1.0 + cos(3.0 * M_PI * z)) / 8.0 + noise * r;
      }
    }
  }
  return MDS_OK;
}

/* r(x) of the grid's radial map (metric_spherical_general.c:149-161, metric_spherical_logr.c) */
static double radius_of(const mds_grid *g, double x)
{
  if (g->rcoord == MDS_RCOORD_LOG) return exp(x);
  if (g->rcoord == MDS_RCOORD_SINH) return g->rtrans * sinh(fabs(x) / g->rtrans);
  return fabs(x);
}

int mds_init_sph_riemann(const mds_grid *g, double gam, double rho_in, double p_in, double r0,
                         double frho, double fp, double *prims)
{
  if (!g || !prims || g->ndim != 1) return MDS_EINVAL;
  size_t nz = nzones(g);
  for (int i = 0; i < g->n[0]; i++) {
    double xc = g->startx[0] + (i + 0.5) * g->dx[0];
    double r = radius_of(g, xc);
    int inner = r < r0;
    double rho = inner ? rho_in : rho_in / frho;
    double p = inner ? p_in : p_in / fp;
    prims[0 * nz + i] = rho;
    prims[1 * nz + i] = p / (gam - 1.0);
    prims[2 * nz + i] = 0.0;
    prims[3 * nz + i] = 0.0;
    prims[4 * nz + i] = 0.0;
  }
  return MDS_OK;
}

int mds_ensemble_params(int nprob, int nmin, int nmax, unsigned long long seed, double *out)
{
  if (nprob < 1 || nmin < 4 || nmax < nmin || !out) return MDS_EINVAL;
  unsigned long long k = 0;
  for (int p = 0; p < nprob; p++) {
    double u0 = mds_uniform(seed, k++);
    int n = nmin + (int)floor(u0 * (double)(nmax - nmin + 1));
    if (n > nmax) n = nmax;
    double *o = &out[7 * (size_t)p];
    o[0] = n;
    o[1] = 1.3 + 0.37 * mds_uniform(seed, k++);
    o[2] = pow(10.0, 8.0 + 2.0 * mds_uniform(seed, k++));
    o[3] = pow(10.0, 26.0 + 2.0 * mds_uniform(seed, k++));
    o[4] = pow(10.0, 6.0 + 2.0 * mds_uniform(seed, k++));
    o[5] = pow(10.0, 1.0 + 2.0 * mds_uniform(seed, k++));
    o[6] = pow(10.0, 1.0 + 2.0 * mds_uniform(seed, k++));
  }
  return MDS_OK;
}

/* ------------------------------------------------------------------------ */
/* Polytropic star (intent of problems/init_ccsn.c:51-394 for a synthetic
 * progenitor): Lane-Emden solution of index n, rho = rho_c theta^n,
 * P = K rho^(1+1/n) inside the first zero, (rho_atm, K rho_atm^(1+1/n))
 * outside; seeded lognormal density perturbations rho *= exp(sigma z) at
 * constant pressure; Ye = 0.45 + 0.05 (1 - (rho/rho_c)^(1/3)).  With a table
 * the temperature solves P_model(rho, T, Ye) = P (bisection in log T over
 * the table's range less 2% at each end, clamped there) and u = rho e_model;
 * without one,
 * u = P / (gam - 1). */

#define LE_H 1.0e-4L
#define LE_MAX 200000

static int lane_emden(long double n, long double *theta, long double *dtheta, int cap)
{
  /* RK4 in xi for y = (theta, theta'); series start at xi = h */
  long double h = LE_H, xi = h;
  long double y0 = 1.0L - h * h / 6.0L + n * h * h * h * h / 120.0L, y1 = -h / 3.0L + n * h * h * h / 30.0L;
  theta[0] = 1.0L; dtheta[0] = 0.0L;
  theta[1] = y0; dtheta[1] = y1;
  int k = 1;
  while (k + 1 < cap) {
    long double a0 = y1, b0 = -2.0L / xi * y1 - powl(fmaxl(y0, 0.0L), n);
    long double ya = y0 + 0.5L * h * a0, yb = y1 + 0.5L * h * b0, xa = xi + 0.5L * h;
    long double a1 = yb, b1 = -2.0L / xa * yb - powl(fmaxl(ya, 0.0L), n);
    ya = y0 + 0.5L * h * a1; yb = y1 + 0.5L * h * b1;
    long double a2 = yb, b2 = -2.0L / xa * yb - powl(fmaxl(ya, 0.0L), n);
    ya = y0 + h * a2; yb = y1 + h * b2;
    long double a3 = yb, b3 = -2.0L / (xi + h) * yb - powl(fmaxl(ya, 0.0L), n);
    y0 += h / 6.0L * (a0 + 2 * a1 + 2 * a2 + a3);
    y1 += h / 6.0L * (b0 + 2 * b1 + 2 * b2 + b3);
    xi += h;
    k++;
    theta[k] = y0; dtheta[k] = y1;
    if (y0 <= 0.0L) break;
  }
  return k;   /* theta[k] <= 0: first zero in (xi_{k-1}, xi_k] */
}

static double temp_for_pressure(const double prm[3], const mds_eos_table *t, double rho, double ye, double P)
{
  /* keep 2% of the log T range clear of the table edges */
  double lo = t->lt0 + 0.02 * (t->lt1 - t->lt0), hi = t->lt1 - 0.02 * (t->lt1 - t->lt0), v[3];
  mds_eos_model_eval(prm, rho, pow(10.0, lo), ye, v);
  if (v[0] >= P) return pow(10.0, lo);
  mds_eos_model_eval(prm, rho, pow(10.0, hi), ye, v);
  if (v[0] <= P) return pow(10.0, hi);
  for (int it = 0; it < 80; it++) {
    double mid = 0.5 * (lo + hi);
    mds_eos_model_eval(prm, rho, pow(10.0, mid), ye, v);
    if (v[0] < P) lo = mid; else hi = mid;
  }
  return pow(10.0, 0.5 * (lo + hi));
}

int mds_init_polytrope(const mds_grid *g, double n_index, double rho_c, double K, double rho_atm, double gam,
                       const mds_eos_table *tab, unsigned long long eos_seed, double sigma,
                       unsigned long long seed, double *prims)
{
  if (!g || !prims || g->geom != MDS_GEOM_SPHERICAL || !(rho_c > 0.0) || !(K > 0.0)) return MDS_EINVAL;
  long double *th = (long double *)malloc(sizeof(long double) * LE_MAX);
  long double *dth = (long double *)malloc(sizeof(long double) * LE_MAX);
  const int n1 = g->n[0];
  double *rho0 = (double *)malloc(sizeof(double) * n1), *e0 = (double *)malloc(sizeof(double) * n1);
  double *ye0 = (double *)malloc(sizeof(double) * n1);
  if (!th || !dth || !rho0 || !e0 || !ye0) {
    free(th); free(dth); free(rho0); free(e0); free(ye0);
    return MDS_ENOMEM;
  }
  const int kz = lane_emden(n_index, th, dth, LE_MAX);
  const double GN = 6.67384e-8;
  const double alpha = sqrt((n_index + 1.0) * K * pow(rho_c, 1.0 / n_index - 1.0) / (4.0 * M_PI * GN));
  const double xi1 = (double)((kz - 1) * LE_H + LE_H * th[kz - 1] / (th[kz - 1] - th[kz]));
  double prm[3] = {0.0, 0.0, 0.0};
  if (tab) mds_eos_model_params(eos_seed, prm);
  /* the unperturbed star on the radial zones */
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n1; i++) {
    const double xc = g->beta[0][i + MDS_NG] / g->Gamma[0][i + MDS_NG];
    const double r = radius_of(g, xc);
    const double xi = r / alpha;
    double rho;
    if (xi < xi1) {
      const double s = xi / (double)LE_H;
      int k = (int)s;
      if (k >= kz) k = kz - 1;
      const double f = s - k;
      /* cubic Hermite on theta, theta' */
      const double h = (double)LE_H, t0 = (double)th[k], t1 = (double)th[k + 1];
      const double d0 = (double)dth[k] * h, d1 = (double)dth[k + 1] * h;
      const double f2 = f * f, f3 = f2 * f;
      double theta = (2 * f3 - 3 * f2 + 1) * t0 + (f3 - 2 * f2 + f) * d0 + (-2 * f3 + 3 * f2) * t1 + (f3 - f2) * d1;
      if (theta < 0.0) theta = 0.0;
      rho = rho_c * pow(theta, n_index);
      if (rho < rho_atm) rho = rho_atm;
    } else {
      rho = rho_atm;
    }
    const double P = K * pow(rho, 1.0 + 1.0 / n_index);
    const double ye = 0.45 + 0.05 * (1.0 - cbrt(fmin(rho / rho_c, 1.0)));
    double e;
    if (tab) {
      const double T = temp_for_pressure(prm, tab, rho, ye, P);
      double v[3];
      mds_eos_model_eval(prm, rho, T, ye, v);
      e = v[1];
    } else {
      e = P / ((gam - 1.0) * rho);
    }
    rho0[i] = rho;
    e0[i] = e;
    ye0[i] = ye;
  }
  /* zones: radial profile times the lognormal density factor at constant
   * specific internal energy and Ye */
  const size_t nz = nzones(g);
#pragma omp parallel for schedule(static)
  for (long long zz = 0; zz < (long long)nz; zz++) {
    const size_t z = (size_t)zz;
    const int i = (int)(z % (size_t)n1);
    double fac = 1.0;
    if (sigma > 0.0) {
      double u1 = 1.0 - mds_uniform(seed, 2 * (unsigned long long)z);
      double u2 = mds_uniform(seed, 2 * (unsigned long long)z + 1);
      fac = exp(sigma * (sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2)));
    }
    const double rho = rho0[i] * fac;
    prims[0 * nz + z] = rho;
    prims[1 * nz + z] = rho * e0[i];
    prims[2 * nz + z] = 0.0;
    prims[3 * nz + z] = 0.0;
    prims[4 * nz + z] = 0.0;
    prims[5 * nz + z] = ye0[i];
  }
  free(th); free(dth); free(rho0); free(e0); free(ye0);
  return MDS_OK;
}
