/*
 * mds_problems.c -- problem initialisers of the benchmark configurations
 * (host C, the reference's init_problem() role, problems/init_*.c).
 */
#include <math.h>
#include <stddef.h>
#include "../../../include/mds_host.h"

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

int mds_init_sph_riemann(const mds_grid *g, double gam, double rho_in, double p_in, double r0,
                         double frho, double fp, double *prims)
{
  if (!g || !prims || g->ndim != 1) return MDS_EINVAL;
  size_t nz = nzones(g);
  for (int i = 0; i < g->n[0]; i++) {
    double xc = g->startx[0] + (i + 0.5) * g->dx[0];
    double r = (g->rcoord == MDS_RCOORD_LOG) ? exp(xc) : fabs(xc);
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
