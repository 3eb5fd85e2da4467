/*
 * mds_host.h -- host-side C layer above the device C-ABI (mds_hydro.h).
 *
 * Mirrors the reference's problem-init and main-loop call surface:
 *   init_grid()/init_problem()  (problems/init_*.c, decs.h:935-951)
 *   estimate_dt() + step()      (main.c:228, :262; step.c:28-203)
 * for the benchmark configurations of this repository.  Problem states are
 * SoA arrays [5][n3][n2][n1] of the reference primitives (rho, u, v1, v2, v3).
 */
#ifndef MDS_HOST_H
#define MDS_HOST_H

#include "mds_hydro.h"
#include "mds_eos.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Counter-based splitmix64 (Vigna): output number k of the stream seeded
 * with `seed`, and a uniform double in [0,1) from its top 53 bits. */
unsigned long long mds_splitmix64(unsigned long long seed, unsigned long long k);
double mds_uniform(unsigned long long seed, unsigned long long k);

/* Sod shock tube on a 1D Cartesian grid (problems/init_sod.c:31-99):
 * rho,p,v = (1,1,0) left of x=0.5, (0.125,0.1,0) right, volume-fraction mix
 * in the straddling zone.  gam is the gamma-law index. */
int mds_init_sod(const mds_grid *g, double gam, double *prims);

/* 2D Kelvin-Helmholtz on [0,1]^2: p uniform, rho=2, vx=+0.5 for |y-0.5|<0.25,
 * rho=1, vx=-0.5 elsewhere, vy = amp*sin(4 pi x) + noise*U(-1,1) with the
 * uniform draw k = zone index of the seeded stream (intent of init_kh.c:23-89). */
int mds_init_kh(const mds_grid *g, double gam, double p0, double amp, double noise,
                unsigned long long seed, double *prims);

/* 3D Rayleigh-Taylor (intent of problems/init_rt_3d.c:94-236): rho=rho_hi for
 * z>0, rho_lo below, hydrostatic p = p0 + gz*rho*z, vz perturbation
 * amp*(1+cos 4 pi x)(1+cos 4 pi y)(1+cos 3 pi z)/8 + noise*U(-1,1). */
int mds_init_rt(const mds_grid *g, double gam, double p0, double gz, double rho_lo, double rho_hi,
                double amp, double noise, unsigned long long seed, double *prims);

/* Spherical Riemann problem on a 1D radial grid: (rho_in, p_in) for zone
 * centre radius r < r0, (rho_in/frho, p_in/fp) outside, at rest. */
int mds_init_sph_riemann(const mds_grid *g, double gam, double rho_in, double p_in, double r0,
                         double frho, double fp, double *prims);

/* Parameters of an ensemble of spherical Riemann problems drawn from one
 * seeded stream in problem order: n ~ U{nmin..nmax}, gam ~ U[1.3,1.67],
 * rho_in = 10^U(8,10), p_in = 10^U(26,28), r0 = 10^U(6,8), frho = 10^U(1,3),
 * fp = 10^U(1,3).  out[7*p + {0..6}] = {n, gam, rho_in, p_in, r0, frho, fp}. */
int mds_ensemble_params(int nprob, int nmin, int nmax, unsigned long long seed, double *out);

/* Polytropic star on a spherical grid (1D column or 3D), intent of
 * problems/init_ccsn.c:51-394 with a synthetic progenitor: Lane-Emden index
 * n_index on the radial zones, rho = rho_c theta^n, P = K rho^(1+1/n) inside
 * the first zero and (rho_atm, K rho_atm^(1+1/n)) outside, at rest; Ye =
 * 0.45 + 0.05 (1 - (rho/rho_c)^(1/3)); specific internal energy e from the
 * table's model at the T that gives P (eos_seed: the table's model
 * parameters), or P / ((gam - 1) rho) without a table.  Each zone's density
 * is then multiplied by exp(sigma z), z standard normal (Box-Muller on the
 * seeded stream, two draws per zone), at constant e and Ye.  prims holds six
 * variables (rho, u, v1, v2, v3, Ye) over the grid's zones. */
int mds_init_polytrope(const mds_grid *g, double n_index, double rho_c, double K, double rho_atm, double gam,
                       const mds_eos_table *tab, unsigned long long eos_seed, double sigma,
                       unsigned long long seed, double *prims);

#ifdef __cplusplus
}
#endif
#endif
