/* mds_mp.h -- grid tables of the multipole gravity (gravity.c:499-1110),
 * shared by the host builder (C) and the device path (HIP).  Everything here
 * depends on the grid and lmax only; the host evaluates it once with the
 * same libm calls and operation order as the CPU restatement. */
#ifndef MDS_MP_H
#define MDS_MP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDS_MP_LMAX 13          /* multipole_lmax + 1 < MULTIPOLE_LMAX (15, decs.h:38) */
#define MDS_MP_LM(l, m) ((l) * ((l) + 1) + (m))

typedef struct mds_mp_tab {
  int lmax, mmax, nlm;          /* mmax = lmax in 3D, 0 in 2D; nlm = (lmax + 1)^2 */
  int n1, n2, n3, ndim;
  int maxjlev, maxklev;         /* 0 on uniform grids */
  /* zone integrals of Y[l,m] sin(th) per theta level lj, phi level lk and
   * level-local (j, k): S[(Soff[lj * (maxklev + 1) + lk] + j * (n3 >> lk) + k) * nlm + lm] */
  double *S;
  long long *Soff;
  long long nS;                 /* doubles in S */
  /* harmonics at the corner angles J = jj * DJS, K = kk * DKS (finest units,
   * J in [0, n2], K in [0, n3]): Y[(J * (n3 + 1) + K) * nlm + lm] */
  double *Y;
  /* per shell i < n1 and l: the radial weights of the inner and outer
   * moments, rc = int r^(l+2) dr and rd = int r^(1-l) dr over the shell */
  double *rc, *rd;              /* [i * (lmax + 1) + l] */
  /* per corner radius ii <= n1 and l: -4 pi G / (2l + 1) / r^(l+1) and
   * -4 pi G / (2l + 1) r^l, and r itself (a corner at r < 1e-16 has no potential) */
  double *Cf, *Df;              /* [ii * (lmax + 1) + l] */
  double *rm;                   /* [ii] */
  /* centroid weights of the corner differences: radial per shell, theta per
   * level and row, phi per level and column: w[2 q + {0: lower, 1: upper}] */
  double *wx;                   /* [2 i] */
  double *wy;                   /* [2 (lj * n2 + j)] */
  double *wz;                   /* [2 (lk * n3 + k)] */
} mds_mp_tab;

/* Build the tables of grid g (2D or 3D spherical) for lmax in [1, MDS_MP_LMAX];
 * returns MDS_OK, MDS_EINVAL or MDS_ENOMEM.  Release with mds_mp_tab_free. */
int mds_mp_tab_build(const struct mds_grid *g, int lmax, mds_mp_tab *t);
void mds_mp_tab_free(mds_mp_tab *t);

#ifdef __cplusplus
}
#endif
#endif
