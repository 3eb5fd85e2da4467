/*
 * oracle.h -- CPU oracle for the explicit hydrodynamics sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded, strict-IEEE
 * restatement of the reference algorithm (MEDUSA/Fornax: step.c, reconstruct.c,
 * riemann.c, fluid.c, rad_fluid.c, bc.c, gravity.c, eos/gamma/eos_gamma.c,
 * metrics/metric_cartesian.c, metrics/metric_spherical_general.c, utils.c).
 * It is used only by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py, always as the checker -- never as the product path.
 *
 * Every function cites the reference file:line it restates.  The arithmetic
 * follows the reference operation order so that a strict-IEEE build
 * (-O2 -fno-fast-math -ffp-contract=off) is a bit-level model of the
 * reference CPU build without -ffast-math.
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed in the header comment of each function and in tests/test_oracle_*.py.
 */
#ifndef MDS_ORACLE_H
#define MDS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NG 2          /* ghost zones, decs.h:61 */
#define ORC_NSIZE 20      /* pencil row capacity, decs.h:36 */

/* primitive mnemonics, decs.h:641-646 */
#define ORC_RHO 0
#define ORC_UU  1
#define ORC_U1  2
#define ORC_U2  3
#define ORC_U3  4
#define ORC_YE  5
/* conserved mnemonics, decs.h:650-654 */
#define ORC_ETOT 1

/* boundary codes, decs.h:64-72 */
#define ORC_BC_OUTFLOW  0
#define ORC_BC_REFLECT  1
#define ORC_BC_EXTRAP   2
#define ORC_BC_PERIODIC 3
#define ORC_BC_PROB     4   /* ghost values from the problem's prob_bounds (bc.c:125-140) */
#define ORC_BC_SPHERICAL_ORIGIN 5   /* theta: across the pole (phi + pi); inner radius: through the
                                       origin (theta -> pi - theta, phi + pi); decs.h:69 */
#define ORC_BC_DISK     8   /* inner radius only: EXTRAP with v1 clamped to <= 0 (bc.c:863-878) */

#define ORC_GEOM_CARTESIAN 0
#define ORC_GEOM_SPHERICAL 1

/* radial map r(x) for spherical grids */
#define ORC_RCOORD_UNIFORM 0   /* r = |x|       (metric_spherical_general.c:149-152) */
#define ORC_RCOORD_LOG     1   /* r = exp(x)    (metrics/metric_spherical_logr.c:3-38) */
#define ORC_RCOORD_SINH    2   /* r = rtrans sinh(|x|/rtrans) (metric_spherical_general.c:155-161) */

/* Geometry of one zone, restating zone_geom (decs.h:144-153) for NDIM<=3.
 * Index 0 of gcov/scale is the zone centre, 1+d the lower face in direction d. */
typedef struct {
  double area[3];
  double conn[3][3][3];
  double gcov[4][3];
  double gcon[3];
  double scale[4][3];
  double volume;
} orc_zone_geom;

typedef struct {
  int ndim;
  int n[3];
  int geom;            /* ORC_GEOM_* */
  int rcoord;          /* ORC_RCOORD_* (spherical only) */
  double startx[3];
  double dx[3];
  int bc_lo[3], bc_hi[3];
  int nvars;           /* 5 (+ncomp) */
  int interp_order;    /* hydro_interp_order, io.c:5054 (3 = para_recon) */
  double gam;          /* gamma-law index */
  double cfl;
  double dt_init_safety, dt_grow_safety, dtmin, dtmax, initial_dt;
  double g0[3];        /* fixed gravity (contravariant), gravity.c:46-64; all zero = none */
  int use_fixed_grav;
  int use_monopole;    /* 1: Newtonian monopole gravity, gravity.c:145-251, 432-498;
                          2: GR effective monopole (GR_MONOPOLE), gravity.c:142-251, 266-375 */
  const struct orc_eos_tab *eos_tab;   /* tabulated EOS (NULL: gamma law); nvars 6 carries Ye */
  double rtrans;       /* ORC_RCOORD_SINH scale */
  int dendritic;       /* 3D spherical: dendritic angular coarsening (utils.c:181-590) */
  /* PROB boundaries: the problem's prob_bounds(i, j, k, p) (e.g.
   * problems/init_rt_3d.c:223-236), called for every ghost zone of a PROB side
   * with p holding nvars entries */
  void (*prob_bounds)(int i, int j, int k, double *p, void *user);
  void *prob_user;
  /* or a hook that reads the global time, as problems/init_spherical_noh.c
   * does: called with the time the reference's fill sees, t in the first
   * stage and t + dt in the second (bc.c:717-721) */
  void (*prob_bounds_t)(int i, int j, int k, double t, double *p, void *user);
  /* multipole gravity (SPHERICAL_MULTIPOLE_GRAV, gravity.c:499-1110): the
   * l = 1..multipole_lmax terms of the potential on top of the monopole
   * (use_monopole must be set); 0 = none.  2D grids take m = 0 only */
  int multipole_lmax;
} orc_config;

typedef struct orc_sim orc_sim;

/* tabulated EOS table (layout of eos/burrows/eos_stuff.c:7: record, Ye,
 * rho, T with T fastest; records 0 e, 1 ln P, 2 Gamma_1), FP32 values */
typedef struct orc_eos_tab {
  int nr, nt, ny;
  double lr0, lr1, lt0, lt1, ye0, ye1;
  const float *data;
  /* table units (eos_stuff.c:530, 618): e / e_scale - e_shift is compared
   * with the energy record, P = exp(ln P) * p_scale; 0 reads as 1 */
  double e_scale, e_shift, p_scale;
} orc_eos_tab;

/* eos_stuff.c:507-650: out = {P, cs, T, Gamma_1}; returns off-table bits
 * (1 below, 2 above the energy range) */
int orc_eos_walk(const orc_eos_tab *t, double rho, double u, double ye, double temp_guess, double out[4]);
/* eos_stellar_collapse.c:294-369 on the same interpolant: log10 T of the root */
int orc_eos_root(const orc_eos_tab *t, double rho, double u, double ye, double lt_guess, double tol,
                 double *lt_out, int *newton_iters);

/* ---- multipole gravity pieces (gravity.c:499-697) ---- */
/* real spherical harmonics at (th, phi), LM index l (l + 1) + m, m in [-l, l] */
void orc_ylms(double *ylm, double th, double phi, int lmax, int mmax);
/* their integrals times sin(th) over th in [thmin, thmax], phi in [phimin, phimax] */
void orc_ylm_ints(double *slm, double thmin, double thmax, double phimin, double phimax, int lmax, int mmax);

/* ---- pointwise kernels (reconstruct.c, riemann.c, fluid.c) ---- */
void orc_para_recon(const double *a, double *al, double *ar,
                    const double *alpha, const double *beta, const double *Gamma,
                    const double *x, int istart, int istop);
void orc_lin_recon(const double *a, double *al, double *ar,
                   const double *beta, const double *Gamma,
                   const double *x, int istart, int istop);
/* face geometry needed by a Riemann solve: area[dir], gcov[1+dir][*], scale[1+dir][dir] */
int orc_riemann_hllc(int nhydro, int nvars, const double *pl, const double *pr,
                     const orc_zone_geom *g, int dir,
                     double *flux, double *sig_speed, double *vriemann);
int orc_riemann_hlle(int nhydro, int nvars, const double *pl, const double *pr,
                     const orc_zone_geom *g, int dir,
                     double *flux, double *sig_speed, double *vriemann);
void orc_p_to_u(const double *p, double *u, const double *gcov, int nhydro);
void orc_u_to_p(double *u, double *p, const double *gcon, int nhydro);
void orc_geom_src(const double *p, double press, double *src,
                  const orc_zone_geom *g, int nvars, int geom);
double orc_dt_source_control(const double *p, const double *u, const double *s);

/* ---- geometry (metric_cartesian.c, metric_spherical_general.c, utils.c) ---- */
double orc_romb(double x0, double x1, double (*f)(double, const void *), const void *ctx);

/* ---- simulation ---- */
orc_sim *orc_create(const orc_config *cfg);
void orc_destroy(orc_sim *s);
/* prims: SoA over interior zones, [nvars][n3][n2][n1] with i fastest */
void orc_set_prims(orc_sim *s, const double *prims);
void orc_get_prims(const orc_sim *s, double *prims);
/* geometry export for cross-checks: per zone incl. ghosts, natural order */
int orc_zone_count_all(const orc_sim *s);
void orc_fill_ghosts(orc_sim *s);                /* bc.c:705-1421 on the current state */
int orc_zone_prims(const orc_sim *s, int i, int j, int k, double *out);   /* any zone, ghosts included */
const orc_zone_geom *orc_geometry(const orc_sim *s);
/* volume moments per direction, length n[d]+2*NG (index 0 = zone -NG) */
const double *orc_moment(const orc_sim *s, int which /*0 alpha,1 beta,2 Gamma*/, int dir);
void orc_update_eos_all(orc_sim *s);           /* main.c:181 */
double orc_estimate_dt(orc_sim *s);             /* step.c:163-203 */
void orc_step(orc_sim *s);                      /* step.c:28-129 */
/* main.c:206-313 loop: runs until t>=tmax or max_steps; returns steps taken */
int orc_run(orc_sim *s, double tmax, int max_steps);
double orc_time(const orc_sim *s);
double orc_last_dt(const orc_sim *s);
int orc_istep(const orc_sim *s);
long orc_shock_face_count(const orc_sim *s);    /* radial faces HLLC tagged as shocks */
long orc_hlle_face_count(const orc_sim *s);     /* transverse faces solved with HLLE */
/* compensated (Neumaier) sums of conserved mass and total energy over interior zones */
void orc_conserved_totals(const orc_sim *s, double out[2]);
/* monopole gravity at the n1+1 radial faces from the current state (gravity.c:420-425) */
void orc_monopole_edges(const orc_sim *s, double *grav);
/* GR effective monopole state after the last call (use_monopole == 2): the
 * TOV factor and mass at the n1+1 faces, the lapses at the n1 zone centroids
 * and the n1+1 faces, the iteration count.  Returns 1 if the last call did
 * not converge in 20 passes, -1 without GR monopole.  Any output may be NULL. */
int orc_gr_state(const orc_sim *s, double *gtov, double *mr, double *lapse, double *lapse_edge, int *iter);
/* GR effective monopole face gravity (n1+1 values) of the current state, after
 * an EOS fill of it; updates the kept TOV factor like a stage does. */
int orc_gr_edges(orc_sim *s, double *grav);
/* multipole potential at the zone corners of the current state, after the
 * radial fluxes of a stage are known or not (the potential needs no flux):
 * phi[(ii * (n2 + 1) + jj) * (n3 + 1) + kk] for corner (ii, jj, kk) of
 * level-local indices, ii in [0, n1], jj in [0, nj(ii)], kk in [0, nk(ii,
 * jj)] (gravity.c:848-906); -1 without multipole gravity */
int orc_multipole_phi(orc_sim *s, double *phi);
/* tests only: the multipole momentum sources -rho dPhi/dx^d of the current
 * state with no mass fluxes (src[3 * zone + d], zones i fastest, box
 * positions; the energy source needs fluxes and is not returned) */
int orc_multipole_sources(orc_sim *s, double *src);
/* dendritic levels: nj[n1], nk[n1*n2] (0 beyond nj[i]); returns maxjlev*256 + maxklev */
int orc_dendritic_levels(const orc_sim *s, int *nj, int *nk);
/* theta (dir 0) / phi (dir 1) moments of refinement level lev (length n + 2 NG), NULL if none */
const double *orc_level_moment(const orc_sim *s, int which, int dir, int lev);
/* tests only: pencil entries of the dendritic sweeps and the polar phi-face
 * states on the current state (step.c:783-1085, 1299-1397); call
 * orc_prepare_entries first (ghost fill + EOS update) */
void orc_prepare_entries(orc_sim *s);
int orc_dend_radial_entry(orc_sim *s, int i, int j, int k, int ii, double *out);
int orc_dend_theta_entry(orc_sim *s, int i, int j, int kk, int jj, double *out);
int orc_polar_face_states(orc_sim *s, int i, int j, int k, double *left, double *right);

#ifdef __cplusplus
}
#endif
#endif
