/*
 * mds_hydro.h -- C-ABI of the MI355X explicit hydrodynamics sweep.
 *
 * One context (mds_ctx) owns the device state of one or more independent
 * problems (a batch) and advances all of them with the reference's SSP-RK2
 * step: boundaries -> EOS -> sources -> fluxes -> push, twice per step
 * (step.c:28-129), with the reference's dt policy (step.c:163-203,
 * main.c:228-252) evaluated on the device.
 *
 * Conventions
 *   - every call returns 0 on success, a negative MDS_E* code on failure;
 *     mds_last_error() gives a message; the library never exits the process.
 *   - primitive records follow the reference (decs.h:641-646):
 *       0 rho, 1 u (internal energy density), 2..4 contravariant coordinate
 *       velocities v^1..v^3 (io.c:176-186 applies scale factors on output only)
 *   - host arrays are SoA over interior zones, [nvars][n3][n2][n1], i fastest.
 *   - one HIP stream per context; mds_run() is asynchronous and performs no
 *     host synchronisation (t, dt and the step counter live in device memory),
 *     so it can be enqueued repeatedly or captured.
 *   - no torch types cross this boundary: plain pointers and sizes only.
 *
 * Reference interface replaced by each entry point is cited per declaration.
 */
#ifndef MDS_HYDRO_H
#define MDS_HYDRO_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDS_NG 2                 /* ghost depth of the reconstruction stencil, decs.h:61 */

/* precision of the device path */
#define MDS_FP64       0         /* FP64 end to end, reference operation order */
#define MDS_FP32_BULK  1         /* FP32 state/reconstruction/Riemann, FP64 update and geometry */

/* boundary codes: same numbers and meaning as decs.h:64-72 */
#define MDS_BC_OUTFLOW   0
#define MDS_BC_REFLECT   1
#define MDS_BC_EXTRAP    2 /* direction 0 only: linear extrapolation from the two edge zones
                              (bc.c:806-821, 937-954; the reference fills no other direction) */
#define MDS_BC_PERIODIC  3
#define MDS_BC_PROB      4 /* ghost values of the problem, set with mds_set_prob_bounds (bc.c:125-140) */
#define MDS_BC_SPHERICAL_ORIGIN 5 /* spherical grids: theta boundary across the pole (3D), or inner
                                     radius through the origin (1D, 3D; theta -> pi - theta,
                                     phi + pi, v^1 and v^3 flipped) */
#define MDS_BC_DISK      8 /* inner radius (direction 0, lower side) only: linear extrapolation from
                              zones 0 and 1 as EXTRAP, with v^1 clamped to 0 where it would point
                              out of the boundary, inflow only (bc.c:863-878); 1D columns and
                              2D/3D Cartesian grids */

#define MDS_GEOM_CARTESIAN 0
#define MDS_GEOM_SPHERICAL 1     /* 1D radial columns or 3D (r, theta, phi) grids */

#define MDS_RCOORD_UNIFORM 0     /* r = |x|     (metric_spherical_general.c:149-152) */
#define MDS_RCOORD_LOG     1     /* r = exp(x)  (metrics/metric_spherical_logr.c:3-38) */
#define MDS_RCOORD_SINH    2     /* r = rtrans sinh(|x|/rtrans) (metric_spherical_general.c:155-161, 186-208) */

/* error codes */
#define MDS_OK          0
#define MDS_EINVAL     -1
#define MDS_ENOMEM     -2
#define MDS_EDEVICE    -3
#define MDS_ENOTIMPL   -4
#define MDS_EIO        -5
#define MDS_ERANGE     -6   /* a size beyond an entry point's fixed maximum */

/* Per-zone geometry record of a 1D spherical column (one record per zone,
 * n1 + 2*MDS_NG records, zone -NG first), restating the fields of zone_geom
 * (decs.h:144-153) that the 1D sweep reads. */
enum {
  MDS_ZG_VOLUME = 0,       /* geom.volume                         */
  MDS_ZG_AREA0,            /* geom.area[0] (lower radial face)    */
  MDS_ZG_GCOV00,           /* geom.gcov[0][0..2] (zone average)   */
  MDS_ZG_GCOV01,
  MDS_ZG_GCOV02,
  MDS_ZG_GCON0,            /* geom.gcon[0..2]                     */
  MDS_ZG_GCON1,
  MDS_ZG_GCON2,
  MDS_ZG_GCOVF0,           /* geom.gcov[1][0..2] (lower face)     */
  MDS_ZG_GCOVF1,
  MDS_ZG_GCOVF2,
  MDS_ZG_SCALEF,           /* geom.scale[1][0]                    */
  MDS_ZG_SCALE0,           /* geom.scale[0][0]                    */
  MDS_ZG_C000,             /* non-zero connection coefficients    */
  MDS_ZG_C011,             /*   conn[l][m][n] -> C<l><m><n>        */
  MDS_ZG_C101,
  MDS_ZG_C110,
  MDS_ZG_C111,
  MDS_ZG_C202,
  MDS_ZG_C220,
  MDS_ZG_C212,
  MDS_ZG_C221,
  MDS_ZG_C022,
  MDS_ZG_C122,
  MDS_ZG_COUNT
};

/* Geometry record of a 3D spherical (r, theta, phi) grid, one per (i, j);
 * 2D (r, theta) grids use it as a 3D grid of one phi zone spanning 2 pi
 * with zero phi-face areas
 * (the metric does not depend on phi): (n1 + 2*NG) * (n2 + 2*NG) records,
 * i fastest, zone (-NG, -NG) first; restating zone_geom (decs.h:144-153) of
 * metric_spherical_general.c:708-1051 on a uniform angular grid. */
enum {
  MDS_ZS_VOLUME = 0,
  MDS_ZS_AREA0, MDS_ZS_AREA1, MDS_ZS_AREA2,        /* area[0..2] of the lower faces   */
  MDS_ZS_GCOV0,                                    /* gcov[0][0..2] (zone average)    */
  MDS_ZS_GCOVF0 = MDS_ZS_GCOV0 + 3,                /* gcov[1][0..2] (radial face)     */
  MDS_ZS_GCOVF1 = MDS_ZS_GCOVF0 + 3,               /* gcov[2][0..2] (theta face)      */
  MDS_ZS_GCON = MDS_ZS_GCOVF1 + 3,                 /* gcon[0..2]                      */
  MDS_ZS_SCALE0 = MDS_ZS_GCON + 3,                 /* scale[0][0..2] (= scale[3][..]) */
  MDS_ZS_SCALEF0 = MDS_ZS_SCALE0 + 3,              /* scale[1][0]                     */
  MDS_ZS_SCALEF1,                                  /* scale[2][1]                     */
  MDS_ZS_C000, MDS_ZS_C011, MDS_ZS_C101, MDS_ZS_C110, MDS_ZS_C202, MDS_ZS_C220,
  MDS_ZS_C212, MDS_ZS_C221, MDS_ZS_C022, MDS_ZS_C122, MDS_ZS_C111,
  MDS_ZS_STH, MDS_ZS_CTH,                          /* sin, cos of theta at the theta centroid */
  MDS_ZS_COUNT
};

/* Reconstruction coefficients of one zone for the FP32-bulk path, in a
 * well-conditioned local form of para_recon (reconstruct.c:127-213):
 *   first fit   al = a + CLM*dqm + CLP*dqp,  ar = a + CRM*dqm + CRP*dqp
 *   refit       b2 = (a - al - (ar-al)*MU1) * KINV,  b1 = (ar - al) - b2,
 *               interior extremum at sigma* = -b1/(2 b2) in (0,1)
 *   relocation  al' = ar + (a-ar)*INUR  (extremum to right face)
 *               ar' = al + (a-al)*IMU2  (extremum to left face)
 * They are evaluated in extended precision from the same rounded volume
 * moments the FP64 path uses. */
enum { MDS_RC_CLM = 0, MDS_RC_CLP, MDS_RC_CRM, MDS_RC_CRP, MDS_RC_MU1, MDS_RC_KINV,
       MDS_RC_INUR, MDS_RC_IMU2, MDS_RC_COUNT };

/* Grid and geometry of one problem.  Built by mds_grid_build() (which follows
 * the reference's init_coords/init_interp_vol/init_volume/init_area/init_gcov/
 * init_gcon/init_scale/init_conn, geometry.c:4-62) or supplied by the caller. */
typedef struct mds_grid {
  int ndim;                /* 1..3 */
  int n[3];                /* interior zones (1 for inactive directions) */
  int geom;                /* MDS_GEOM_* */
  int rcoord;              /* MDS_RCOORD_* (spherical) */
  double startx[3];
  double dx[3];
  /* volume moments per direction: alpha = int x^2 dV, beta = int x dV,
   * Gamma = int dV (per unit transverse measure), length n[d] + 2*NG,
   * zone -NG first (reference alpha0/beta0/Gamma0, alpha1s..., decs.h:360-364) */
  double *alpha[3], *beta[3], *Gamma[3];
  /* FP32-bulk reconstruction coefficients per direction: (n[d]+2*NG) * MDS_RC_COUNT */
  double *rcoef[3];
  /* Cartesian: uniform volume and areas (metric_cartesian.c:134-187) */
  double volume;
  double area[3];
  /* spherical: 1D (n[0]+2*NG) * MDS_ZG_COUNT per-zone records, 3D
   * (n[0]+2*NG) * (n[1]+2*NG) * MDS_ZS_COUNT per-(i,j) records; else NULL.
   * 3D spherical grids span theta over x1 in [-1,1] (startx[1] = -1,
   * dx[1] = 2/n2, POLYTH map) and phi over [0, 2 pi). */
  double *zone;
  /* input: scale of MDS_RCOORD_SINH */
  double rtrans;
  /* input: 3D spherical dendritic grid (utils.c:181-590): theta zones
   * coarsened by dj(i) = n2/nj(i) near the origin, phi zones by
   * dk(i,j) = n3/nk(i,j) near the poles, 2:1 between neighbours.  Zone
   * (i, j, k) then uses level-local j < nj(i) and k < nk(i, j); per-(i, j)
   * records (zone) and state arrays keep the uniform grid's (i, j, k)
   * positions, and positions beyond a level's count are not zones.  Needs
   * n2 and n3 powers of two (n3 >= 4), theta over [-1, 1]. */
  int dendritic;
  /* built for dendritic grids, else NULL / 0: */
  int *dj;                 /* n[0] + 2*NG, zone -NG first */
  int *dk;                 /* (n[0] + 2*NG) x (n[1] + 2*NG), [i + NG][j + NG], ghosts included */
  int maxjlev, maxklev;    /* finest level l has dj or dk = 2^l */
  /* theta (0) and phi (1) moments and FP32-bulk coefficients per level l
   * (dj or dk = 2^l), level l at offset l * (n + 2*NG) (times MDS_RC_COUNT
   * for lrcoef); level 0 equals alpha/beta/Gamma/rcoef[1 or 2] */
  double *lalpha[2], *lbeta[2], *lGamma[2], *lrcoef[2];
} mds_grid;

/* Build the reference geometry for a grid.  geom/rcoord/ndim/n/startx/dx
 * must be set on input; allocates the arrays (release with mds_grid_free).
 * 2D spherical (r, theta) grids: startx[2] and dx[2] are set to 0 and 2 pi
 * on return (one phi zone spanning the circle, as the reference's
 * init_coords sets the absent direction); a phi wedge is not expressible. */
int mds_grid_build(mds_grid *g);
void mds_grid_free(mds_grid *g);
/* Build count independent grids (host threads); returns the lowest error code. */
int mds_grid_build_batch(mds_grid *g, int count);

/* monopole gravity (mds_problem.monopole) */
#define MDS_MONOPOLE_NEWTON 1
#define MDS_MONOPOLE_GR     2

/* Physics and boundary description of one problem. */
typedef struct mds_problem {
  const mds_grid *grid;    /* copied into the context at create */
  int bc_lo[3], bc_hi[3];  /* MDS_BC_* per direction (all variables) */
  double gam;              /* gamma-law index (eos_gamma.c:5-12) */
  double g0[3];            /* fixed gravity, contravariant (gravity.c:46-64); zeros = none */
  int nvel;                /* velocity components carried; 0 = auto: ndim, except FP64
                              spherical columns, which carry 3 (the connection terms give
                              the transverse components round-off sources); FP32-bulk
                              spherical columns carry the radial component by default, or
                              all three with nvel = 3 (rotating columns).  3D grids
                              and 2D/3D spherical grids always carry 3: any other explicit
                              value is rejected with MDS_EINVAL */
  const struct mds_eos_table *eos;  /* tabulated EOS (mds_eos.h), NULL = gamma law; with a table
                              the state carries Ye as a sixth variable (decs.h:646).
                              Spherical grids (1D columns, 2D and 3D); every problem of a
                              batch shares it */
  int monopole;            /* monopole self-gravity, spherical grids (1D columns, 2D and 3D):
                              0 none; MDS_MONOPOLE_NEWTON (gravity.c:145-498);
                              MDS_MONOPOLE_GR, the GR effective monopole of the
                              reference's GR_MONOPOLE build (gravity.c:142-375: TOV
                              mass and factor, pressure and inertia terms, lapses) */
  int multipole_lmax;      /* multipole gravity (SPHERICAL_MULTIPOLE_GRAV, gravity.c:499-1110):
                              the l = 1..multipole_lmax terms of the potential on top
                              of the monopole (needs monopole != 0), at most
                              MDS_MULTIPOLE_LMAX; 2D and 3D spherical grids (2D: m = 0
                              only).  0 = none; the reference's default is 12 */
} mds_problem;
#define MDS_MULTIPOLE_LMAX 13

/* Runtime parameters with the reference's names and defaults (io.c:5018-5041). */
typedef struct mds_run_params {
  double cfl;              /* 0.9 */
  double dt_init_safety;   /* 0.1 */
  double dt_grow_safety;   /* 1.1 */
  double dtmin;            /* 1e-200 */
  double dtmax;            /* 1e200 */
  double initial_dt;       /* 1e100 */
  double tmax;             /* problems freeze once t >= tmax */
  int hydro_interp_order;  /* 3: para_recon (the default, io.c:5054); 2: lin_recon; 1: piecewise
                              constant (reconstruct.c:246-266).  FP32-bulk runs order 3 only */
} mds_run_params;

typedef struct mds_ctx mds_ctx;

/* Default run parameters (io.c:5018-5041). */
void mds_default_run_params(mds_run_params *rp);

/* Create a context holding nprob problems.  Replaces the reference's
 * allocate_memory + init_geometry + device residency (main.c:78-165,
 * utils.c:1500-1630).  precision: MDS_FP64 or MDS_FP32_BULK. */
int mds_create(const mds_problem *probs, int nprob, const mds_run_params *rp,
               int precision, mds_ctx **out);
int mds_destroy(mds_ctx *ctx);
const char *mds_last_error(const mds_ctx *ctx);

/* Host <-> device copies of the primitive state (reference: omp target
 * update to/from, main.c:180, step.c:16-24).  prims: nvars*n3*n2*n1 doubles. */
int mds_upload_prims(mds_ctx *ctx, int prob, const double *prims, int nvars);
int mds_download_prims(mds_ctx *ctx, int prob, double *prims, int nvars);

/* The ghost values of the PROB sides of problem prob: fn(i, j, k, p, user) is
 * called once for every ghost zone (i, j, k) of every PROB side -- the
 * reference's per-problem hook prob_bounds(i, j, k, p) (problems/init_*.c,
 * e.g. problems/init_rt_3d.c:223-236), which its fill calls per ghost zone
 * (bc.c:125-140 and the like) -- and p receives rho, u, v1, v2, v3 and, with a
 * table EOS, Ye.  (i, j, k): the ghost index along the side's direction
 * (-2, -1 below the grid, n, n + 1 above), interior indices along the others
 * (on a dendritic grid, positions of the ghost shell's own level; 0 for
 * absent directions).  The values stay on the device for every later stage
 * (time-independent; call again to change them); mds_run refuses a problem
 * with a PROB side that was never set.  PROB is accepted on every side of
 * Cartesian grids and 1D spherical columns and at the outer radius of 3D
 * spherical grids; EXTRAP on both sides of direction 0, except the inner
 * radius of 3D spherical grids. */
typedef void (*mds_prob_bounds_fn)(int i, int j, int k, double *p, void *user);
int mds_set_prob_bounds(mds_ctx *ctx, int prob, mds_prob_bounds_fn fn, void *user);

/* A PROB hook that reads the global time, as the reference's Noh problems do
 * (problems/init_spherical_noh.c prob_bounds: rho = (1 + t/r)^2): fn(i, j, k,
 * t, p, user) fills the same ghost zones as above at the time the
 * reference's fill sees -- t in the first RK stage and t + dt in the second
 * (bc.c:717-721).  The library keeps the hook: every step of mds_run and
 * mds_run_graph then computes dt on the device, waits for it, evaluates the
 * hook on the host at both times and copies the values in before each stage,
 * so such a context steps synchronously (one host round trip per step). */
typedef void (*mds_prob_bounds_t_fn)(int i, int j, int k, double t, double *p, void *user);
int mds_set_prob_bounds_t(mds_ctx *ctx, int prob, mds_prob_bounds_t_fn fn, void *user);

/* Initial EOS fill and first dt estimate input (main.c:181, eos.c:4-18);
 * also resets t, the step counter and the dt history of every problem. */
int mds_update_eos(mds_ctx *ctx);

/* Enqueue nsteps iterations of the reference main loop body for every active
 * problem: estimate_dt (step.c:163-203) + dt policy (main.c:228-252) +
 * step() (step.c:28-129).  Asynchronous; no host synchronisation. */
int mds_run(mds_ctx *ctx, int nsteps);
/* Same, replayed from a captured hipGraph of one step (captured on first use). */
int mds_run_graph(mds_ctx *ctx, int nsteps);
int mds_sync(mds_ctx *ctx);

/* Launch-shape parameters of the sweep kernels (no counterpart in the
 * reference, whose OpenMP runtime picks launch shapes).  Chosen by
 * mds_autotune() or set explicitly; every value gives identical results. */
#define MDS_TILE_ROWS2D  0       /* rows marched by one wavefront in the 2D sweep */
#define MDS_TILE_TY3D    1       /* output rows per workgroup in the 3D sweep (4, 8 or 12; 4 and
                                    halo mode 0 when a problem has PROB/EXTRAP sides or
                                    hydro_interp_order < 3) */
#define MDS_TILE_KC3D    2       /* planes marched by one workgroup in the 3D sweep */
#define MDS_TILE_FUSE1D  3       /* 1D: 1 = both RK stages of a step in one launch (default
                                    when no problem is periodic in x), 0 = one launch per stage,
                                    2 = both stages in one launch with two zones per lane
                                    (FP32-bulk only) */
#define MDS_TILE_GRID1D  4       /* 1D grid: 0 = one wavefront per segment of the segment table,
                                    1 = persistent (the workgroups the device holds at once walk
                                    the whole table) */
#define MDS_TILE_HALO3D  5       /* 3D: 0 = the four halo rows of a tile on wavefronts of their own,
                                    1 = loaded by the output wavefronts beside their rows */
#define MDS_TILE_UPD3S   6       /* 2D/3D spherical update workgroup: 0 = 256 zones in storage
                                    order, 1 = 64 radial zones x 4 theta rows, 2 = 64 x 2 rows x
                                    2 phi positions, 3 = 64 x 4 phi positions (default) */
int mds_set_tile(mds_ctx *ctx, int param, int value);
int mds_get_tile(mds_ctx *ctx, int param, int *value);

/* Time each candidate launch shape on the context's current state for
 * nsteps steps and keep the fastest; the state, clocks and counters are
 * restored afterwards, so the simulation is unaffected.  best_ms (optional)
 * receives the best time per step. */
int mds_autotune(mds_ctx *ctx, int nsteps, double *best_ms);

/* Measurement variant of mds_run (no counterpart in the reference, which
 * times with host wall-clock timers, timer.c:34-48): enqueues nsteps steps
 * eagerly with HIP events between the kernels on the context stream and
 * waits.  ms[0] = elapsed time of the whole region, ms[1] = summed duration
 * of the first-stage sweep kernels, ms[2] = of the second-stage sweep kernels. */
int mds_run_timed(mds_ctx *ctx, int nsteps, double ms[3]);

/* Device copy kernel over nbytes (rounded down to 16 B) repeated reps times;
 * returns the mean duration in ms.  Used to calibrate HBM counters and to
 * measure the achievable copy bandwidth on the device. */
int mds_copy_probe(long long nbytes, int reps, double *ms);

/* VALU issue rate of the device (no counterpart in the reference): eight
 * independent chains per lane over a grid of 16 workgroups per compute unit,
 * reps launches; returns wavefront instructions per second.  kind 0: FP32
 * fused multiply-add, 1: FP64 fused multiply-add (the capacity against which
 * the stage kernels' SQ_INSTS_VALU counts are read), 2/3: FP32/FP64
 * reciprocal, 4/5: FP32/FP64 square root (the cost of the transcendental
 * instructions beside the FMA). */
int mds_valu_probe(int kind, int reps, double *wave_instr_per_s);

/* Clock of a problem: time, last dt, steps taken (reference globals t, dt, istep). */
int mds_get_clock(mds_ctx *ctx, int prob, double *t, double *dt, int *istep);

/* Compensated (Neumaier) FP64 totals of rho*V and E*V over interior zones
 * (reference calculate_total_mass/energy, utils.c:2337-2366, 2581-2613). */
int mds_conserved_totals(mds_ctx *ctx, int prob, double out[2]);

/* Device health of one problem, read with its clock.  The reference has no
 * error channel on this path: it prints EOS off-table warnings and goes on
 * (eos/burrows/eos_stuff.c:571-604) and exits when dt < dtmin (main.c:229-238).
 * Here the stage kernels count, since mds_update_eos:
 *   bad_state      zone updates (one per zone and RK stage) whose new state
 *                  has a non-finite or non-positive density or a non-finite
 *                  internal energy -- or, with the gamma law, a negative one
 *                  (negative pressure).  A NaN state makes no CFL limit (its
 *                  key loses every minimum), so this is where it shows;
 *   eos_off_table  table EOS evaluations of the stage's zones with density,
 *                  Ye or temperature outside the table (clamped to its edge
 *                  in density and Ye, extrapolated in temperature);
 *   dt_below_min   1 once dt fell below dtmin (the run froze there). */
typedef struct {
  double t, dt;
  int istep;
  int dt_below_min;
  long long bad_state;
  long long eos_off_table;
} mds_health;

int mds_get_health(mds_ctx *ctx, int prob, mds_health *out);

/* State of the GR effective monopole of problem prob (MDS_MONOPOLE_GR) as
 * the last stage left it, the reference's gtov, mr, gr_lapse and
 * gr_lapse_edge (gravity.c:142-375; geom[].lapse[0..1], dumped by
 * io.c:198): n1 + 1 face values of the TOV factor Gamma, of the TOV mass and
 * of the lapse, and n1 zone-centred lapses (any pointer may be NULL).
 * passes: passes of the last call's iteration; nonconverged: calls whose 21
 * passes did not converge (the reference exits there, gravity.c:320-332). */
int mds_get_gr_state(mds_ctx *ctx, int prob, double *gtov, double *mr, double *lapse_edge,
                     double *lapse, int *passes, int *nonconverged);

/* Diagnostic counters since mds_update_eos: [0] radial faces tagged as shocks
 * by HLLC, [1] transverse faces solved with HLLE. */
int mds_counters(mds_ctx *ctx, long long out[2]);

/* The HIP stream (hipStream_t) all work of the context is enqueued on. */
void *mds_stream(mds_ctx *ctx);

/* Pointwise parity entry points: run the device reconstruction / Riemann
 * solvers on host arrays (copied in and out).  Replace para_recon
 * (reconstruct.c:127-213) and riemann_flux_HLLC/HLLE (riemann.c:178-499).
 *   recon: ncell pencils of 3 values a[3*c+q], moments mom[9*c + 3*q + {0:alpha,1:beta,2:Gamma}],
 *          faces x[2*c+{0,1}]  ->  al[c], ar[c]
 *   riemann: states pl/pr[7*f + {rho,u,v1,v2,v3,P,Gamma}], face geometry
 *          fg[5*f + {area, scale, gcov0, gcov1, gcov2}], dir  ->  flux[5*f+..], sig[f], isshock[f] */
int mds_recon_points(int precision, int ncell, const double *a, const double *mom,
                     const double *x, const double *rcoef, double *al, double *ar);
int mds_riemann_points(int precision, int solver /*0 HLLC, 1 HLLE*/, int nface,
                       const double *pl, const double *pr, const double *fg, int dir,
                       double *flux, double *sig, int *isshock);

/* Allocation-free variant of mds_riemann_points (no counterpart in the
 * reference): for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results; the device staging of the inputs and outputs lives in a fixed
 * static device array of the library (about 6.7 MB per GPU, held until the
 * process ends) instead of per-call allocations, and the call makes no
 * hipMalloc/hipFree.  nface > MDS_NOALLOC_MAX_FACES is rejected with
 * MDS_ERANGE.  Runs on the current device, synchronously on its null stream.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several host threads) race and must be serialised by the caller. */
#define MDS_NOALLOC_MAX_FACES 32768
int mds_riemann_points_noalloc(int precision, int solver, int nface,
                               const double *pl, const double *pr, const double *fg, int dir,
                               double *flux, double *sig, int *isshock);

#ifdef __cplusplus
}
#endif
#endif
