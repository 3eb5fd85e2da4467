/*
 * mds_eos.h -- tabulated equation of state of the hydro sweep.
 *
 * The table follows the layout of the reference's Burrows table
 * (eos/burrows/eos_stuff.c:7, TAB_ELEM: record, Ye, log rho, log T with T
 * fastest) and its use (eos_given_rex_tianshu, eos_stuff.c:507-650):
 * bilinear weights in (log10 rho, Ye), the temperature bracket of the
 * specific internal energy along log T, the exact inverse of the piecewise
 * linear e(log T), then P, Gamma_1 and cs = sqrt(Gamma_1 P / rho) at the
 * recovered temperature.  Values are stored in FP32; the inversion runs in
 * FP64.  The reference walks the bracket from the previous temperature; the
 * device finds the same bracket by bisection over the T nodes (e is strictly
 * increasing in T), so no temperature state is carried.  Off-table (rho, Ye)
 * are clamped to the table edges (the reference reads outside the table).
 *
 * The benchmark table is synthetic: a seeded smooth analytic
 * electron-degeneracy + thermal electrons + ideal ions + radiation model,
 * evaluated in extended precision at the nodes, in cgs and K.  A real table
 * in the Burrows binary layout loads with mds_eos_table_load_burrows (the
 * table files themselves are an external download).
 */
#ifndef MDS_EOS_H
#define MDS_EOS_H

#ifdef __cplusplus
extern "C" {
#endif

/* records of the table (reference record numbers in brackets) */
#define MDS_EOS_E       0   /* specific internal energy in table units, linear (record 1) */
#define MDS_EOS_LNP     1   /* ln P in table units (record 2, log-transformed at load) */
#define MDS_EOS_GAM1    2   /* adiabatic index Gamma_1 (record 12) */
#define MDS_EOS_NFIELD  3

/* Table units.  A state's specific internal energy e = u / rho [erg/g] is
 * looked up as e / e_scale - e_shift, and the pressure is exp(ln P) *
 * p_scale [dyn/cm^2]; log10 T is in the table's own temperature unit.  The
 * synthetic table is in cgs and K (scales 1, shift 0); the Burrows table is
 * in MeV per baryon (e_scale 0.95655684e18 erg/g, e_shift const_hack),
 * MeV/fm^3 (p_scale 1.60217733e33) and MeV (eos_stuff.c:530-532, 618).  A
 * scale of 0 reads as 1. */
typedef struct mds_eos_table {
  int nr, nt, ny;                        /* nodes in log10 rho, log10 T, Ye */
  double lr0, lr1;                       /* first and last log10 rho node [g/cm^3] */
  double lt0, lt1;                       /* first and last log10 T node (table unit) */
  double ye0, ye1;                       /* first and last Ye node */
  float *data;                           /* MDS_EOS_NFIELD * ny * nr * nt, T fastest */
  double e_scale, e_shift, p_scale;      /* table units (above) */
} mds_eos_table;

/* The reference's Burrows table (eos_stuff.c:8-21, DEFINE_EOSCONST_1): 300
 * log10 rho nodes over [1, 15.5], 300 log10 T [MeV] nodes over [-2, 2], 50
 * Ye nodes over [0.035, 0.56]; MeV units with const_hack = 0 (:375). */
#define MDS_BURROWS_NREC 16
void mds_eos_burrows_default(mds_eos_table *t);

/* Load a table file in the reference's Burrows binary layout
 * (collapse_init, eos_stuff.c:76-133): MDS_BURROWS_NREC records of
 * ny*nr*nt doubles each, record irec (1-based) at byte offset
 * (irec-1)*8*ny*nr*nt, T fastest within a record (TAB_ELEM, :7).  Records 1
 * (energy per baryon), 2 (pressure, log-transformed here as the reference
 * does at load) and 12 (Gamma_1) fill t->data in FP32 (allocated here).
 * Dimensions, node ranges and units come from *t as set by the caller
 * (mds_eos_burrows_default, or other dimensions of the same layout).
 * Returns MDS_EIO when the file is missing or short. */
int mds_eos_table_load_burrows(const char *path, mds_eos_table *t);

/* The table at a given temperature (host): eos_given_rtx
 * (eos/burrows/eos_stuff.c:1020-1044) through eos(eos_input_rt)
 * (:1189-1215) and findthis (:253-313).  The records are trilinear in
 * (log10 rho, log10 T, Ye) with the reference's weights and summation
 * order; temp is in the table's temperature unit.  out = {u = rho e
 * [erg/cm^3] (u_given_rtx, eos_collapse.c:172-181), P [dyn/cm^2], Gamma_1,
 * cs = sqrt(Gamma_1 P / rho)}.  rho, temp and Ye outside the table are
 * clamped to its edges (the reference reads beyond its arrays there).
 * Returns MDS_EINVAL for a non-positive rho or temp, or a table with fewer
 * than two nodes along an axis. */
int mds_eos_rtx(const mds_eos_table *t, double rho, double temp, double ye, double out[4]);

/* Default synthetic table dimensions: 180 x 180 x 50 over log10 rho in
 * [3,15], log10 T in [9,12], Ye in [0.05,0.55]. */
void mds_eos_table_default(mds_eos_table *t);

/* Model parameters drawn from the seeded stream (uniform jitter k = 0,1,2):
 * out = {A (mean ion mass number), K53 (degenerate-electron constant, cgs),
 *        rho_x (relativistic transition of rho*Ye, g/cm^3)}. */
void mds_eos_model_params(unsigned long long seed, double out[3]);

/* Evaluate the analytic model in extended precision at one point:
 * out = {P, e, Gamma_1}. */
void mds_eos_model_eval(const double prm[3], double rho, double temp, double ye, double out[3]);

/* Fill t->data (allocated here; dimensions and ranges must be set) from the
 * model with parameters mds_eos_model_params(seed).  Host threads. */
int mds_eos_table_build(mds_eos_table *t, unsigned long long seed);
void mds_eos_table_free(mds_eos_table *t);

/* Pointwise device EOS (parity entry point; replaces eos_fill /
 * eos_given_rex_per_zone, eos_stuff.c:768-901): for each of n states
 * (rho, u = rho*e, Ye) -> P, cs, T (table unit), Gamma_1.  precision: MDS_FP64 (all
 * FP64 after the FP32 table reads) or MDS_FP32_BULK (FP64 inversion, FP32
 * interpolation of ln P increments and Gamma_1).  Any output may be NULL. */
int mds_eos_points(int precision, const mds_eos_table *t, int n, const double *rho, const double *u,
                   const double *ye, double *press, double *cs, double *temp, double *gam1);

#ifdef __cplusplus
}
#endif
#endif
