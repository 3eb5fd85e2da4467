/*
 * mds_driver.c -- a C host program in the shape of the reference's main.c,
 * driving the device path through the C-ABI only.
 *
 * The reference's host calls (decs.h:754-756, 936-937) keep their names and
 * order here: init_grid() -> init_problem() -> update_eos (main.c:181) ->
 * loop { estimate_dt(); step(); } (main.c:206-264) -> update_cpu() ->
 * dump_ascii1() (io.c:109-230).  They act on file-scope globals named as in
 * defs.h (t, dt, istep, tmax, sim_p).  step() enqueues one step on the device
 * without synchronising; estimate_dt() refreshes the host copies of t, dt and
 * istep from the device clock, where the reference dt policy (step.c:163-203,
 * main.c:228-252) has already been applied.
 *
 * Usage:
 *   mds_driver sod  <n> <tmax> <dump>            1D Sod (problems/init_sod.c), FP64
 *   mds_driver star <n> <steps> <dump> [fp32]    polytropic star column with the
 *                                                synthetic table EOS and monopole gravity
 * The dump is the reference's 1D ASCII layout: "# t", then per zone
 * x (r for spherical grids), the primitives times scale (rho, u, v1, v2, v3,
 * [Ye]), the eos record (P, cs, T, entropy, Gamma), the enclosed mass at the
 * zone centre and the two lapses (1 1 without GR).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/mds_hydro.h"
#include "../../../include/mds_host.h"
#include "../../../include/mds_eos.h"

/* ---- reference globals (defs.h) ---- */
static double t, dt, tmax;
static int istep, max_steps;
static int nvars;
static double *sim_p;            /* SoA host copy [nvars][n1] */

/* ---- this program's state ---- */
static mds_grid grid;
static mds_problem prob;
static mds_ctx *ctx;
static mds_eos_table table;
static int use_table, precision = MDS_FP64;
static double gam;

static void die(const char *what, int rc)
{
  fprintf(stderr, "mds_driver: %s failed (%d): %s\n", what, rc, ctx ? mds_last_error(ctx) : "");
  exit(1);
}

/* init_grid (decs.h:936): geometry of the chosen problem */
static void init_grid(const char *kind, int n)
{
  memset(&grid, 0, sizeof(grid));
  grid.ndim = 1;
  grid.n[0] = n;
  grid.n[1] = grid.n[2] = 1;
  if (!strcmp(kind, "sod")) {
    grid.geom = MDS_GEOM_CARTESIAN;
    grid.startx[0] = 0.0;
    grid.dx[0] = 1.0 / n;
  } else {
    grid.geom = MDS_GEOM_SPHERICAL;
    grid.rcoord = MDS_RCOORD_LOG;
    grid.startx[0] = log(1.0e5);
    grid.dx[0] = (log(2.0e9) - log(1.0e5)) / n;
  }
  grid.dx[1] = grid.dx[2] = 1.0;
  int rc = mds_grid_build(&grid);
  if (rc) die("mds_grid_build", rc);
}

/* init_problem (decs.h:937): initial state, boundaries, EOS, gravity */
static void init_problem(const char *kind)
{
  memset(&prob, 0, sizeof(prob));
  prob.grid = &grid;
  const int n = grid.n[0];
  nvars = 6;
  sim_p = (double *)calloc((size_t)nvars * n, sizeof(double));
  int rc;
  if (!strcmp(kind, "sod")) {
    gam = 1.4;
    prob.bc_lo[0] = prob.bc_hi[0] = MDS_BC_OUTFLOW;
    nvars = 5;
    rc = mds_init_sod(&grid, gam, sim_p);
  } else {
    gam = 4.0 / 3.0;
    use_table = 1;
    mds_eos_table_default(&table);
    rc = mds_eos_table_build(&table, 0x5EED0005ull);
    if (rc) die("mds_eos_table_build", rc);
    prob.bc_lo[0] = MDS_BC_REFLECT;
    prob.bc_hi[0] = MDS_BC_OUTFLOW;
    prob.eos = &table;
    prob.monopole = 1;
    rc = mds_init_polytrope(&grid, 3.0, 1.0e10, 6.0e14, 1.0e4, gam, &table, 0x5EED0005ull, 0.01,
                            0x5EED0006ull, sim_p);
  }
  if (rc) die("problem init", rc);
  prob.gam = gam;
  mds_run_params rp;
  mds_default_run_params(&rp);
  rp.cfl = 0.5;
  rp.tmax = tmax;
  rc = mds_create(&prob, 1, &rp, precision, &ctx);
  if (rc) die("mds_create", rc);
  rc = mds_upload_prims(ctx, 0, sim_p, nvars);
  if (rc) die("mds_upload_prims", rc);
}

/* estimate_dt (step.c:163-203): the device folds it into each step; here
 * the host copies of t, dt and istep are brought up to date */
static void estimate_dt(void)
{
  int rc = mds_get_clock(ctx, 0, &t, &dt, &istep);
  if (rc) die("mds_get_clock", rc);
}

/* step (step.c:28-129): one SSP-RK2 step, asynchronous */
static void step(void)
{
  int rc = mds_run(ctx, 1);
  if (rc) die("mds_run", rc);
}

/* update_cpu (decs.h:754): device state -> sim_p */
static void update_cpu(void)
{
  int rc = mds_download_prims(ctx, 0, sim_p, nvars);
  if (rc) die("mds_download_prims", rc);
}

/* dump_ascii1 (io.c:109-230) */
static void dump_ascii1(const char *name)
{
  const int n = grid.n[0];
  double *press = calloc(n, sizeof(double)), *cs = calloc(n, sizeof(double));
  double *temp = calloc(n, sizeof(double)), *g1 = calloc(n, sizeof(double));
  if (use_table) {
    int rc = mds_eos_points(precision, &table, n, sim_p, sim_p + n, sim_p + 5 * (size_t)n, press, cs, temp, g1);
    if (rc) die("mds_eos_points", rc);
  } else {
    /* eos_gamma.c:45-57 */
    for (int i = 0; i < n; i++) {
      press[i] = (gam - 1.0) * sim_p[n + i];
      cs[i] = sqrt(gam * press[i] / sim_p[i]);
      temp[i] = press[i] / (2.0 * sim_p[i] * 1.3806488e-16) * (1.67262158e-24 + 9.10938291e-28);
      g1[i] = gam;
    }
  }
  FILE *fp = fopen(name, "w");
  if (!fp) { fprintf(stderr, "mds_driver: cannot open %s\n", name); exit(1); }
  fprintf(fp, "# %g\n", t);
  double mass = 0.0;
  for (int i = 0; i < n; i++) {
    const int q = i + MDS_NG;
    double x = grid.startx[0] + (i + 0.5) * grid.dx[0];
    double scale1 = 1.0, vol = grid.volume;
    if (grid.geom == MDS_GEOM_SPHERICAL) {
      x = exp(x);
      scale1 = grid.zone[(size_t)q * MDS_ZG_COUNT + MDS_ZG_SCALE0];
      vol = grid.zone[(size_t)q * MDS_ZG_COUNT + MDS_ZG_VOLUME];
    }
    fprintf(fp, "%16.14e ", x);
    for (int v = 0; v < nvars; v++) {
      double s = (v == 2) ? scale1 : 1.0;
      fprintf(fp, "%16.14e ", sim_p[(size_t)v * n + i] * s);
    }
    fprintf(fp, "%16.15e %16.15e %16.15e %16.15e %16.15e ", press[i], cs[i], temp[i], 0.0, g1[i]);
    const double dm = vol * sim_p[i];
    fprintf(fp, "%16.15e ", mass + 0.5 * dm);
    mass += dm;
    fprintf(fp, "%16.15e %16.15e \n", 1.0, 1.0);
  }
  fclose(fp);
  free(press); free(cs); free(temp); free(g1);
}

int main(int argc, char **argv)
{
  if (argc < 5) {
    fprintf(stderr, "usage: %s sod <n> <tmax> <dump> | star <n> <steps> <dump> [fp32]\n", argv[0]);
    return 2;
  }
  const char *kind = argv[1];
  const int n = atoi(argv[2]);
  if (!strcmp(kind, "sod")) {
    tmax = atof(argv[3]);
    max_steps = 1 << 30;
  } else if (!strcmp(kind, "star")) {
    tmax = 1.0e300;
    max_steps = atoi(argv[3]);
    if (argc > 5 && !strcmp(argv[5], "fp32")) precision = MDS_FP32_BULK;
  } else {
    fprintf(stderr, "unknown problem %s\n", kind);
    return 2;
  }
  if (n < 4) return 2;
  init_grid(kind, n);
  init_problem(kind);
  int rc = mds_update_eos(ctx);                 /* main.c:181 */
  if (rc) die("mds_update_eos", rc);
  t = 0.0;
  istep = 0;
  /* main.c:206-264; steps are enqueued in batches between clock reads */
  while (t < tmax && istep < max_steps) {
    for (int k = 0; k < 64 && istep + k < max_steps; k++) step();
    estimate_dt();
  }
  update_cpu();
  dump_ascii1(argv[4]);
  fprintf(stderr, "mds_driver: %s n=%d steps=%d t=%.17g\n", kind, n, istep, t);
  mds_destroy(ctx);
  mds_grid_free(&grid);
  if (use_table) mds_eos_table_free(&table);
  free(sim_p);
  return 0;
}
