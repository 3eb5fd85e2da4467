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
 *   mds_driver star <n> <steps> <dump> [fp32] [name=value ...]
 *                                                polytropic star column with the
 *                                                synthetic table EOS and monopole gravity
 * The dump is the reference's 1D ASCII layout: "# t", then per zone
 * x (r for spherical grids), the primitives times scale (rho, u, v1, v2, v3,
 * [Ye]), the eos record (P, cs, T, entropy, Gamma), the enclosed mass at the
 * zone centre and the two lapses (1 1 without GR).
 *
 * Star parameters, named as in the reference's input file (io.c:5018-5102):
 *   gr_monopole=1        the GR effective monopole (GR_MONOPOLE; the dump
 *                        prints its lapses)
 *   detect_tbounce=1     bounce time: the first step after which the central
 *                        density is below 0.95 of its running maximum
 *                        (set_tbounce, step.c:130-153)
 *   restart_create_1d=1  write restarts/<model>.<n1>.<nr1>g.<ms>ms.restart_1d
 *                        in the dump layout once t >= tbounce + restart_delay,
 *                        and exit with status 1 (main.c:268-274, io.c:139-147);
 *                        implies detect_tbounce
 *   restart_delay=<s>, perturb_delay=<s>, model_name=<path>
 *   restart_from_1d=<file>  initial state from such a file (init_ccsn.c:157-264):
 *                        n lines after the header, radius, rho, v1 (physical),
 *                        Ye and T interpolated in r onto the zone centres,
 *                        u from the table at (rho, T, Ye) (u_given_rtx)
 *   vexp=<1/s>           initial radial velocity v = vexp r exp(-(r / 1e7 cm)^2)
 *                        added to the polytrope: the core expands (vexp > 0)
 *                        or collapses homologously, the envelope stays at rest
 *                        (a set-up that exercises the bounce detector)
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <sys/stat.h>
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
static double tbounce = -1.0, rho_cent_max = 0.0;

/* ---- reference run parameters (io.c:5041-5102 defaults) ---- */
static int detect_tbounce, restart_create_1d, gr_monopole;
static double restart_delay = 0.010, perturb_delay = 0.010, vexp;
static const char *model_name = "star", *restart_from_1d;

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

/* zone-centre radius r_of_x(startx + (i + 1/2) dx) of the log-r column, and
 * its scale factor (dr/dx) */
static double zone_radius(int i) { return exp(grid.startx[0] + (i + 0.5) * grid.dx[0]); }
static double zone_scale(int i) { return grid.zone[(size_t)(i + MDS_NG) * MDS_ZG_COUNT + MDS_ZG_SCALE0]; }

/* model_interp (init_ccsn.c:72-78): quadratic in r through two points */
static double model_interp(double r, double r0, double r1, double p0, double p1)
{
  const double A = (p1 - p0) / (r1 * r1 - r0 * r0);
  const double B = p0 - A * (r0 * r0);
  return A * (r * r) + B;
}

/* The 1D restart reader of init_ccsn.c:157-264 (restart_from_1d): skip the
 * header line, require n data lines, read radius, rho, v1 (physical), Ye
 * and T (columns 1, 2, 4, 7, 10 of the dump), then per zone interpolate
 * linearly in r between the bracketing lines (inside the first line's
 * radius: model_interp for rho, Ye and T and v scaled by r / r0) and take u
 * from the table at (rho, T, Ye).  Past the last line the last two lines
 * extrapolate (the reference reads beyond its arrays there). */
static int restart_read_1d(const char *path)
{
  const int n = grid.n[0];
  FILE *fp = fopen(path, "r");
  if (!fp) { fprintf(stderr, "[init_problem]:  problem opening the 1-d model %s!\n", path); exit(1); }
  char *line = NULL;
  size_t cap = 0;
  int lines = 0;
  if (getline(&line, &cap, fp) < 0) { fprintf(stderr, "[init_problem]:  %s is empty\n", path); exit(1); }
  while (getline(&line, &cap, fp) > 0) lines++;
  if (lines != n) {
    fprintf(stderr, "[init_problem]:  Found %d lines in file %s, but require %d\n", lines, path, n);
    exit(1);
  }
  rewind(fp);
  if (getline(&line, &cap, fp) < 0) exit(1);
  double *rm = malloc(sizeof(double) * 5 * (size_t)lines);
  double *rhom = rm + lines, *vm = rm + 2 * lines, *yem = rm + 3 * lines, *tm = rm + 4 * lines;
  for (int i = 0; i < lines; i++) {
    double col[15];
    if (getline(&line, &cap, fp) < 0) exit(1);
    char *p = line, *e;
    for (int c = 0; c < 15; c++) {
      col[c] = strtod(p, &e);
      if (e == p) { fprintf(stderr, "[init_problem]:  %s line %d: need 15 columns\n", path, i + 2); exit(1); }
      p = e;
    }
    rm[i] = col[0]; rhom[i] = col[1]; vm[i] = col[3]; yem[i] = col[6]; tm[i] = col[9];
  }
  free(line);
  fclose(fp);
  for (int ii = 0; ii < n; ii++) {
    const double r = zone_radius(ii);
    int i = 0;
    while (i < lines && r > rm[i]) i++;
    i--;
    double rho, ye, temp, v;
    if (i < 0) {
      rho = model_interp(r, rm[0], rm[1], rhom[0], rhom[1]);
      ye = model_interp(r, rm[0], rm[1], yem[0], yem[1]);
      temp = model_interp(r, rm[0], rm[1], tm[0], tm[1]);
      v = r / rm[0] * vm[0];
    } else {
      if (i > lines - 2) i = lines - 2;
      const double del = (r - rm[i]) / (rm[i + 1] - rm[i]);
      rho = (1.0 - del) * rhom[i] + del * rhom[i + 1];
      ye = (1.0 - del) * yem[i] + del * yem[i + 1];
      temp = (1.0 - del) * tm[i] + del * tm[i + 1];
      v = (1.0 - del) * vm[i] + del * vm[i + 1];
    }
    double eo[4];
    const int rc = mds_eos_rtx(&table, rho, temp, ye, eo);          /* get_model_u, init_ccsn.c:462-469 */
    if (rc) { free(rm); return rc; }
    sim_p[ii] = rho;
    sim_p[(size_t)n + ii] = eo[0];
    sim_p[2 * (size_t)n + ii] = v / zone_scale(ii);
    sim_p[3 * (size_t)n + ii] = 0.0;
    sim_p[4 * (size_t)n + ii] = 0.0;
    sim_p[5 * (size_t)n + ii] = ye;
  }
  free(rm);
  return 0;
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
    prob.monopole = gr_monopole ? MDS_MONOPOLE_GR : MDS_MONOPOLE_NEWTON;
    if (restart_from_1d) {
      rc = restart_read_1d(restart_from_1d);
    } else {
      rc = mds_init_polytrope(&grid, 3.0, 1.0e10, 6.0e14, 1.0e4, gam, &table, 0x5EED0005ull, 0.01,
                              0x5EED0006ull, sim_p);
      for (int i = 0; rc == 0 && vexp != 0.0 && i < n; i++)
        sim_p[2 * (size_t)n + i] += vexp * zone_radius(i) * exp(-pow(zone_radius(i) / 1.0e7, 2)) / zone_scale(i);
    }
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
  double *lapse = NULL, *lapse_edge = NULL;
  if (gr_monopole) {
    lapse = calloc(n, sizeof(double));
    lapse_edge = calloc(n + 1, sizeof(double));
    int rc = mds_get_gr_state(ctx, 0, NULL, NULL, lapse_edge, lapse, NULL, NULL);
    if (rc) die("mds_get_gr_state", rc);
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
    /* geom lapse[0], lapse[1]: zone-centre and lower-face lapse (io.c:198-202) */
    fprintf(fp, "%16.15e %16.15e \n", lapse ? lapse[i] : 1.0, lapse_edge ? lapse_edge[i] : 1.0);
  }
  fclose(fp);
  free(press); free(cs); free(temp); free(g1); free(lapse); free(lapse_edge);
}

/* set_tbounce (step.c:130-153): the central (first) zone's density against
 * its running maximum; t is the time at the start of the step just taken */
static void set_tbounce(double t_step)
{
  update_cpu();
  const double rho_cent = sim_p[0];
  if (rho_cent > rho_cent_max) rho_cent_max = rho_cent;
  if (rho_cent < 0.95 * rho_cent_max) {
    tbounce = t_step;
    fprintf(stderr, "[set_tbounce]:  Bounce detected! tbounce=%1.6f\n", tbounce);
  }
}

/* main.c:268-274 with io.c:139-147: the 1D ASCII restart file, then exit(1) */
static void restart_dump_1d(void)
{
  char name[1024];
  const char *str = strrchr(model_name, '/');
  str = str ? str + 1 : model_name;
  mkdir("restarts", 0777);
  snprintf(name, sizeof name, "restarts/%s.%d.%dg.%dms.restart_1d", str, grid.n[0], 0, (int)(1000 * perturb_delay));
  fprintf(stderr, "Writing 1-d ASCII restart file...\n");
  update_cpu();
  dump_ascii1(name);
  fprintf(stderr, "mds_driver: restart %s at t=%.17g, step %d\n", name, t, istep);
  exit(1);
}

static int parse_param(const char *a)
{
  const char *eq = strchr(a, '=');
  if (!eq) return -1;
  const size_t k = (size_t)(eq - a);
  const char *v = eq + 1;
  if (k == 14 && !strncmp(a, "detect_tbounce", k)) detect_tbounce = atoi(v);
  else if (k == 17 && !strncmp(a, "restart_create_1d", k)) restart_create_1d = atoi(v);
  else if (k == 11 && !strncmp(a, "gr_monopole", k)) gr_monopole = atoi(v);
  else if (k == 13 && !strncmp(a, "restart_delay", k)) restart_delay = atof(v);
  else if (k == 13 && !strncmp(a, "perturb_delay", k)) perturb_delay = atof(v);
  else if (k == 10 && !strncmp(a, "model_name", k)) model_name = v;
  else if (k == 15 && !strncmp(a, "restart_from_1d", k)) restart_from_1d = v;
  else if (k == 4 && !strncmp(a, "vexp", k)) vexp = atof(v);
  else return -1;
  return 0;
}

int main(int argc, char **argv)
{
  if (argc < 5) {
    fprintf(stderr, "usage: %s sod <n> <tmax> <dump> | star <n> <steps> <dump> [fp32] [name=value ...]\n", argv[0]);
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
    for (int k = 5; k < argc; k++) {
      if (!strcmp(argv[k], "fp32")) precision = MDS_FP32_BULK;
      else if (parse_param(argv[k])) { fprintf(stderr, "unknown parameter %s\n", argv[k]); return 2; }
    }
    /* io.c:5718-5771: a restart file needs the bounce time; a restarted run
     * does not look for a bounce */
    if (restart_create_1d) detect_tbounce = 1;
    if (restart_from_1d) detect_tbounce = 0;
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
  /* main.c:206-264; steps are enqueued in batches between clock reads,
   * one at a time while a bounce is looked for */
  while (t < tmax && istep < max_steps) {
    if (detect_tbounce) {
      const double t_step = t;
      step();
      estimate_dt();
      if (tbounce < 0.0) set_tbounce(t_step);
      if (restart_create_1d && tbounce > 0.0 && t >= tbounce + restart_delay) restart_dump_1d();
      continue;
    }
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
