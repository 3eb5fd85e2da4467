/*
 * orc_eos.c -- CPU oracle of the tabulated EOS (TEST INFRASTRUCTURE ONLY).
 *
 * orc_eos_walk restates the reference's table use in
 * eos/burrows/eos_stuff.c:507-650 (eos_given_rex_tianshu): bilinear weights in
 * (log10 rho, Ye), the walk along the T nodes from the previous temperature
 * until the energy is bracketed, the linear inverse in log T, and trilinear
 * ln P and Gamma_1 at the recovered temperature, with the walk's unit
 * handling: the specific energy divided by e_scale minus e_shift (MeV per
 * baryon and const_hack for a Burrows table, eos_stuff.c:530) is what the
 * energy record is compared with, and P = exp(ln P) * p_scale (MeV/fm^3 to
 * dyn/cm^2, :618); T is returned in the table's unit (MeV there, :531-532,
 * 645).  Differences, all stated in the library header as well: 0-based
 * nodes; FP32 table values; (rho, Ye) are clamped to the table edges (the
 * reference reads outside the table there).
 *
 * orc_eos_root restates the Newton iteration with central-difference
 * derivative and the bisection fallback of
 * eos/stellarcollapse/eos_stellar_collapse.c:294-369 (eos_bisect, eos_root),
 * applied to the same piecewise-linear interpolant of e in log T, so that
 * its root and orc_eos_walk's must agree to the iteration tolerance.
 *
 * Pinned by: the table nodes against an independent numpy evaluation of the
 * model; the walk's independence from its starting temperature; the round
 * trip e(T(e)) = e; Newton/bisection against the walk (tests/test_eos.py).
 */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

static double escale(const orc_eos_tab *t) { return t->e_scale > 0.0 ? t->e_scale : 1.0; }
static double pscale(const orc_eos_tab *t) { return t->p_scale > 0.0 ? t->p_scale : 1.0; }

static double tab(const orc_eos_tab *t, int rec, int jy, int jr, int jt)
{
  return (double)t->data[(((size_t)rec * t->ny + jy) * t->nr + jr) * t->nt + jt];
}

/* bilinear weights and corner indices (eos_stuff.c:534-560) */
typedef struct { int jy, jyp, jr, jrp; double c0, c1, c2, c3; } orc_w;

static orc_w weights(const orc_eos_tab *t, double rho, double ye)
{
  orc_w w;
  const double dr = (t->lr1 - t->lr0) / (t->nr - 1);
  const double dy = (t->ye1 - t->ye0) / (t->ny - 1);
  double rl = log10(rho);
  if (!(rl >= t->lr0)) rl = t->lr0;     /* also catches NaN */
  if (rl > t->lr1) rl = t->lr1;
  if (!(ye >= t->ye0)) ye = t->ye0;
  if (ye > t->ye1) ye = t->ye1;
  w.jr = (int)((rl - t->lr0) / dr);
  w.jy = (int)((ye - t->ye0) / dy);
  if (w.jr > t->nr - 1) w.jr = t->nr - 1;
  if (w.jy > t->ny - 1) w.jy = t->ny - 1;
  if (w.jr < 0) w.jr = 0;
  if (w.jy < 0) w.jy = 0;
  w.jyp = w.jy + 1;
  w.jrp = w.jr + 1;
  double drho = (rl - (t->lr0 + w.jr * dr)) / dr;
  double dye = (ye - (t->ye0 + w.jy * dy)) / dy;
  if (w.jy == t->ny - 1) { w.jyp = w.jy; dye = 0.0; }
  if (w.jr == t->nr - 1) { w.jrp = w.jr; drho = 0.0; }
  w.c0 = (1.0 - drho) * (1.0 - dye);
  w.c1 = dye * (1.0 - drho);
  w.c2 = (1.0 - dye) * drho;
  w.c3 = dye * drho;
  return w;
}

/* energy projected onto T node jt (eos_stuff.c:566-569) */
static double e_at(const orc_eos_tab *t, const orc_w *w, int jt)
{
  double e = w->c0 * tab(t, 0, w->jy, w->jr, jt);
  e = e + w->c1 * tab(t, 0, w->jyp, w->jr, jt);
  e = e + w->c2 * tab(t, 0, w->jy, w->jrp, jt);
  e = e + w->c3 * tab(t, 0, w->jyp, w->jrp, jt);
  return e;
}

/* trilinear record at (jt, dtemp) in the reference's summation order (eos_stuff.c:616-643) */
static double tri(const orc_eos_tab *t, const orc_w *w, int rec, int jt, double dtemp)
{
  const int jtp = jt + 1;
  double p = w->c0 * (1 - dtemp) * tab(t, rec, w->jy, w->jr, jt);
  p += w->c1 * (1 - dtemp) * tab(t, rec, w->jyp, w->jr, jt);
  p += w->c2 * (1 - dtemp) * tab(t, rec, w->jy, w->jrp, jt);
  p += w->c3 * (1 - dtemp) * tab(t, rec, w->jyp, w->jrp, jt);
  p += w->c0 * dtemp * tab(t, rec, w->jy, w->jr, jtp);
  p += w->c1 * dtemp * tab(t, rec, w->jyp, w->jr, jtp);
  p += w->c2 * dtemp * tab(t, rec, w->jy, w->jrp, jtp);
  p += w->c3 * dtemp * tab(t, rec, w->jyp, w->jrp, jtp);
  return p;
}

int orc_eos_walk(const orc_eos_tab *t, double rho, double u, double ye, double temp_guess, double out[4])
{
  const orc_w w = weights(t, rho, ye);
  const double dt = (t->lt1 - t->lt0) / (t->nt - 1);
  const double e = (u / rho) / escale(t) - t->e_shift;          /* eos_stuff.c:527, 530 */
  double temp = temp_guess < 1.0e-10 ? 1.0 : temp_guess;       /* eos_stuff.c:528 */
  const double st = (log10(temp) - t->lt0) / dt;
  int jt = !(st >= 0.0) ? 0 : (st > t->nt - 2 ? t->nt - 2 : (int)st);
  int status = 0;
  /* walk down to the last node below e (eos_stuff.c:563-577) */
  double elo, ehi = -1e99;
  while (1) {
    elo = e_at(t, &w, jt);
    if (elo < e || jt == 0) break;
    ehi = elo;
    jt = jt - 1;
  }
  if (elo > e) status |= 1;                                     /* off the bottom */
  /* walk up until the next node is above e (eos_stuff.c:579-590) */
  if (ehi < -1e98) {
    while (1) {
      ehi = e_at(t, &w, jt + 1);
      if (ehi > e || jt == t->nt - 2) break;
      elo = ehi;
      jt = jt + 1;
    }
  }
  if (ehi < e) status |= 2;                                     /* off the top */
  /* linear model in log T (eos_stuff.c:597-605) */
  const double tlo = t->lt0 + jt * dt;
  const double thi = tlo + dt;
  const double dtemp = (e - elo) / (ehi - elo);
  const double lt = tlo + (thi - tlo) * dtemp;
  const double lnp = tri(t, &w, 1, jt, dtemp);
  const double g1 = tri(t, &w, 2, jt, dtemp);
  const double p = exp(lnp) * pscale(t);                       /* eos_stuff.c:618 */
  out[0] = p;
  out[1] = sqrt((g1 * p) / rho);
  out[2] = pow(10.0, lt);
  out[3] = g1;
  return status;
}

/* the interpolant e(log T) whose root both inversions find */
static double e_of_lt(const orc_eos_tab *t, const orc_w *w, double lt)
{
  const double dt = (t->lt1 - t->lt0) / (t->nt - 1);
  double s = (lt - t->lt0) / dt;
  int jt = (int)floor(s);
  if (jt < 0) jt = 0;
  if (jt > t->nt - 2) jt = t->nt - 2;
  const double f = s - jt;
  return (1.0 - f) * e_at(t, w, jt) + f * e_at(t, w, jt + 1);
}

/* eos_bisect, eos_stellar_collapse.c:294-332 */
static int bisect(const orc_eos_tab *t, const orc_w *w, double lt, double val, double tol, double *res)
{
  double ltl, ltr, fl, fr, grow = 0.01;
  do {
    double d = fabs(grow * lt) + 1.e-10;
    ltl = lt - d;
    fl = e_of_lt(t, w, ltl) - val;
    ltr = lt + d;
    fr = e_of_lt(t, w, ltr) - val;
    grow *= 1.1;
  } while (fl * fr > 0 && ltl > t->lt0 && ltr < t->lt1);
  if (fl * fr > 0) {
    ltl = t->lt0;
    fl = e_of_lt(t, w, ltl) - val;
    ltr = t->lt1;
    fr = e_of_lt(t, w, ltr) - val;
    if (fl * fr > 0) return 0;
  }
  do {
    double ltm = 0.5 * (ltl + ltr);
    double fm = e_of_lt(t, w, ltm) - val;
    if (fl * fm <= 0.) { ltr = ltm; fr = fm; }
    else { ltl = ltm; fl = fm; }
  } while (ltr - ltl > tol);
  *res = 0.5 * (ltl + ltr);
  return 1;
}

/* eos_root, eos_stellar_collapse.c:334-369: Newton with central-difference
 * derivative (at most 10 iterations), bisection if it fails.  Returns the
 * status; *lt_out = log10 T. */
int orc_eos_root(const orc_eos_tab *t, double rho, double u, double ye, double lt_guess, double tol,
                 double *lt_out, int *newton_iters)
{
  const orc_w w = weights(t, rho, ye);
  const double val = (u / rho) / escale(t) - t->e_shift;
  double lt = lt_guess, lt_last;
  int iter = 0;
  do {
    double f = e_of_lt(t, &w, lt) - val;
    double dlt = fabs(1.e-6 * lt) + 1.e-10;
    double df = (e_of_lt(t, &w, lt + dlt) - e_of_lt(t, &w, lt - dlt)) / (2. * dlt);
    lt_last = lt;
    lt -= f / df;
    iter++;
  } while (iter < 10 && fabs(lt - lt_last) / (fabs(lt) + 1.e-10) > tol);
  *lt_out = lt;
  if (newton_iters) *newton_iters = iter;
  if ((iter == 10 && fabs(lt - lt_last) > tol) || isnan(lt) || isinf(lt)) {
    if (newton_iters) *newton_iters = -1;
    return bisect(t, &w, lt_guess, val, tol, lt_out);
  }
  return 1;
}
