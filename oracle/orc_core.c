/*
 * orc_core.c -- pointwise pieces of the CPU oracle (test infrastructure only).
 *
 * Restates, in the reference's operation order:
 *   reconstruct.c:127-213  para_recon  (volume-moment parabola, "generalised PPM")
 *   reconstruct.c:89-124   lin_recon
 *   riemann.c:4-62         p_uflux / pflux
 *   riemann.c:178-325      riemann_flux_HLLE
 *   riemann.c:328-499      riemann_flux_HLLC
 *   fluid.c:4-53           hydro_p_to_u / hydro_u_to_p   (geometry.c:75-91 lower/raise)
 *   fluid.c:56-73, rad_fluid.c:164-192   stress tensor and geometric source
 *   step.c:205-223         dt_source_control
 *
 * Every parenthesisation below is deliberate: it is the left-to-right C
 * evaluation order of the reference expression, so that a strict-IEEE build
 * reproduces the reference's rounding.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define OMIN(a, b) ((a) < (b) ? (a) : (b))   /* decs.h:627 */
#define OMAX(a, b) ((a) > (b) ? (a) : (b))   /* decs.h:628 */
#define OSIGN(a) ((a) < 0.0 ? -1.0 : 1.0)    /* decs.h:629 */
#define OSQR(x) ((x) * (x))

/* reconstruct.c:127-213.  Reconstructs zones istart+NG .. istop+NG of the
 * pencil a[] (x[] holds the lower-face coordinate of each pencil zone).
 * Pinned by: volume-average reproduction, exactness on linear data, the
 * extremum / face-clamp / relocation branches (tests/test_oracle_recon.py). */
void orc_para_recon(const double *a, double *al, double *ar,
                    const double *alpha, const double *beta, const double *Gamma,
                    const double *x, int istart, int istop)
{
  for (int i = istart + ORC_NG; i <= istop + ORC_NG; i++) {
    double dqm = a[i] - a[i - 1];
    double dqp = a[i + 1] - a[i];
    if (dqm * dqp <= 0.0) {          /* local extremum: flat */
      al[i] = ar[i] = a[i];
      continue;
    }
    double dqc = a[i + 1] - a[i - 1];
    double t012 = (alpha[i - 1] * beta[i]) * Gamma[i + 1];
    double t021 = (alpha[i - 1] * beta[i + 1]) * Gamma[i];
    double t102 = (alpha[i] * beta[i - 1]) * Gamma[i + 1];
    double t120 = (alpha[i] * beta[i + 1]) * Gamma[i - 1];
    double t201 = (alpha[i + 1] * beta[i - 1]) * Gamma[i];
    double t210 = (alpha[i + 1] * beta[i]) * Gamma[i - 1];
    double den = 1.0 / (((((t012 + t120) + t201) - t021) - t102) - t210);
    double A = ((((dqm * beta[i + 1]) * Gamma[i - 1]) * Gamma[i])
                + Gamma[i + 1] * (((dqp * beta[i - 1]) * Gamma[i]) - ((dqc * beta[i]) * Gamma[i - 1])))
               * den;
    double B = (-((((dqm * alpha[i + 1]) * Gamma[i - 1]) * Gamma[i])
                  + Gamma[i + 1] * (((dqp * alpha[i - 1]) * Gamma[i]) - ((dqc * alpha[i]) * Gamma[i - 1]))))
               * den;
    double C = a[i] - ((alpha[i] * A) + (beta[i] * B)) / Gamma[i];

    double l = x[i] * ((A * x[i]) + B) + C;
    double r = x[i + 1] * ((A * x[i + 1]) + B) + C;

    /* clamp each face value between the zone mean and 0.75*nbr+0.25*self */
    double mean = (0.75 * a[i - 1]) + (0.25 * a[i]);
    if (((a[i] - l) * (l - mean)) < 0.0) l = mean;
    mean = (0.75 * a[i + 1]) + (0.25 * a[i]);
    if (((mean - r) * (r - a[i])) < 0.0) r = mean;

    /* refit the parabola through (l, r, mean) and test for an interior extremum */
    double den2 = (x[i] - x[i + 1]) * ((alpha[i] + ((Gamma[i] * x[i]) * x[i + 1])) - (beta[i] * (x[i] + x[i + 1])));
    den2 = 1.0 / den2;
    double A2 = ((beta[i] * (r - l)) + Gamma[i] * (((a[i] * (x[i] - x[i + 1])) - (r * x[i])) + (l * x[i + 1]))) * den2;
    double B2 = (((alpha[i] * (l - r)) + ((Gamma[i] * (r - a[i])) * OSQR(x[i])))
                 + ((Gamma[i] * (a[i] - l)) * OSQR(x[i + 1]))) * den2;
    double xzero = (-B2) / ((2 * A2) + 1.0e-16);
    if (isinf(xzero) || isnan(xzero)) xzero = -1.0e-200;
    double xlo = x[i], xhi = x[i + 1];
    double xmid = 0.5 * (x[i] + x[i + 1]);
    if (xzero > xlo && xzero < xhi) {
      if (xzero > xmid) {            /* push the extremum onto the right face */
        double A3 = (Gamma[i] * (a[i] - r)) / (alpha[i] + (xhi * ((Gamma[i] * xhi) - (2.0 * beta[i]))));
        double B3 = (-2.0 * A3) * xhi;
        double C3 = a[i] - ((alpha[i] * A3) + (beta[i] * B3)) / Gamma[i];
        l = x[i] * ((A3 * x[i]) + B3) + C3;
      } else {                       /* ... or onto the left face */
        double A3 = (Gamma[i] * (a[i] - l)) / (alpha[i] + (xlo * ((Gamma[i] * xlo) - (2.0 * beta[i]))));
        double B3 = (-2.0 * A3) * xlo;
        double C3 = a[i] - ((alpha[i] * A3) + (beta[i] * B3)) / Gamma[i];
        r = x[i + 1] * ((A3 * x[i + 1]) + B3) + C3;
      }
    }
    al[i] = l;
    ar[i] = r;
  }
}

/* reconstruct.c:89-124 (interp_order 2). */
void orc_lin_recon(const double *a, double *al, double *ar,
                   const double *beta, const double *Gamma,
                   const double *x, int istart, int istop)
{
  for (int i = istart + ORC_NG; i <= istop + ORC_NG; i++) {
    double xm = beta[i - 1] / Gamma[i - 1];
    double xp = beta[i + 1] / Gamma[i + 1];
    double x0 = beta[i] / Gamma[i];
    double sm = (a[i] - a[i - 1]) / (x0 - (0.5 * (x[i] + xm)));
    double sp = (a[i + 1] - a[i]) / ((0.5 * (x[i + 1] + xp)) - x0);
    if (sm * sp <= 0.0) {
      al[i] = ar[i] = a[i];
      continue;
    }
    double sc = (a[i + 1] - a[i - 1]) / (xp - xm);
    double A = OSIGN(sc) * OMIN(OMIN(fabs(sm), fabs(sp)), fabs(sc));
    double B = a[i] - ((A * beta[i]) / Gamma[i]);
    al[i] = (A * x[i]) + B;
    ar[i] = (A * x[i + 1]) + B;
  }
}

/* riemann.c:4-37: conserved state and physical flux of one side. */
static void p_uflux(int nhydro, const double *p, const double *gcov, double press,
                    int dir, double *u, double *f)
{
  double vcov[3];
  double v = p[ORC_U1 + dir];
  for (int d = 0; d < 3; d++) vcov[d] = gcov[d] * p[ORC_U1 + d];
  u[ORC_RHO] = p[ORC_RHO];
  double vsq = 0.0;
  for (int d = 0; d < 3; d++) vsq += p[ORC_U1 + d] * vcov[d];
  double etot = ((0.5 * u[ORC_RHO]) * vsq) + p[ORC_UU];
  u[ORC_ETOT] = etot;
  for (int d = 0; d < 3; d++) u[ORC_U1 + d] = u[ORC_RHO] * vcov[d];
  for (int vv = ORC_U1 + 3; vv < nhydro; vv++) u[vv] = u[ORC_RHO] * p[vv];
  f[ORC_RHO] = u[ORC_RHO] * v;
  f[ORC_ETOT] = v * (etot + press);
  for (int d = 0; d < 3; d++) f[ORC_U1 + d] = f[ORC_RHO] * vcov[d];
  f[ORC_U1 + dir] += press;
  for (int vv = ORC_U1 + 3; vv < nhydro; vv++) f[vv] = f[ORC_RHO] * p[vv];
}

/* riemann.c:39-62: area-weighted physical flux (supersonic upwinding). */
static void pflux(int nhydro, const double *p, double press, double area,
                  const double *gcov, int dir, double *f)
{
  double vcov[3];
  double v = p[ORC_U1 + dir];
  for (int d = 0; d < 3; d++) vcov[d] = gcov[d] * p[ORC_U1 + d];
  f[ORC_RHO] = (area * p[ORC_RHO]) * v;
  for (int d = 0; d < 3; d++) f[ORC_U1 + d] = f[ORC_RHO] * vcov[d];
  f[ORC_U1 + dir] += area * press;
  double vsq = 0.0;
  for (int d = 0; d < 3; d++) vsq += p[ORC_U1 + d] * vcov[d];
  double etot = ((0.5 * p[ORC_RHO]) * vsq) + p[ORC_UU];
  f[ORC_ETOT] = (area * v) * (etot + press);
  for (int vv = ORC_U1 + 3; vv < nhydro; vv++) f[vv] = f[ORC_RHO] * p[vv];
}

static double shock_q(double ps, double pk, double Gk)
{
  /* riemann.c:389-390 (Toro eq. 9.33 style wave-speed factor) */
  return (ps <= pk) ? 1.0 : sqrt(1.0 + (((Gk + 1.0) / (2.0 * Gk)) * ((ps / pk) - 1.0)));
}

/* Wave-speed estimate shared by HLLC and HLLE (riemann.c:210-303, 368-429).
 * hllc_floor selects MAX(0,p_s) (HLLC, riemann.c:386) vs MAX(p_s,0.0)
 * (HLLE, riemann.c:226), which differ only for NaN. */
static void wave_speeds(const double *pl, const double *pr, double p_l, double p_r,
                        double G_l, double G_r, double u_l, double u_r,
                        double cs_l, double cs_r, int hllc_floor,
                        double *sl_o, double *sr_o, double *ss_o, double *ps_o,
                        int *isshock)
{
  double p_max, p_min;
  if (p_l > p_r) { p_max = p_l; p_min = p_r; } else { p_max = p_r; p_min = p_l; }
  double pratio = p_max / p_min;
  *isshock = ((p_max - p_min) / p_min > 0.5) ? 1 : 0;
  double rl = pl[ORC_RHO], rr = pr[ORC_RHO];
  double p_s = (0.5 * (p_l + p_r)) - (((0.125 * (u_r - u_l)) * (rl + rr)) * (cs_l + cs_r));
  if (hllc_floor) p_s = OMAX(0, p_s);
  else p_s = OMAX(p_s, 0.0);
  double sl, sr, ss;
  if (pratio < 2.0 && p_s >= p_min && p_s <= p_max) {
    /* PVRS estimate accepted */
  } else if (p_s < p_min) {
    /* two rarefactions, Toro eqs. 9.35-9.36 */
    double G_avg = 0.5 * (G_l + G_r);
    double plr = pow(p_l / p_r, (G_avg - 1.0) / (2.0 * G_avg));
    ss = ((((plr * u_l) / cs_l) + (u_r / cs_r)) + ((2.0 * (plr - 1.0)) / (G_avg - 1.0)))
         / ((plr / cs_l) + (1.0 / cs_r));
    p_s = 0.5 * ((p_l * pow(1.0 + (((G_l - 1.0) / (2.0 * cs_l)) * (u_l - ss)), (2.0 * G_l) / (G_l - 1.0)))
                 + (p_r * pow(1.0 + (((G_r - 1.0) / (2.0 * cs_r)) * (ss - u_r)), (2.0 * G_r) / (G_r - 1.0))));
    p_s = OMAX(p_s, 0.0);
  } else {
    /* two shocks, Toro eqs. 9.42-9.43 */
    double a_l = 2.0 / ((G_l + 1.0) * rl);
    double a_r = 2.0 / ((G_r + 1.0) * rr);
    double b_l = ((G_l - 1.0) / (G_l + 1.0)) * p_l;
    double b_r = ((G_r - 1.0) / (G_r + 1.0)) * p_r;
    double g_l = sqrt(a_l / (p_s + b_l));
    double g_r = sqrt(a_r / (p_s + b_r));
    p_s = (((g_l * p_l) + (g_r * p_r)) - (u_r - u_l)) / (g_l + g_r);
    p_s = OMAX(p_s, 0.0);
  }
  double q_l = shock_q(p_s, p_l, G_l);
  double q_r = shock_q(p_s, p_r, G_r);
  sl = u_l - (cs_l * q_l);
  sr = u_r + (cs_r * q_r);
  ss = (((p_r - p_l) + ((rl * u_l) * (sl - u_l))) - ((rr * u_r) * (sr - u_r)))
       / ((rl * (sl - u_l)) - (rr * (sr - u_r)));
  *sl_o = sl; *sr_o = sr; *ss_o = ss; *ps_o = p_s;
}

/* riemann.c:328-499.  Returns isshock ((pmax-pmin)/pmin > 0.5).
 * Pinned by: consistency F(U,U)=F(U), supersonic upwinding, mirror symmetry,
 * stationary contact, Toro branch selection (tests/test_oracle_riemann.py). */
int orc_riemann_hllc(int nhydro, int nvars, const double *pl, const double *pr,
                     const orc_zone_geom *g, int dir,
                     double *flux, double *sig_speed, double *vriemann)
{
  double p_l = pl[nvars], p_r = pr[nvars];
  double G_l = pl[nvars + 1], G_r = pr[nvars + 1];
  double area = g->area[dir];
  const double *gcov = g->gcov[1 + dir];
  double scale = g->scale[1 + dir][dir];

  if (area < 1.0e-16) {
    *sig_speed = 1.0e-200;
    for (int d = 0; d < 3; d++) vriemann[d] = 0.5 * (pl[ORC_U1 + d] + pr[ORC_U1 + d]);
    for (int vv = 0; vv < nhydro; vv++) flux[vv] = 0.0;
    return 0;
  }
  double u_l = pl[ORC_U1 + dir] * scale;
  double u_r = pr[ORC_U1 + dir] * scale;
  double cs_l = sqrt((G_l * p_l) / pl[ORC_RHO]);
  double cs_r = sqrt((G_r * p_r) / pr[ORC_RHO]);
  double sl, sr, ss, p_s;
  int isshock;
  wave_speeds(pl, pr, p_l, p_r, G_l, G_r, u_l, u_r, cs_l, cs_r, 1, &sl, &sr, &ss, &p_s, &isshock);

  if (ss > sr) sr = 1.001 * ss;
  if (ss < sl) sl = 1.001 * ss;
  *sig_speed = OMAX(fabs(sl), fabs(sr)) / scale;

  if (sl > 0.0) {
    pflux(nhydro, pl, p_l, area, gcov, dir, flux);
    for (int d = 0; d < 3; d++) vriemann[d] = pl[ORC_U1 + d];
    return isshock;
  } else if (sr < 0.0) {
    pflux(nhydro, pr, p_r, area, gcov, dir, flux);
    for (int d = 0; d < 3; d++) vriemann[d] = pr[ORC_U1 + d];
    return isshock;
  }

  double u[ORC_NSIZE], fx[ORC_NSIZE], Us[ORC_NSIZE], vcov[3];
  double slr;
  const double *pk = (ss > 0.0) ? pl : pr;
  if (ss > 0.0) {
    p_uflux(nhydro, pl, gcov, p_l, dir, u, fx);
    Us[0] = (pl[ORC_RHO] * (sl - u_l)) / (sl - ss);
    Us[ORC_ETOT] = (((u[ORC_ETOT] * (sl - u_l)) - (p_l * u_l)) + (p_s * ss)) / (sl - ss);
    slr = sl / scale;
  } else {
    p_uflux(nhydro, pr, gcov, p_r, dir, u, fx);
    Us[0] = (pr[ORC_RHO] * (sr - u_r)) / (sr - ss);
    Us[ORC_ETOT] = (((u[ORC_ETOT] * (u_r - sr)) + (p_r * u_r)) - (p_s * ss)) / (ss - sr);
    slr = sr / scale;
  }
  for (int d = 0; d < 3; d++) Us[ORC_U1 + d] = pk[ORC_U1 + d];
  Us[ORC_U1 + dir] = ss / scale;
  for (int d = 0; d < 3; d++) vcov[d] = gcov[d] * Us[ORC_U1 + d];
  for (int d = 0; d < 3; d++) {
    vriemann[d] = Us[ORC_U1 + d];
    Us[ORC_U1 + d] = vcov[d] * Us[0];
  }
  for (int vv = ORC_U1 + 3; vv < nhydro; vv++) Us[vv] = Us[0] * pk[vv];
  for (int vv = 0; vv < nhydro; vv++) flux[vv] = area * (fx[vv] + (slr * (Us[vv] - u[vv])));
  return isshock;
}

/* riemann.c:178-325 (HLLE, used transverse to tagged shocks). */
int orc_riemann_hlle(int nhydro, int nvars, const double *pl, const double *pr,
                     const orc_zone_geom *g, int dir,
                     double *flux, double *sig_speed, double *vriemann)
{
  double p_l = pl[nvars], p_r = pr[nvars];
  double G_l = pl[nvars + 1], G_r = pr[nvars + 1];
  double area = g->area[dir];
  const double *gcov = g->gcov[1 + dir];
  double scale = g->scale[1 + dir][dir];

  if (area < 1.0e-16) {
    *sig_speed = 1.0e-200;
    for (int d = 0; d < 3; d++) vriemann[d] = 0.5 * (pl[ORC_U1 + d] + pr[ORC_U1 + d]);
    for (int vv = 0; vv < nhydro; vv++) flux[vv] = 0.0;
    return 0;
  }
  double u_l = pl[ORC_U1 + dir] * scale;
  double u_r = pr[ORC_U1 + dir] * scale;
  double cs_l = sqrt((G_l * p_l) / pl[ORC_RHO]);
  double cs_r = sqrt((G_r * p_r) / pr[ORC_RHO]);
  double sl, sr, ss, p_s;
  int isshock;
  wave_speeds(pl, pr, p_l, p_r, G_l, G_r, u_l, u_r, cs_l, cs_r, 0, &sl, &sr, &ss, &p_s, &isshock);

  double rhowl = sqrt(pl[ORC_RHO]);
  double rhowr = sqrt(pr[ORC_RHO]);
  for (int d = 0; d < 3; d++)
    vriemann[d] = ((rhowl * pl[ORC_U1 + d]) + (rhowr * pr[ORC_U1 + d])) / (rhowl + rhowr);

  double ul[ORC_NSIZE], ur[ORC_NSIZE], fl[ORC_NSIZE], fr[ORC_NSIZE];
  p_uflux(nhydro, pl, gcov, p_l, dir, ul, fl);
  p_uflux(nhydro, pr, gcov, p_r, dir, ur, fr);
  double cmax = OMAX(sr / scale, 0.0);
  double cmin = fabs(OMIN(sl / scale, 0.0));
  *sig_speed = OMAX(cmax, cmin);
  for (int vv = 0; vv < nhydro; vv++)
    flux[vv] = (area * (((cmin * fr[vv]) + (cmax * fl[vv])) - ((cmax * cmin) * (ur[vv] - ul[vv]))))
               / (cmax + cmin);
  return 0;
}

/* fluid.c:4-25 with geometry.c:75-82 */
void orc_p_to_u(const double *p, double *u, const double *gcov, int nhydro)
{
  double vcov[3];
  for (int d = 0; d < 3; d++) vcov[d] = gcov[d] * p[ORC_U1 + d];
  u[ORC_RHO] = p[ORC_RHO];
  double vsq = 0.0;
  for (int d = 0; d < 3; d++) {
    vsq += p[d + ORC_U1] * vcov[d];
    u[d + ORC_U1] = u[ORC_RHO] * vcov[d];
  }
  u[ORC_ETOT] = ((0.5 * u[ORC_RHO]) * vsq) + p[ORC_UU];
  for (int d = ORC_U1 + 3; d < nhydro; d++) u[d] = u[ORC_RHO] * p[d];
}

/* fluid.c:27-53 with geometry.c:84-91 (floors off: ENFORCE_FLOORS is not set
 * for any gamma-law setup, setups/kh_2d.mk:19). */
void orc_u_to_p(double *u, double *p, const double *gcon, int nhydro)
{
  p[ORC_RHO] = u[ORC_RHO];
  for (int d = ORC_U1; d < ORC_U1 + 3; d++) u[d] /= u[ORC_RHO];
  for (int d = 0; d < 3; d++) p[ORC_U1 + d] = gcon[d] * u[ORC_U1 + d];
  double vsq = 0.0;
  for (int d = 0; d < 3; d++) vsq += u[ORC_U1 + d] * p[ORC_U1 + d];
  p[ORC_UU] = u[ORC_ETOT] - ((0.5 * p[ORC_RHO]) * vsq);
  for (int d = ORC_U1 + 3; d < nhydro; d++) p[d] = u[d] / u[ORC_RHO];
}

/* rad_fluid.c:164-192 with fluid.c:56-73.  Zero unless spherical. */
void orc_geom_src(const double *p, double press, double *src,
                  const orc_zone_geom *g, int nvars, int geom)
{
  for (int vv = 0; vv < nvars; vv++) src[vv] = 0;
  if (geom != ORC_GEOM_SPHERICAL) return;
  double T[3][3], vcov[3];
  for (int d = 0; d < 3; d++) vcov[d] = g->gcov[0][d] * p[ORC_U1 + d];
  for (int l = 0; l < 3; l++) {
    for (int m = 0; m < 3; m++) T[l][m] = (p[ORC_RHO] * p[ORC_U1 + m]) * vcov[l];
    T[l][l] += press;
  }
  for (int l = 0; l < 3; l++)
    for (int m = 0; m < 3; m++)
      for (int n = 0; n < 3; n++) src[ORC_U1 + m] += T[l][n] * g->conn[l][m][n];
}

/* step.c:205-223 */
double orc_dt_source_control(const double *p, const double *u, const double *s)
{
  const double alpha = 0.02;
  double mdt = 1.0e100;
  double local_dt = (alpha * u[0]) / (fabs(s[0]) + 1.0e-16);
  mdt = OMIN(mdt, local_dt);
  local_dt = (alpha * u[1]) / (fabs(s[1]) + 1.0e-16);
  mdt = OMIN(mdt, local_dt);
  double momterm = 0.0;
  for (int d = 0; d < 3; d++) momterm += p[ORC_U1 + d] * s[ORC_U1 + d];
  local_dt = ((100 * alpha) * p[ORC_UU]) / (fabs(fabs(s[1]) - fabs(momterm)) + 1.0e-16);
  mdt = OMIN(mdt, local_dt);
  return mdt;
}
