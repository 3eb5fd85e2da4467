// mds_physics.h -- pointwise device physics of the hydro sweep (gfx950).
//
// Two families:
//   *_x64  FP64, the reference's operation order (reconstruct.c:127-213,
//          riemann.c:4-62,178-499, fluid.c:4-53, rad_fluid.c:164-192,
//          step.c:205-223).  Built with -ffp-contract=off so that every
//          product and sum rounds exactly as in the reference CPU build.
//   *_f32  FP32-bulk: the same algorithms in single precision with the
//          well-conditioned local reconstruction (MDS_RC_* coefficients),
//          hardware reciprocal / square root, FMA contraction allowed.
//
// Velocity components beyond NV are identically zero and skipped; every
// skipped term is an exact +-0 contribution in the reference arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MDS_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// geometry of one face as seen by a Riemann solve: area, scale[1+dir][dir],
// gcov[1+dir][0..2].  CART: unit metric, so every multiply/divide by scale or
// gcov is the identity and is dropped at compile time.
template <class R>
struct FaceGeo {
  R area;
  R scale;
  R gcov[3];
};

template <class R, int NV>
struct PState {          // reconstructed primitive state on one side of a face
  R rho, uu, v[3], P, G;
  R ye;                  // passive scalar (electron fraction), used when NS = 1
};

// slot of the passive scalar in conserved/flux records (decs.h:646: YE after U3)
#define MDS_YI 5

// field-wise select of one of two states
template <class R, int NV>
__device__ __forceinline__ PState<R, NV> pick_state(bool first, const PState<R, NV> &a, const PState<R, NV> &b)
{
  PState<R, NV> k;
  k.rho = first ? a.rho : b.rho;
  k.uu = first ? a.uu : b.uu;
  for (int d = 0; d < 3; d++) k.v[d] = first ? a.v[d] : b.v[d];
  k.P = first ? a.P : b.P;
  k.G = first ? a.G : b.G;
  k.ye = first ? a.ye : b.ye;
  return k;
}

template <class R, int NV>
struct Flux {
  R f[5];                // rho, E, m1..m3 (area weighted)
};

// ---------------------------------------------------------------------------
// FP64, reference operation order
// ---------------------------------------------------------------------------

// para_recon for one zone (reconstruct.c:142-201).  am, a0, ap: zone means;
// al_m..: volume moments alpha/beta/Gamma of zones i-1, i, i+1; xlo, xhi:
// faces of zone i.  den and den2 depend on the geometry only; the reference
// recomputes them per variable with identical operands, so hoisting them is
// bit-exact.
struct ReconGeo64 {
  double am, a0, ap;     // alpha
  double bm, b0, bp;     // beta
  double gm, g0, gp;     // Gamma
  double xlo, xhi;
  double den, den2;
  int order = 3;         // hydro_interp_order: 3 para_recon, 2 lin_recon, 1 piecewise constant
};

MDS_DEV void recon_geo_x64(ReconGeo64 &g)
{
  double t012 = (g.am * g.b0) * g.gp;
  double t021 = (g.am * g.bp) * g.g0;
  double t102 = (g.a0 * g.bm) * g.gp;
  double t120 = (g.a0 * g.bp) * g.gm;
  double t201 = (g.ap * g.bm) * g.g0;
  double t210 = (g.ap * g.b0) * g.gm;
  g.den = 1.0 / (((((t012 + t120) + t201) - t021) - t102) - t210);
  double d2 = (g.xlo - g.xhi) * ((g.a0 + ((g.g0 * g.xlo) * g.xhi)) - (g.b0 * (g.xlo + g.xhi)));
  g.den2 = 1.0 / d2;
}

MDS_DEV void para_recon_x64(double qm, double q0, double qp, const ReconGeo64 &g,
                            double &al, double &ar)
{
  double dqm = q0 - qm;
  double dqp = qp - q0;
  if (dqm * dqp <= 0.0) { al = q0; ar = q0; return; }
  double dqc = qp - qm;
  double A = ((((dqm * g.bp) * g.gm) * g.g0) + g.gp * (((dqp * g.bm) * g.g0) - ((dqc * g.b0) * g.gm))) * g.den;
  double B = (-((((dqm * g.ap) * g.gm) * g.g0) + g.gp * (((dqp * g.am) * g.g0) - ((dqc * g.a0) * g.gm)))) * g.den;
  double C = q0 - ((g.a0 * A) + (g.b0 * B)) / g.g0;
  double l = g.xlo * ((A * g.xlo) + B) + C;
  double r = g.xhi * ((A * g.xhi) + B) + C;
  double mean = (0.75 * qm) + (0.25 * q0);
  if (((q0 - l) * (l - mean)) < 0.0) l = mean;
  mean = (0.75 * qp) + (0.25 * q0);
  if (((mean - r) * (r - q0)) < 0.0) r = mean;
  double A2 = ((g.b0 * (r - l)) + g.g0 * (((q0 * (g.xlo - g.xhi)) - (r * g.xlo)) + (l * g.xhi))) * g.den2;
  double B2 = (((g.a0 * (l - r)) + ((g.g0 * (r - q0)) * (g.xlo * g.xlo))) + ((g.g0 * (q0 - l)) * (g.xhi * g.xhi))) * g.den2;
  double xz = (-B2) / ((2.0 * A2) + 1.0e-16);
  if (__builtin_isinf(xz) || __builtin_isnan(xz)) xz = -1.0e-200;
  double xmid = 0.5 * (g.xlo + g.xhi);
  if (xz > g.xlo && xz < g.xhi) {
    if (xz > xmid) {
      double A3 = (g.g0 * (q0 - r)) / (g.a0 + (g.xhi * ((g.g0 * g.xhi) - (2.0 * g.b0))));
      double B3 = (-2.0 * A3) * g.xhi;
      double C3 = q0 - ((g.a0 * A3) + (g.b0 * B3)) / g.g0;
      l = g.xlo * ((A3 * g.xlo) + B3) + C3;
    } else {
      double A3 = (g.g0 * (q0 - l)) / (g.a0 + (g.xlo * ((g.g0 * g.xlo) - (2.0 * g.b0))));
      double B3 = (-2.0 * A3) * g.xlo;
      double C3 = q0 - ((g.a0 * A3) + (g.b0 * B3)) / g.g0;
      r = g.xhi * ((A3 * g.xhi) + B3) + C3;
    }
  }
  al = l;
  ar = r;
}

// lin_recon for one zone (reconstruct.c:89-124): a limited linear profile
// through the zone's centroid x0 = beta/Gamma, the slope the smallest of the
// one-sided slopes to the neighbours' centroid-face midpoints and the centred
// slope; flat at a local extremum
MDS_DEV void lin_recon_x64(double qm, double q0, double qp, const ReconGeo64 &g, double &al, double &ar)
{
  const double xm = g.bm / g.gm;
  const double xp = g.bp / g.gp;
  const double x0 = g.b0 / g.g0;
  const double sm = (q0 - qm) / (x0 - (0.5 * (g.xlo + xm)));
  const double sp = (qp - q0) / ((0.5 * (g.xhi + xp)) - x0);
  if (sm * sp <= 0.0) { al = q0; ar = q0; return; }
  const double sc = (qp - qm) / (xp - xm);
  const double asm_ = fabs(sm), asp = fabs(sp), asc = fabs(sc);
  const double m1 = asm_ < asp ? asm_ : asp;                // MIN of decs.h:627, NaN order included
  const double A = (sc < 0.0 ? -1.0 : 1.0) * (m1 < asc ? m1 : asc);
  const double B = q0 - ((A * g.b0) / g.g0);
  al = (A * g.xlo) + B;
  ar = (A * g.xhi) + B;
}

// the reconstruction of hydro_interp_order (reconstruct.c:246-266): every
// zone of a launch shares the order, so the branch is uniform
MDS_DEV void recon_x64(double qm, double q0, double qp, const ReconGeo64 &g, double &al, double &ar)
{
  if (g.order == 3) para_recon_x64(qm, q0, qp, g, al, ar);
  else if (g.order == 2) lin_recon_x64(qm, q0, qp, g, al, ar);
  else { al = q0; ar = q0; }
}

// The quotients of Gamma alone in the wave-speed estimate, for a gamma law
// (G_l = G_r = G, so G_avg is G itself).  One set per launch, in scalar
// registers: a quotient of the same operands has the same bits wherever it
// is formed, so the faces' results do not change.
struct GamF64 {
  double ravg;  // (G - 1) / (2 G)
  double pex;   // (2 G) / (G - 1)
  double bgm;   // (G - 1) / (G + 1)
  double kq;    // (G + 1) / (2 G)
};
MDS_DEV double uniform_f64(double x)
{
  const long long v = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
MDS_DEV GamF64 gam_f64_uniform(double G)
{
  GamF64 k;
  k.ravg = uniform_f64((G - 1.0) / (2.0 * G));
  k.pex = uniform_f64((2.0 * G) / (G - 1.0));
  k.bgm = uniform_f64((G - 1.0) / (G + 1.0));
  k.kq = uniform_f64((G + 1.0) / (2.0 * G));
  return k;
}

// wave-speed estimate shared by HLLC/HLLE (riemann.c:210-303, 368-429);
// UNIG: G_l = G_r, with its quotients in *gk (a gamma law)
template <bool HLLC_FLOOR, bool UNIG = false>
MDS_DEV void wave_speeds_x64(double rl, double rr, double p_l, double p_r, double G_l, double G_r,
                             double u_l, double u_r, double cs_l, double cs_r,
                             double &sl, double &sr, double &ss, double &p_s, int &isshock,
                             const GamF64 *gk = nullptr)
{
  double p_max, p_min;
  if (p_l > p_r) { p_max = p_l; p_min = p_r; } else { p_max = p_r; p_min = p_l; }
  double pratio = p_max / p_min;
  isshock = ((p_max - p_min) / p_min > 0.5) ? 1 : 0;
  p_s = (0.5 * (p_l + p_r)) - (((0.125 * (u_r - u_l)) * (rl + rr)) * (cs_l + cs_r));
  if (HLLC_FLOOR) p_s = (0 > p_s) ? 0.0 : p_s;
  else p_s = (p_s > 0.0) ? p_s : 0.0;
  if (pratio < 2.0 && p_s >= p_min && p_s <= p_max) {
  } else if (p_s < p_min) {
    double G_avg = 0.5 * (G_l + G_r);
    double plr = pow(p_l / p_r, UNIG ? gk->ravg : (G_avg - 1.0) / (2.0 * G_avg));
    double s0 = ((((plr * u_l) / cs_l) + (u_r / cs_r)) + ((2.0 * (plr - 1.0)) / (G_avg - 1.0)))
                / ((plr / cs_l) + (1.0 / cs_r));
    const double el = UNIG ? gk->pex : (2.0 * G_l) / (G_l - 1.0);
    const double er = UNIG ? gk->pex : (2.0 * G_r) / (G_r - 1.0);
    p_s = 0.5 * ((p_l * pow(1.0 + (((G_l - 1.0) / (2.0 * cs_l)) * (u_l - s0)), el))
                 + (p_r * pow(1.0 + (((G_r - 1.0) / (2.0 * cs_r)) * (s0 - u_r)), er)));
    p_s = (p_s > 0.0) ? p_s : 0.0;
  } else {
    double a_l = 2.0 / ((G_l + 1.0) * rl);
    double a_r = 2.0 / ((G_r + 1.0) * rr);
    double b_l = (UNIG ? gk->bgm : (G_l - 1.0) / (G_l + 1.0)) * p_l;
    double b_r = (UNIG ? gk->bgm : (G_r - 1.0) / (G_r + 1.0)) * p_r;
    double g_l = sqrt(a_l / (p_s + b_l));
    double g_r = sqrt(a_r / (p_s + b_r));
    p_s = (((g_l * p_l) + (g_r * p_r)) - (u_r - u_l)) / (g_l + g_r);
    p_s = (p_s > 0.0) ? p_s : 0.0;
  }
  const double kl = UNIG ? gk->kq : (G_l + 1.0) / (2.0 * G_l);
  const double kr = UNIG ? gk->kq : (G_r + 1.0) / (2.0 * G_r);
  double q_l = (p_s <= p_l) ? 1.0 : sqrt(1.0 + (kl * ((p_s / p_l) - 1.0)));
  double q_r = (p_s <= p_r) ? 1.0 : sqrt(1.0 + (kr * ((p_s / p_r) - 1.0)));
  sl = u_l - (cs_l * q_l);
  sr = u_r + (cs_r * q_r);
  ss = (((p_r - p_l) + ((rl * u_l) * (sl - u_l))) - ((rr * u_r) * (sr - u_r)))
       / ((rl * (sl - u_l)) - (rr * (sr - u_r)));
}

// conserved state and physical flux of one side (riemann.c:4-37)
template <int NV, bool CART, int NS = 0>
MDS_DEV void p_uflux_x64(const PState<double, NV> &p, const FaceGeo<double> &g, int dir,
                         double u[6], double f[6])
{
  double vcov[3];
  for (int d = 0; d < NV; d++) vcov[d] = CART ? p.v[d] : g.gcov[d] * p.v[d];
  double v = p.v[dir];
  u[0] = p.rho;
  double vsq = 0.0;
  for (int d = 0; d < NV; d++) vsq += p.v[d] * vcov[d];
  double etot = ((0.5 * u[0]) * vsq) + p.uu;
  u[1] = etot;
  for (int d = 0; d < NV; d++) u[2 + d] = u[0] * vcov[d];
  if (NS) u[MDS_YI] = u[0] * p.ye;
  f[0] = u[0] * v;
  f[1] = v * (etot + p.P);
  for (int d = 0; d < NV; d++) f[2 + d] = f[0] * vcov[d];
  f[2 + dir] += p.P;
  if (NS) f[MDS_YI] = f[0] * p.ye;
}

// area-weighted physical flux (riemann.c:39-62)
template <int NV, bool CART, int NS = 0>
MDS_DEV void pflux_x64(const PState<double, NV> &p, const FaceGeo<double> &g, int dir, double f[6])
{
  double vcov[3];
  for (int d = 0; d < NV; d++) vcov[d] = CART ? p.v[d] : g.gcov[d] * p.v[d];
  double v = p.v[dir];
  f[0] = (g.area * p.rho) * v;
  for (int d = 0; d < NV; d++) f[2 + d] = f[0] * vcov[d];
  f[2 + dir] += g.area * p.P;
  double vsq = 0.0;
  for (int d = 0; d < NV; d++) vsq += p.v[d] * vcov[d];
  double etot = ((0.5 * p.rho) * vsq) + p.uu;
  f[1] = (g.area * v) * (etot + p.P);
  if (NS) f[MDS_YI] = f[0] * p.ye;
}

// HLLC (riemann.c:328-499).  Returns isshock.
template <int NV, bool CART, int NS = 0, bool UNIG = false>
MDS_DEV int hllc_x64(const PState<double, NV> &L, const PState<double, NV> &Rt,
                     const FaceGeo<double> &g, int dir, double flux[6], double &sig, const GamF64 *gk = nullptr)
{
  if (g.area < 1.0e-16) {
    sig = 1.0e-200;
    for (int q = 0; q < 2 + NV; q++) flux[q] = 0.0;
    if (NS) flux[MDS_YI] = 0.0;
    return 0;
  }
  double u_l = CART ? L.v[dir] : L.v[dir] * g.scale;
  double u_r = CART ? Rt.v[dir] : Rt.v[dir] * g.scale;
  double cs_l = sqrt((L.G * L.P) / L.rho);
  double cs_r = sqrt((Rt.G * Rt.P) / Rt.rho);
  double sl, sr, ss, p_s;
  int isshock;
  wave_speeds_x64<true, UNIG>(L.rho, Rt.rho, L.P, Rt.P, L.G, Rt.G, u_l, u_r, cs_l, cs_r, sl, sr, ss, p_s, isshock, gk);
  if (ss > sr) sr = 1.001 * ss;
  if (ss < sl) sl = 1.001 * ss;
  double smax = (fabs(sl) > fabs(sr)) ? fabs(sl) : fabs(sr);
  sig = CART ? smax : smax / g.scale;
  if (sl > 0.0) { pflux_x64<NV, CART, NS>(L, g, dir, flux); return isshock; }
  if (sr < 0.0) { pflux_x64<NV, CART, NS>(Rt, g, dir, flux); return isshock; }
  double u[6], fx[6], Us[6], vs[3];
  double slr;
  // the side upwind of the contact, selected by value (a reference to one of
  // two locals would force both into scratch memory)
  const PState<double, NV> K = pick_state(ss > 0.0, L, Rt);
  if (ss > 0.0) {
    p_uflux_x64<NV, CART, NS>(L, g, dir, u, fx);
    Us[0] = (L.rho * (sl - u_l)) / (sl - ss);
    Us[1] = (((u[1] * (sl - u_l)) - (L.P * u_l)) + (p_s * ss)) / (sl - ss);
    slr = CART ? sl : sl / g.scale;
  } else {
    p_uflux_x64<NV, CART, NS>(Rt, g, dir, u, fx);
    Us[0] = (Rt.rho * (sr - u_r)) / (sr - ss);
    Us[1] = (((u[1] * (u_r - sr)) + (Rt.P * u_r)) - (p_s * ss)) / (ss - sr);
    slr = CART ? sr : sr / g.scale;
  }
  for (int d = 0; d < NV; d++) vs[d] = K.v[d];
  vs[dir] = CART ? ss : ss / g.scale;
  for (int d = 0; d < NV; d++) Us[2 + d] = (CART ? vs[d] : g.gcov[d] * vs[d]) * Us[0];
  if (NS) Us[MDS_YI] = Us[0] * K.ye;
  for (int q = 0; q < 2 + NV; q++) flux[q] = g.area * (fx[q] + (slr * (Us[q] - u[q])));
  if (NS) flux[MDS_YI] = g.area * (fx[MDS_YI] + (slr * (Us[MDS_YI] - u[MDS_YI])));
  return isshock;
}

// HLLE (riemann.c:178-325)
template <int NV, bool CART, int NS = 0, bool UNIG = false>
MDS_DEV void hlle_x64(const PState<double, NV> &L, const PState<double, NV> &Rt,
                      const FaceGeo<double> &g, int dir, double flux[6], double &sig, const GamF64 *gk = nullptr)
{
  if (g.area < 1.0e-16) {
    sig = 1.0e-200;
    for (int q = 0; q < 2 + NV; q++) flux[q] = 0.0;
    if (NS) flux[MDS_YI] = 0.0;
    return;
  }
  double u_l = CART ? L.v[dir] : L.v[dir] * g.scale;
  double u_r = CART ? Rt.v[dir] : Rt.v[dir] * g.scale;
  double cs_l = sqrt((L.G * L.P) / L.rho);
  double cs_r = sqrt((Rt.G * Rt.P) / Rt.rho);
  double sl, sr, ss, p_s;
  int isshock;
  wave_speeds_x64<false, UNIG>(L.rho, Rt.rho, L.P, Rt.P, L.G, Rt.G, u_l, u_r, cs_l, cs_r, sl, sr, ss, p_s, isshock, gk);
  double ul[6], ur[6], fl[6], fr[6];
  p_uflux_x64<NV, CART, NS>(L, g, dir, ul, fl);
  p_uflux_x64<NV, CART, NS>(Rt, g, dir, ur, fr);
  double srs = CART ? sr : sr / g.scale;
  double sls = CART ? sl : sl / g.scale;
  double cmax = (srs > 0.0) ? srs : 0.0;
  double cmin = fabs((sls < 0.0) ? sls : 0.0);
  sig = (cmax > cmin) ? cmax : cmin;
  for (int q = 0; q < 2 + NV; q++)
    flux[q] = (g.area * (((cmin * fr[q]) + (cmax * fl[q])) - ((cmax * cmin) * (ur[q] - ul[q])))) / (cmax + cmin);
  if (NS) {
    const int q = MDS_YI;
    flux[q] = (g.area * (((cmin * fr[q]) + (cmax * fl[q])) - ((cmax * cmin) * (ur[q] - ul[q])))) / (cmax + cmin);
  }
}

// fluid.c:4-25
template <int NV, bool CART>
MDS_DEV void p_to_u_x64(double rho, double uu, const double v[3], const double gcov[3], double u[5])
{
  double vcov[3];
  for (int d = 0; d < NV; d++) vcov[d] = CART ? v[d] : gcov[d] * v[d];
  u[0] = rho;
  double vsq = 0.0;
  for (int d = 0; d < NV; d++) {
    vsq += v[d] * vcov[d];
    u[2 + d] = u[0] * vcov[d];
  }
  u[1] = ((0.5 * u[0]) * vsq) + uu;
}

// fluid.c:27-53 (no floors)
template <int NV, bool CART>
MDS_DEV void u_to_p_x64(double u[5], const double gcon[3], double &rho, double &uu, double v[3])
{
  rho = u[0];
  for (int d = 0; d < NV; d++) u[2 + d] /= u[0];
  for (int d = 0; d < NV; d++) v[d] = CART ? u[2 + d] : gcon[d] * u[2 + d];
  double vsq = 0.0;
  for (int d = 0; d < NV; d++) vsq += u[2 + d] * v[d];
  uu = u[1] - ((0.5 * rho) * vsq);
}

// step.c:205-223
template <int NV>
MDS_DEV double dt_source_control_x64(double uu, const double v[3], const double u[5], const double s[5])
{
  const double alpha = 0.02;
  double mdt = 1.0e100;
  double local_dt = (alpha * u[0]) / (fabs(s[0]) + 1.0e-16);
  mdt = (mdt < local_dt) ? mdt : local_dt;
  local_dt = (alpha * u[1]) / (fabs(s[1]) + 1.0e-16);
  mdt = (mdt < local_dt) ? mdt : local_dt;
  double momterm = 0.0;
  for (int d = 0; d < NV; d++) momterm += v[d] * s[2 + d];
  local_dt = (2.0 * uu) / (fabs(fabs(s[1]) - fabs(momterm)) + 1.0e-16);
  mdt = (mdt < local_dt) ? mdt : local_dt;
  return mdt;
}

// dt_source_control with reciprocals instead of divisions (FP32-bulk path,
// where the limit only needs to be close to the reference value)
template <int NV>
MDS_DEV double dt_source_control_fast(double uu, const double v[3], const double u[5], const double s[5]);

// ---------------------------------------------------------------------------
// FP32-bulk
// ---------------------------------------------------------------------------
struct ReconGeo32 {
  float clm, clp, crm, crp, mu1, kinv, inur, imu2;
};

MDS_DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// hardware square root and reciprocal square root (v_sqrt_f32, v_rsq_f32,
// about 1 ulp): the FP32-bulk path needs no correctly rounded result, and the
// IEEE sequence costs some 16 VALU instructions per call.  Their arguments
// must stay in the FP32 normal range: the hardware flushes denormal inputs.
MDS_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
MDS_DEV float frsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// x^y for x > 0 as 2^(y log2 x) (v_log_f32, v_exp_f32): three instructions
// instead of the library powf's range reduction; the relative error grows to
// a few 1e-7 for |y log2 x| near 10, within what FP32 storage already makes
MDS_DEV float fpow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }

// para_recon in local coordinates: identical branch structure, well-conditioned.
MDS_DEV void para_recon_f32(float qm, float q0, float qp, const ReconGeo32 &g, float &al, float &ar)
{
#pragma clang fp contract(fast)
  float dqm = q0 - qm;
  float dqp = qp - q0;
  if (dqm * dqp <= 0.0f) { al = q0; ar = q0; return; }
  float l = q0 + g.clm * dqm + g.clp * dqp;
  float r = q0 + g.crm * dqm + g.crp * dqp;
  float mean = 0.75f * qm + 0.25f * q0;
  if ((q0 - l) * (l - mean) < 0.0f) l = mean;
  mean = 0.75f * qp + 0.25f * q0;
  if ((mean - r) * (r - q0) < 0.0f) r = mean;
  float b2 = (q0 - l - (r - l) * g.mu1) * g.kinv;
  float b1 = (r - l) - b2;
  // extremum sigma* = -b1/(2 b2) strictly inside (0,1)?
  float sz = -b1 * frcp(2.0f * b2);
  if (b2 != 0.0f && sz > 0.0f && sz < 1.0f) {
    if (sz > 0.5f) l = r + (q0 - r) * g.inur;
    else r = l + (q0 - l) * g.imu2;
  }
  al = l;
  ar = r;
}

// Gamma-only factors of the FP32 wave-speed estimate, each formed exactly as
// wave_speeds_f32 forms it from a side's Gamma_1 (with G_l = G_r = G, G_avg
// is G itself).  A gamma law needs one set per launch: gam_f32_uniform makes
// it once and keeps it in scalar registers, and the shock and rarefaction
// branches then take no reciprocal of Gamma per face.
struct GamF32 {
  float ravg;  // (G - 1) rcp(2 G): exponent of the two-rarefaction pressure ratio
  float igm1;  // rcp(G - 1)
  float hgm1;  // (G - 1) / 2
  float pex;   // 2 G rcp(G - 1): power of the two-rarefaction star pressure
  float gp1;   // G + 1
  float bgm;   // (G - 1) rcp(G + 1): the two-shock b / p
  float kq;    // (G + 1) rcp(2 G): the shock factor's slope
};
MDS_DEV float uniform_f32(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
MDS_DEV GamF32 gam_f32_uniform(float G)
{
  GamF32 k;
  k.ravg = uniform_f32((G - 1.0f) * frcp(2.0f * G));
  k.igm1 = uniform_f32(frcp(G - 1.0f));
  k.hgm1 = uniform_f32((G - 1.0f) * 0.5f);
  k.pex = uniform_f32(2.0f * G * frcp(G - 1.0f));
  k.gp1 = uniform_f32(G + 1.0f);
  k.bgm = uniform_f32((G - 1.0f) * frcp(G + 1.0f));
  k.kq = uniform_f32((G + 1.0f) * frcp(2.0f * G));
  return k;
}

// both precisions' sets for a gamma law (a kernel uses the one it needs)
struct GamSet {
  GamF32 f32;
  GamF64 f64;
};
MDS_DEV GamSet gam_set_uniform(double G)
{
  GamSet k;
  k.f32 = gam_f32_uniform((float)G);
  k.f64 = gam_f64_uniform(G);
  return k;
}

// EAGER: both pressure reciprocals up front, serving the ratio and both
// shock factors (faster where shock factors are common, as on Cartesian
// boxes); otherwise the shock factors divide only where a wavefront needs them.
// UNIG: G_l = G_r, with its factors in *gk (a gamma law).
template <bool HLLC_FLOOR, bool EAGER = false, bool UNIG = false>
MDS_DEV void wave_speeds_f32(float rl, float rr, float p_l, float p_r, float G_l, float G_r,
                             float u_l, float u_r, float cs_l, float cs_r,
                             float &sl, float &sr, float &ss, float &p_s, int &isshock,
                             const GamF32 *gk = nullptr)
{
#pragma clang fp contract(fast)
  float p_max = fmaxf(p_l, p_r), p_min = fminf(p_l, p_r);
  float ipl = 0.0f, ipr = 0.0f, ipmin;
  if constexpr (EAGER) {
    ipl = frcp(p_l);
    ipr = frcp(p_r);
    ipmin = p_l <= p_r ? ipl : ipr;
  } else {
    ipmin = frcp(p_min);
  }
  float pratio = p_max * ipmin;
  isshock = ((p_max - p_min) * ipmin > 0.5f) ? 1 : 0;
  p_s = 0.5f * (p_l + p_r) - 0.125f * (u_r - u_l) * (rl + rr) * (cs_l + cs_r);
  p_s = fmaxf(p_s, 0.0f);
  if (pratio < 2.0f && p_s >= p_min && p_s <= p_max) {
  } else if (p_s < p_min) {
    float G_avg = 0.5f * (G_l + G_r);
    const float ex = UNIG ? gk->ravg : (G_avg - 1.0f) * frcp(2.0f * G_avg);
    float plr = fpow(p_l * (EAGER ? ipr : frcp(p_r)), ex);
    float icl = frcp(cs_l), icr = frcp(cs_r);
    const float igm1 = UNIG ? gk->igm1 : frcp(G_avg - 1.0f);
    float s0 = (plr * u_l * icl + u_r * icr + 2.0f * (plr - 1.0f) * igm1) * frcp(plr * icl + icr);
    const float hl = UNIG ? gk->hgm1 : (G_l - 1.0f) * 0.5f, hr = UNIG ? gk->hgm1 : (G_r - 1.0f) * 0.5f;
    const float el = UNIG ? gk->pex : 2.0f * G_l * frcp(G_l - 1.0f);
    const float er = UNIG ? gk->pex : 2.0f * G_r * frcp(G_r - 1.0f);
    p_s = 0.5f * (p_l * fpow(1.0f + hl * icl * (u_l - s0), el) + p_r * fpow(1.0f + hr * icr * (s0 - u_r), er));
    p_s = fmaxf(p_s, 0.0f);
  } else {
    float a_l = 2.0f * frcp((UNIG ? gk->gp1 : G_l + 1.0f) * rl);
    float a_r = 2.0f * frcp((UNIG ? gk->gp1 : G_r + 1.0f) * rr);
    float b_l = (UNIG ? gk->bgm : (G_l - 1.0f) * frcp(G_l + 1.0f)) * p_l;
    float b_r = (UNIG ? gk->bgm : (G_r - 1.0f) * frcp(G_r + 1.0f)) * p_r;
    // sqrt(a / (p + b)) as sqrt(a) / sqrt(p + b): the quotient itself falls
    // below the FP32 normal range at stellar-core densities and pressures
    // (2 / (2.3 rho) / (p + b) near 1e-38 at rho 1e10, p 1e28), where the
    // hardware square root flushes it to zero
    float g_l = fsqrt(a_l) * frsqrt(p_s + b_l);
    float g_r = fsqrt(a_r) * frsqrt(p_s + b_r);
    p_s = (g_l * p_l + g_r * p_r - (u_r - u_l)) * frcp(g_l + g_r);
    p_s = fmaxf(p_s, 0.0f);
  }
  const float kl = UNIG ? gk->kq : (G_l + 1.0f) * frcp(2.0f * G_l);
  const float kr = UNIG ? gk->kq : (G_r + 1.0f) * frcp(2.0f * G_r);
  float q_l = (p_s <= p_l) ? 1.0f : fsqrt(1.0f + kl * (p_s * (EAGER ? ipl : frcp(p_l)) - 1.0f));
  float q_r = (p_s <= p_r) ? 1.0f : fsqrt(1.0f + kr * (p_s * (EAGER ? ipr : frcp(p_r)) - 1.0f));
  sl = u_l - cs_l * q_l;
  sr = u_r + cs_r * q_r;
  ss = (p_r - p_l + rl * u_l * (sl - u_l) - rr * u_r * (sr - u_r)) * frcp(rl * (sl - u_l) - rr * (sr - u_r));
}

// Per-unit-area fluxes in FP32 (the FP64 caller applies the FP64 area).
template <int NV, bool CART, int NS = 0>
MDS_DEV void pflux_f32(const PState<float, NV> &p, const FaceGeo<float> &g, int dir, float f[6])
{
#pragma clang fp contract(fast)
  float vcov[3];
  for (int d = 0; d < NV; d++) vcov[d] = CART ? p.v[d] : g.gcov[d] * p.v[d];
  float v = p.v[dir];
  f[0] = p.rho * v;
  for (int d = 0; d < NV; d++) f[2 + d] = f[0] * vcov[d];
  f[2 + dir] += p.P;
  float vsq = 0.0f;
  for (int d = 0; d < NV; d++) vsq += p.v[d] * vcov[d];
  f[1] = v * (0.5f * p.rho * vsq + p.uu + p.P);
  if (NS) f[MDS_YI] = f[0] * p.ye;
}

template <int NV, bool CART, int NS = 0>
MDS_DEV void p_uflux_f32(const PState<float, NV> &p, const FaceGeo<float> &g, int dir, float u[6], float f[6])
{
#pragma clang fp contract(fast)
  float vcov[3];
  for (int d = 0; d < NV; d++) vcov[d] = CART ? p.v[d] : g.gcov[d] * p.v[d];
  float v = p.v[dir];
  u[0] = p.rho;
  float vsq = 0.0f;
  for (int d = 0; d < NV; d++) vsq += p.v[d] * vcov[d];
  float etot = 0.5f * p.rho * vsq + p.uu;
  u[1] = etot;
  for (int d = 0; d < NV; d++) u[2 + d] = p.rho * vcov[d];
  f[0] = p.rho * v;
  f[1] = v * (etot + p.P);
  for (int d = 0; d < NV; d++) f[2 + d] = f[0] * vcov[d];
  f[2 + dir] += p.P;
  if (NS) { u[MDS_YI] = p.rho * p.ye; f[MDS_YI] = f[0] * p.ye; }
}

// HLLC in FP32; flux per unit area.  Returns isshock.  UNIG: a gamma law,
// whose wave-speed factors *gk carries.
template <int NV, bool CART, int NS = 0, bool UNIG = false>
MDS_DEV int hllc_f32(const PState<float, NV> &L, const PState<float, NV> &Rt,
                     const FaceGeo<float> &g, int dir, float flux[6], float &sig, const GamF32 *gk = nullptr)
{
#pragma clang fp contract(fast)
  float u_l = CART ? L.v[dir] : L.v[dir] * g.scale;
  float u_r = CART ? Rt.v[dir] : Rt.v[dir] * g.scale;
  float cs_l = fsqrt(L.G * L.P * frcp(L.rho));
  float cs_r = fsqrt(Rt.G * Rt.P * frcp(Rt.rho));
  float sl, sr, ss, p_s;
  int isshock;
  wave_speeds_f32<true, CART, UNIG>(L.rho, Rt.rho, L.P, Rt.P, L.G, Rt.G, u_l, u_r, cs_l, cs_r, sl, sr, ss, p_s,
                                    isshock, gk);
  if (ss > sr) sr = 1.001f * ss;
  if (ss < sl) sl = 1.001f * ss;
  float iscale = CART ? 1.0f : frcp(g.scale);
  sig = fmaxf(fabsf(sl), fabsf(sr)) * iscale;
  // One evaluation for every face, no divergent exits: the physical flux of
  // the upwind state (the left state where sl > 0, the right where sr < 0,
  // else the side of the contact), plus the star-state term where neither
  // wave is supersonic (riemann.c:437-495).  The term is selected, not
  // weighted, so an undefined star state of a supersonic face stays out.
  const bool sup_l = sl > 0.0f, sup_r = sr < 0.0f, star = !(sup_l || sup_r);
  const bool left = sup_l || (!sup_r && ss > 0.0f);
  float u[6], fx[6], Us[6];
  const PState<float, NV> K = pick_state(left, L, Rt);
  float sk = left ? sl : sr, uk = left ? u_l : u_r;
  p_uflux_f32<NV, CART>(K, g, dir, u, fx);
  float iden = frcp(sk - ss);
  Us[0] = K.rho * (sk - uk) * iden;
  Us[1] = (u[1] * (sk - uk) - K.P * uk + p_s * ss) * iden;
  float slr = sk * iscale;
  for (int d = 0; d < NV; d++) {
    float vs = (d == dir) ? ss * iscale : K.v[d];
    Us[2 + d] = (CART ? vs : g.gcov[d] * vs) * Us[0];
  }
  for (int q = 0; q < 2 + NV; q++) flux[q] = star ? fx[q] + slr * (Us[q] - u[q]) : fx[q];
  if (NS) flux[MDS_YI] = flux[0] * K.ye;     // passive scalar: upwind by the contact
  return isshock;
}

template <int NV, bool CART, int NS = 0, bool UNIG = false>
MDS_DEV void hlle_f32(const PState<float, NV> &L, const PState<float, NV> &Rt,
                      const FaceGeo<float> &g, int dir, float flux[6], float &sig, const GamF32 *gk = nullptr)
{
#pragma clang fp contract(fast)
  float u_l = CART ? L.v[dir] : L.v[dir] * g.scale;
  float u_r = CART ? Rt.v[dir] : Rt.v[dir] * g.scale;
  float cs_l = fsqrt(L.G * L.P * frcp(L.rho));
  float cs_r = fsqrt(Rt.G * Rt.P * frcp(Rt.rho));
  float sl, sr, ss, p_s;
  int isshock;
  wave_speeds_f32<false, CART, UNIG>(L.rho, Rt.rho, L.P, Rt.P, L.G, Rt.G, u_l, u_r, cs_l, cs_r, sl, sr, ss, p_s,
                                     isshock, gk);
  float ul[6], ur[6], fl[6], fr[6];
  p_uflux_f32<NV, CART, NS>(L, g, dir, ul, fl);
  p_uflux_f32<NV, CART, NS>(Rt, g, dir, ur, fr);
  float iscale = CART ? 1.0f : frcp(g.scale);
  float cmax = fmaxf(sr * iscale, 0.0f);
  float cmin = fabsf(fminf(sl * iscale, 0.0f));
  sig = fmaxf(cmax, cmin);
  float iden = frcp(cmax + cmin);
  for (int q = 0; q < 2 + NV; q++) flux[q] = (cmin * fr[q] + cmax * fl[q] - cmax * cmin * (ur[q] - ul[q])) * iden;
  if (NS) {
    const int q = MDS_YI;
    flux[q] = (cmin * fr[q] + cmax * fl[q] - cmax * cmin * (ur[q] - ul[q])) * iden;
  }
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// positive double -> order-preserving 64-bit key for atomicMin
MDS_DEV unsigned long long dt_key(double v) { return (unsigned long long)__double_as_longlong(v); }

// Running minimum keys shared by a whole launch.  One same-address atomic per
// wavefront serializes at the L2 and can cost more than the kernel's own
// work, so a wavefront reads the current minimum (device-coherent load, the
// L2 copy) when it starts and issues its atomic only if its value beats it.
// A stale read is larger than the true minimum, so no needed update is lost.
MDS_DEV unsigned long long min_key_seen(unsigned long long *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
MDS_DEV void fold_min_key(unsigned long long *p, unsigned long long seen, double v)
{
  const unsigned long long k = dt_key(v);
  if (k < seen) atomicMin(p, k);
}

// Device health: lanes of the wavefront whose new state is unphysical --
// a non-finite or non-positive density, a non-finite internal energy, and
// with the gamma law a negative one (negative pressure).  Call with every
// lane of the wavefront active; returns the wave-uniform count.
MDS_DEV unsigned bad_state_lanes(bool valid, double rho, double uu, bool gamma_law)
{
  const double inf = __builtin_huge_val();
  const bool ok = (rho > 0.0) && (rho < inf) && (uu > -inf) && (uu < inf) && (!gamma_law || uu >= 0.0);
  return (unsigned)__popcll(__ballot(valid && !ok));
}

// Cross-lane neighbours with whole-wave DPP shifts (GFX9 wave_shr:1 /
// wave_shl:1): lane l reads lane l-1 (from_below) or l+1 (from_above); the
// edge lane keeps its own value.  Must be called with every lane active.
MDS_DEV int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
MDS_DEV int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
MDS_DEV float lane_from_below(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
MDS_DEV float lane_from_above(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
MDS_DEV int lane_from_below(int v) { return dpp_shr1(v); }
MDS_DEV int lane_from_above(int v) { return dpp_shl1(v); }
MDS_DEV double lane_from_below(double v)
{
  long long b = __double_as_longlong(v);
  int lo = dpp_shr1((int)(b & 0xffffffffll)), hi = dpp_shr1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
MDS_DEV double lane_from_above(double v)
{
  long long b = __double_as_longlong(v);
  int lo = dpp_shl1((int)(b & 0xffffffffll)), hi = dpp_shl1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/x to full double precision: hardware estimate + one Newton step
MDS_DEV double drcp(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

MDS_DEV double wave_max(double v)
{
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = (o > v) ? o : v;
  }
  return v;
}

MDS_DEV double wave_min(double v)
{
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  return v;
}

// The FP32-bulk source limit: hardware reciprocal estimates (v_rcp_f64, no
// Newton steps).  The limit bounds dt; a relative error near 1e-8 in it is
// far below what FP32 state storage already changes.
template <int NV>
MDS_DEV double dt_source_control_fast(double uu, const double v[3], const double u[5], const double s[5])
{
  double mdt = (0.02 * u[0]) * __builtin_amdgcn_rcp(fabs(s[0]) + 1.0e-16);
  double t = (0.02 * u[1]) * __builtin_amdgcn_rcp(fabs(s[1]) + 1.0e-16);
  mdt = fmin(mdt, t);
  double momterm = 0.0;
  for (int d = 0; d < NV; d++) momterm += v[d] * s[2 + d];
  t = (2.0 * uu) * __builtin_amdgcn_rcp(fabs(fabs(s[1]) - fabs(momterm)) + 1.0e-16);
  return fmin(mdt, t);
}
