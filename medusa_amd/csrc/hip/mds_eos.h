// mds_eos.h -- device side of the tabulated EOS (gfx950).
//
// Restates the reference's table use, eos/burrows/eos_stuff.c:507-650: the
// bilinear (log10 rho, Ye) weights, the temperature bracket of the specific
// internal energy, the linear inverse in log T and the trilinear ln P and
// Gamma_1.  The bracket is found by bisection over the T nodes instead of the
// reference's walk from the previous temperature: e is strictly increasing in
// T, so both give the unique node interval with E(j) < e <= E(j+1).  Where a
// zone's previous bracket is kept (the 3D spherical stage), the search starts
// there and bisects only when the bracket moved by more than one node.  All of
// the inversion is FP64; the FP32-bulk variant interpolates ln P increments
// and Gamma_1 in FP32.
#pragma once
#include "mds_internal.h"
#include "mds_physics.h"

// bilinear weights in (log10 rho, Ye) and the four T rows they touch
struct EosW {
  int o00, o01, o10, o11;        // (jy,jr), (jyp,jr), (jy,jrp), (jyp,jrp)
  double c0, c1, c2, c3;
};

MDS_DEV EosW eos_weights(const DevEos &t, double rho, double ye)
{
  EosW w;
  // clamp to the table (a NaN lands on the first node: the indices below
  // stay inside the table whatever the state)
  double rl = log10(rho);
  rl = !(rl >= t.lr0) ? t.lr0 : (rl > t.lr1 ? t.lr1 : rl);
  ye = !(ye >= t.ye0) ? t.ye0 : (ye > t.ye1 ? t.ye1 : ye);
  int jr = (int)((rl - t.lr0) / t.dlr);
  int jy = (int)((ye - t.ye0) / t.dye);
  jr = jr > t.nr - 1 ? t.nr - 1 : (jr < 0 ? 0 : jr);
  jy = jy > t.ny - 1 ? t.ny - 1 : (jy < 0 ? 0 : jy);
  int jyp = jy + 1, jrp = jr + 1;
  double drho = (rl - (t.lr0 + jr * t.dlr)) / t.dlr;
  double dye = (ye - (t.ye0 + jy * t.dye)) / t.dye;
  if (jy == t.ny - 1) { jyp = jy; dye = 0.0; }
  if (jr == t.nr - 1) { jrp = jr; drho = 0.0; }
  w.c0 = (1.0 - drho) * (1.0 - dye);
  w.c1 = dye * (1.0 - drho);
  w.c2 = (1.0 - dye) * drho;
  w.c3 = dye * drho;
  w.o00 = (jy * t.nr + jr) * t.nt;
  w.o01 = (jyp * t.nr + jr) * t.nt;
  w.o10 = (jy * t.nr + jrp) * t.nt;
  w.o11 = (jyp * t.nr + jrp) * t.nt;
  return w;
}

// energy projected onto T node jt (eos_stuff.c:566-569)
MDS_DEV double eos_e_at(const DevEos &t, const EosW &w, int jt)
{
  double e = w.c0 * (double)t.e[w.o00 + jt];
  e = e + w.c1 * (double)t.e[w.o01 + jt];
  e = e + w.c2 * (double)t.e[w.o10 + jt];
  e = e + w.c3 * (double)t.e[w.o11 + jt];
  return e;
}

// temperature bracket and linear inverse (eos_stuff.c:562-605):
// jt = largest node in [0, nt-2] with E(jt) < e (0 if none)
MDS_DEV void eos_invert(const DevEos &t, const EosW &w, double e, int &jt, double &dtemp, double &lt)
{
  int lo = 0, hi = t.nt - 1;
  double elo = eos_e_at(t, w, 0);
  if (elo < e) {
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      const double em = eos_e_at(t, w, mid);
      if (em < e) { lo = mid; elo = em; }
      else hi = mid;
    }
  }
  jt = lo;
  const double ehi = eos_e_at(t, w, jt + 1);
  const double tlo = t.lt0 + jt * t.dlt;
  const double thi = tlo + t.dlt;
  dtemp = (e - elo) / (ehi - elo);
  lt = tlo + (thi - tlo) * dtemp;
}

// eos_invert from a guessed bracket g (the zone's bracket of the previous
// call, as the reference walks from the previous temperature): E(g) and
// E(g+1) settle it, one node either side is tried next, and the bisection
// settles anything else.  E is strictly increasing in T, so every route ends
// on the same interval, with the same elo and ehi.  g returns the bracket.
MDS_DEV void eos_invert_guess(const DevEos &t, const EosW &w, double e, int &g, double &dtemp, double &lt)
{
  const int top = t.nt - 2;
  int c = g < 0 ? 0 : (g > top ? top : g);
  const double eg = eos_e_at(t, w, c), eg1 = eos_e_at(t, w, c + 1);
  int jt = -1;
  double elo = 0.0, ehi = 0.0;
  if (eg < e) {
    if (c == top || eg1 >= e) {
      jt = c; elo = eg; ehi = eg1;
    } else {                                   // E(c+1) < e, c+1 <= top
      const double eg2 = eos_e_at(t, w, c + 2);
      if (c + 1 == top || eg2 >= e) { jt = c + 1; elo = eg1; ehi = eg2; }
    }
  } else if (c == 0) {                         // E(0) >= e
    jt = 0; elo = eg; ehi = eg1;
  } else {                                     // E(c) >= e
    const double em = eos_e_at(t, w, c - 1);
    if (em < e || c == 1) { jt = c - 1; elo = em; ehi = eg; }
  }
  if (jt < 0) {
    eos_invert(t, w, e, g, dtemp, lt);
    return;
  }
  g = jt;
  const double tlo = t.lt0 + jt * t.dlt;
  const double thi = tlo + t.dlt;
  dtemp = (e - elo) / (ehi - elo);
  lt = tlo + (thi - tlo) * dtemp;
}

// trilinear record at (jt, dtemp), reference summation order (eos_stuff.c:616-643)
MDS_DEV double eos_tri_x64(const float *rec, const EosW &w, int jt, double dtemp)
{
  const int jtp = jt + 1;
  double p = w.c0 * (1 - dtemp) * (double)rec[w.o00 + jt];
  p += w.c1 * (1 - dtemp) * (double)rec[w.o01 + jt];
  p += w.c2 * (1 - dtemp) * (double)rec[w.o10 + jt];
  p += w.c3 * (1 - dtemp) * (double)rec[w.o11 + jt];
  p += w.c0 * dtemp * (double)rec[w.o00 + jtp];
  p += w.c1 * dtemp * (double)rec[w.o01 + jtp];
  p += w.c2 * dtemp * (double)rec[w.o10 + jtp];
  p += w.c3 * dtemp * (double)rec[w.o11 + jtp];
  return p;
}

// The state lies outside the table: density or Ye beyond its nodes (the
// weights clamp them to the edge) or a temperature beyond the T nodes (the
// linear inverse extrapolates).  NaN counts as outside.
MDS_DEV bool eos_off_table(const DevEos &t, double rho, double ye, double lt)
{
  return !((rho >= t.rmin) && (rho <= t.rmax) && (ye >= t.ye0) && (ye <= t.ye1) && (lt >= t.lt0) && (lt <= t.lt1));
}

// specific internal energy in table units (eos_stuff.c:530): e / e_scale - e_shift
MDS_DEV double eos_e_table(const DevEos &t, double rho, double u) { return (u / rho) / t.escale - t.eshift; }

// FP64 EOS fill of one zone: (rho, u, Ye) -> P, Gamma_1, cs, log10 T.  With
// guess non-null the bracket search starts from *guess and returns it there.
MDS_DEV void eos_fill_x64(const DevEos &t, double rho, double u, double ye, double &press, double &gam1,
                          double &cs, double &lt, int *guess = nullptr)
{
  const EosW w = eos_weights(t, rho, ye);
  int jt;
  double dtemp;
  if (guess) {
    eos_invert_guess(t, w, eos_e_table(t, rho, u), *guess, dtemp, lt);
    jt = *guess;
  } else {
    eos_invert(t, w, eos_e_table(t, rho, u), jt, dtemp, lt);
  }
  press = exp(eos_tri_x64(t.lnp, w, jt, dtemp)) * t.pscale;       // eos_stuff.c:618
  gam1 = eos_tri_x64(t.g1, w, jt, dtemp);
  cs = sqrt((gam1 * press) / rho);
}

// FP32-bulk interpolation: increments over the first corner in FP32
MDS_DEV float eos_tri_f32(const float *rec, const EosW &w, int jt, double dtemp, float &base)
{
#pragma clang fp contract(fast)
  const int jtp = jt + 1;
  const float a = (float)(1.0 - dtemp), b = (float)dtemp;
  const float c0 = (float)w.c0, c1 = (float)w.c1, c2 = (float)w.c2, c3 = (float)w.c3;
  base = rec[w.o00 + jt];
  float s = c1 * a * (rec[w.o01 + jt] - base);
  s += c2 * a * (rec[w.o10 + jt] - base);
  s += c3 * a * (rec[w.o11 + jt] - base);
  s += c0 * b * (rec[w.o00 + jtp] - base);
  s += c1 * b * (rec[w.o01 + jtp] - base);
  s += c2 * b * (rec[w.o10 + jtp] - base);
  s += c3 * b * (rec[w.o11 + jtp] - base);
  return s;
}

// FP32-bulk EOS fill: FP64 inversion, FP32 interpolation; P in FP64 from the
// FP32 increment of ln P over an exact table value.  guess as in eos_fill_x64.
MDS_DEV void eos_fill_f32(const DevEos &t, double rho, double u, double ye, double &press, float &gam1, float &cs,
                          double &lt, int *guess = nullptr)
{
  const EosW w = eos_weights(t, rho, ye);
  int jt;
  double dtemp;
  if (guess) {
    eos_invert_guess(t, w, eos_e_table(t, rho, u), *guess, dtemp, lt);
    jt = *guess;
  } else {
    eos_invert(t, w, eos_e_table(t, rho, u), jt, dtemp, lt);
  }
  float b0, b1;
  const float dl = eos_tri_f32(t.lnp, w, jt, dtemp, b0);
  press = exp((double)b0 + (double)dl) * t.pscale;
  gam1 = b1 = 0.0f;
  const float dg = eos_tri_f32(t.g1, w, jt, dtemp, b1);
  gam1 = b1 + dg;
  cs = __builtin_amdgcn_sqrtf(gam1 * (float)(press / rho));
}
