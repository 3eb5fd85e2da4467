/*
 * mds_eos_table.c -- synthetic tabulated EOS (host C).
 *
 * Model, per unit mass, at density rho, temperature T and electron fraction
 * Ye (cgs, extended precision):
 *   degenerate electrons  P_d = K53 (rho Ye)^(5/3) / (1 + x),  x = (rho Ye / rho_x)^(1/3)
 *                         e_d = P_d / (rho (g_d - 1)),  g_d = 5/3 - (1/3) x / (1 + x)
 *   thermal electrons     P_e = n_e k T * T / (T + T_F),  T_F = 3e9 K (rho Ye / 1e6)^(2/3) / (1 + x)
 *                         e_e = 1.5 P_e / rho
 *   ions                  P_i = rho k T / (A m_u),  e_i = 1.5 P_i / rho
 *   radiation             P_r = a T^4 / 3,  e_r = a T^4 / rho
 * P and e are sums of the four; e is strictly increasing in T.  Gamma_1 =
 * chi_rho + P chi_T^2 / (rho T c_v) from central differences of the model.
 * A, K53 and rho_x carry a seeded +-10%, +-5%, +-10% jitter.
 */
#define _FILE_OFFSET_BITS 64
#define _POSIX_C_SOURCE 200809L
#include <sys/types.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../../include/mds_hydro.h"
#include "../../../include/mds_host.h"
#include "../../../include/mds_eos.h"

void mds_eos_table_default(mds_eos_table *t)
{
  memset(t, 0, sizeof(*t));
  t->nr = 180;  t->lr0 = 3.0;  t->lr1 = 15.0;
  t->nt = 180;  t->lt0 = 9.0;  t->lt1 = 12.0;
  t->ny = 50;   t->ye0 = 0.05; t->ye1 = 0.55;
  t->e_scale = 1.0;
  t->e_shift = 0.0;
  t->p_scale = 1.0;
}

void mds_eos_burrows_default(mds_eos_table *t)
{
  memset(t, 0, sizeof(*t));
  t->nr = 300;  t->lr0 = 1.0;    t->lr1 = 15.5;
  t->nt = 300;  t->lt0 = -2.0;   t->lt1 = 2.0;
  t->ny = 50;   t->ye0 = 0.035;  t->ye1 = 0.56;
  t->e_scale = 0.95655684e18;    /* erg/g per MeV/baryon (eos_stuff.c:530) */
  t->e_shift = 0.0;              /* const_hack (eos_stuff.c:375) */
  t->p_scale = 1.60217733e33;    /* dyn/cm^2 per MeV/fm^3 (eos_stuff.c:618) */
}

/* collapse_init, eos_stuff.c:76-133: the records are read one at a time at
 * their byte offsets; record 2 is log-transformed after the read.  Only the
 * three records the sweep uses are kept. */
int mds_eos_table_load_burrows(const char *path, mds_eos_table *t)
{
  if (!path || !t || t->nr < 2 || t->nt < 3 || t->ny < 2) return MDS_EINVAL;
  const size_t rec = (size_t)t->ny * t->nr * t->nt;
  static const int want[MDS_EOS_NFIELD] = {1, 2, 12};     /* reference record numbers (1-based) */
  FILE *f = fopen(path, "rb");
  if (!f) return MDS_EIO;
  /* the file must hold all MDS_BURROWS_NREC records */
  if (fseeko(f, 0, SEEK_END) != 0 || ftello(f) < (off_t)MDS_BURROWS_NREC * 8 * (off_t)rec) {
    fclose(f);
    return MDS_EIO;
  }
  float *data = (float *)malloc(MDS_EOS_NFIELD * rec * sizeof(float));
  double *buf = (double *)malloc(rec * sizeof(double));
  int rc = (data && buf) ? MDS_OK : MDS_ENOMEM;
  for (int k = 0; k < MDS_EOS_NFIELD && rc == MDS_OK; k++) {
    const long long off = (long long)(want[k] - 1) * 8LL * (long long)rec;
    if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fread(buf, sizeof(double), rec, f) != rec) {
      rc = MDS_EIO;
      break;
    }
    float *dst = data + (size_t)k * rec;
    for (size_t q = 0; q < rec; q++) dst[q] = (float)(k == MDS_EOS_LNP ? log(buf[q]) : buf[q]);
  }
  fclose(f);
  free(buf);
  if (rc != MDS_OK) {
    free(data);
    return rc;
  }
  t->data = data;
  return MDS_OK;
}

void mds_eos_model_params(unsigned long long seed, double out[3])
{
  double u0 = 2.0 * mds_uniform(seed, 0) - 1.0;
  double u1 = 2.0 * mds_uniform(seed, 1) - 1.0;
  double u2 = 2.0 * mds_uniform(seed, 2) - 1.0;
  out[0] = 56.0 * (1.0 + 0.10 * u0);
  out[1] = 1.0036e13 * (1.0 + 0.05 * u1);
  out[2] = 2.0e6 * (1.0 + 0.10 * u2);
}

static void model(const double prm[3], long double rho, long double T, long double ye, long double *P,
                  long double *e)
{
  const long double kB = 1.380649e-16L, mu = 1.66053906660e-24L, arad = 7.565723e-15L;
  const long double A = prm[0], K53 = prm[1], rhox = prm[2];
  const long double rye = rho * ye;
  const long double x = cbrtl(rye / rhox);
  const long double Pd = K53 * powl(rye, 5.0L / 3.0L) / (1.0L + x);
  const long double gd = 5.0L / 3.0L - (x / (1.0L + x)) / 3.0L;
  const long double ed = Pd / (rho * (gd - 1.0L));
  const long double TF = 3.0e9L * powl(rye / 1.0e6L, 2.0L / 3.0L) / (1.0L + x);
  const long double ne = rye / mu;
  const long double Pe = ne * kB * T * (T / (T + TF));
  const long double Pi = rho * kB * T / (A * mu);
  const long double T4 = (T * T) * (T * T);
  const long double Pr = arad * T4 / 3.0L;
  *P = ((Pd + Pe) + Pi) + Pr;
  *e = ((ed + 1.5L * Pe / rho) + 1.5L * Pi / rho) + arad * T4 / rho;
}

void mds_eos_model_eval(const double prm[3], double rho, double temp, double ye, double out[3])
{
  long double P, e, Pp, Pm, ep, em, d;
  const long double h = 1.0e-5L;
  model(prm, rho, temp, ye, &P, &e);
  // chi_rho = dlnP/dlnrho at fixed T
  model(prm, rho * expl(h), temp, ye, &Pp, &d);
  model(prm, rho * expl(-h), temp, ye, &Pm, &d);
  const long double chr = (logl(Pp) - logl(Pm)) / (2.0L * h);
  // chi_T = dlnP/dlnT and c_v = de/dT at fixed rho
  model(prm, rho, temp * expl(h), ye, &Pp, &ep);
  model(prm, rho, temp * expl(-h), ye, &Pm, &em);
  const long double cht = (logl(Pp) - logl(Pm)) / (2.0L * h);
  const long double cv = (ep - em) / ((long double)temp * (expl(h) - expl(-h)));
  const long double g1 = chr + P * cht * cht / ((long double)rho * (long double)temp * cv);
  out[0] = (double)P;
  out[1] = (double)e;
  out[2] = (double)g1;
}

int mds_eos_table_build(mds_eos_table *t, unsigned long long seed)
{
  if (!t || t->nr < 2 || t->nt < 3 || t->ny < 2) return MDS_EINVAL;
  if (!(t->lr1 > t->lr0) || !(t->lt1 > t->lt0) || !(t->ye1 > t->ye0)) return MDS_EINVAL;
  const size_t rec = (size_t)t->ny * t->nr * t->nt;
  t->data = (float *)malloc(MDS_EOS_NFIELD * rec * sizeof(float));
  if (!t->data) return MDS_ENOMEM;
  double prm[3];
  mds_eos_model_params(seed, prm);
  const double dlr = (t->lr1 - t->lr0) / (t->nr - 1);
  const double dlt = (t->lt1 - t->lt0) / (t->nt - 1);
  const double dye = (t->ye1 - t->ye0) / (t->ny - 1);
  const int ny = t->ny, nr = t->nr, nt = t->nt;
  float *data = t->data;
#pragma omp parallel for schedule(static) collapse(2)
  for (int jy = 0; jy < ny; jy++) {
    for (int jr = 0; jr < nr; jr++) {
      const double ye = t->ye0 + jy * dye;
      const double rho = pow(10.0, t->lr0 + jr * dlr);
      for (int jt = 0; jt < nt; jt++) {
        const double temp = pow(10.0, t->lt0 + jt * dlt);
        double v[3];
        mds_eos_model_eval(prm, rho, temp, ye, v);
        const size_t q = ((size_t)jy * nr + jr) * nt + jt;
        data[MDS_EOS_E * rec + q] = (float)v[1];
        data[MDS_EOS_LNP * rec + q] = (float)log(v[0]);
        data[MDS_EOS_GAM1 * rec + q] = (float)v[2];
      }
    }
  }
  return MDS_OK;
}

void mds_eos_table_free(mds_eos_table *t)
{
  if (!t) return;
  free(t->data);
  t->data = NULL;
}

/* eos_given_rtx / findthis (eos_stuff.c:253-313, 1189-1215): node indices
 * and fractions along each axis, the eight trilinear weights, and each
 * record summed in the reference's corner order */
int mds_eos_rtx(const mds_eos_table *t, double rho, double temp, double ye, double out[4])
{
  if (!t || !t->data || !out || !(rho > 0.0) || !(temp > 0.0) || t->nr < 2 || t->nt < 2 || t->ny < 2)
    return MDS_EINVAL;
  const double dr = (t->lr1 - t->lr0) / (t->nr - 1);
  const double dtt = (t->lt1 - t->lt0) / (t->nt - 1);
  const double dy = (t->ye1 - t->ye0) / (t->ny - 1);
  double rl = log10(rho), tl = log10(temp);
  rl = rl < t->lr0 ? t->lr0 : (rl > t->lr1 ? t->lr1 : rl);
  tl = tl < t->lt0 ? t->lt0 : (tl > t->lt1 ? t->lt1 : tl);
  ye = ye < t->ye0 ? t->ye0 : (ye > t->ye1 ? t->ye1 : ye);
  int jr = (int)((rl - t->lr0) / dr), jt = (int)((tl - t->lt0) / dtt), jy = (int)((ye - t->ye0) / dy);
  jr = jr > t->nr - 1 ? t->nr - 1 : jr;
  jt = jt > t->nt - 1 ? t->nt - 1 : jt;
  jy = jy > t->ny - 1 ? t->ny - 1 : jy;
  int jrp = jr + 1, jtp = jt + 1, jyp = jy + 1;
  double drho = (rl - (t->lr0 + jr * dr)) / dr;
  double dtemp = (tl - (t->lt0 + jt * dtt)) / dtt;
  double dye = (ye - (t->ye0 + jy * dy)) / dy;
  if (jy == t->ny - 1) { jyp = jy; dye = 0.0; }
  if (jr == t->nr - 1) { jrp = jr; drho = 0.0; }
  if (jt == t->nt - 1) { jtp = jt; dtemp = 0.0; }
  double c[8];
  c[0] = (1.0 - drho) * (1.0 - dtemp) * (1.0 - dye);
  c[1] = (1.0 - drho) * (1.0 - dtemp) * dye;
  c[2] = drho * (1.0 - dtemp) * (1.0 - dye);
  c[3] = drho * (1.0 - dtemp) * dye;
  c[4] = (1.0 - drho) * dtemp * (1.0 - dye);
  c[5] = (1.0 - drho) * dtemp * dye;
  c[6] = drho * dtemp * (1.0 - dye);
  c[7] = drho * dtemp * dye;
  const int cy[8] = {jy, jyp, jy, jyp, jy, jyp, jy, jyp};
  const int cr[8] = {jr, jr, jrp, jrp, jr, jr, jrp, jrp};
  const int ct[8] = {jt, jt, jt, jt, jtp, jtp, jtp, jtp};
  double f[MDS_EOS_NFIELD];
  for (int rec = 0; rec < MDS_EOS_NFIELD; rec++) {
    double v = 0.0;
    for (int k = 0; k < 8; k++)
      v += c[k] * (double)t->data[(((size_t)rec * t->ny + cy[k]) * t->nr + cr[k]) * t->nt + ct[k]];
    f[rec] = v;
  }
  const double es = t->e_scale > 0.0 ? t->e_scale : 1.0, ps = t->p_scale > 0.0 ? t->p_scale : 1.0;
  const double eint = (f[MDS_EOS_E] + t->e_shift) * es;        /* eos_stuff.c:1204 */
  const double P = exp(f[MDS_EOS_LNP]) * ps;                  /* :312, 1209 */
  out[0] = rho * eint;
  out[1] = P;
  out[2] = f[MDS_EOS_GAM1];
  out[3] = sqrt(out[2] * P / rho);
  return MDS_OK;
}
