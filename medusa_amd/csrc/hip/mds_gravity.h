// mds_gravity.h -- GR effective monopole gravity (gravity.c:142-375 with
// GR_MONOPOLE, no radiation: the radiation pressure, energy and v.F shell
// terms are zero), evaluated by one workgroup per problem.
//
// Input: the volume-weighted shell sums of rho, P, u and v^1 in the
// problem's GR block (DevProb::gr, MDS_GR_* arrays of n1 + 1 doubles), and
// the host-made shell geometry (face radius, (4 pi) r^3, dr/dx, shell
// volume).  Output: the zone-centred field gz[0..n1), the TOV factor Gamma
// and mass at the faces (Gamma is kept: the next call starts from it), and
// the lapses exp(phi / c^2) at the faces and zone centres.
//
// Steps, in the reference's order and operation order:
//   averages  sum * (1 / V_shell)                           (gravity.c:277-287)
//   Gamma     first call: the Newtonian estimate             (gravity.c:292-301)
//             passes over the faces ii = 1..n1 until the outer TOV mass moves
//             by less than 1e-10 (at most 21 passes)       (gravity.c:303-336)
//   g         face gravity with the pressure and inertia terms (gravity.c:339-346)
//   lapse     potential integrated inwards from -G M / r_out (gravity.c:349-352)
//   zone      g and the potential linear in x across each zone at the volume
//             centroid beta/Gamma; lapses exponentiated     (gravity.c:355-375)
// The pass over the faces is a recurrence (Gamma at face ii needs the mass
// below it), so it runs on thread 0, which reads each chunk of MDS_GR_CH
// faces from LDS after the whole workgroup has staged it; the per-face
// terms that do not depend on the recurrence are formed while staging.
#pragma once
#include "mds_internal.h"

#define MDS_GR_CH 256
#define MDS_GR_MAXPASS 20

namespace gr {

constexpr double GNEWT = 6.67384e-8;       // constants.h
constexpr double CLIGHT = 2.99792458e10;   // constants.h:10

__device__ __forceinline__ double *arr(const DevProb &pb, int k) { return pb.gr + (size_t)k * (pb.n[0] + 1); }

// zone moments beta and Gamma of radial zone i (ghost zones included)
__device__ __forceinline__ double mom_be(const DevProb &pb, int i) { return pb.mom[0][3 * (i + MDS_NG) + 1]; }
__device__ __forceinline__ double mom_ga(const DevProb &pb, int i) { return pb.mom[0][3 * (i + MDS_NG) + 2]; }

}  // namespace gr

// The field of one problem from its shell sums (see above); NT threads.
template <int NT>
__device__ void gr_monopole_block(const DevProb &pb, double *gz)
{
  using namespace gr;
  const int n1 = pb.n[0];
  const int tid = threadIdx.x;
  const double c2 = CLIGHT * CLIGHT;
  const double dx = pb.dx[0];
  double *redge = arr(pb, MDS_GR_REDGE), *r34 = arr(pb, MDS_GR_R34), *drdx = arr(pb, MDS_GR_DRDX);
  double *vsh = arr(pb, MDS_GR_VSH), *gtov = arr(pb, MDS_GR_GTOV), *mr = arr(pb, MDS_GR_MR);
  double *le = arr(pb, MDS_GR_LAPSE_E), *lc = arr(pb, MDS_GR_LAPSE), *gf = arr(pb, MDS_GR_GFACE);
  double *rho = arr(pb, MDS_GR_RHO), *pg = arr(pb, MDS_GR_P), *uu = arr(pb, MDS_GR_U), *vv = arr(pb, MDS_GR_V);

  // volume averages over the exact shell volume
  for (int i = tid; i < n1; i += NT) {
    const double iv = 1.0 / vsh[i];
    rho[i] *= iv;
    pg[i] *= iv;
    uu[i] *= iv;
    vv[i] *= iv;
  }
  __syncthreads();

  __shared__ double s_go[MDS_GR_CH], s_vq[MDS_GR_CH], s_w[MDS_GR_CH], s_be[MDS_GR_CH], s_ga[MDS_GR_CH];
  __shared__ double s_rc[MDS_GR_CH], s_vs[MDS_GR_CH], s_ra[MDS_GR_CH];
  __shared__ double s_mstar;
  __shared__ int s_done;

  // thread 0 carries the recurrence across the chunks and passes
  double mstar = 0.0, mlast = 0.0;
  int iter = 0;
  if (tid == 0) {
    mr[0] = 0.0;
    if (!pb.grst[0]) {
      // first call: Gamma from the Newtonian mass (gravity.c:292-301)
      gtov[0] = 1.0;
      double ms = 0.0;
      for (int ii = 1; ii <= n1; ii++) {
        ms += vsh[ii - 1] * rho[ii - 1];
        gtov[ii] = sqrt(1.0 - ((2.0 * GNEWT) * ms) / (c2 * redge[ii]));
      }
      pb.grst[0] = 1;
    }
  }
  __syncthreads();

  while (true) {
    double gp = 0.0, mp = 0.0;            // thread 0: Gamma and mass at face ii - 1
    if (tid == 0) {
      gp = gtov[0];
      mp = mr[0];
      mstar = 0.0;
    }
    for (int base = 1; base <= n1; base += MDS_GR_CH) {
      const int cnt = min(MDS_GR_CH, n1 - base + 1);
      for (int t = tid; t < cnt; t += NT) {
        const int ii = base + t;
        s_go[t] = gtov[ii];
        const double ve = (ii < n1) ? 0.5 * (vv[ii - 1] + vv[ii]) : vv[ii - 1];
        const double vc = ve / CLIGHT;
        s_vq[t] = 1.0 + vc * vc;
        s_w[t] = (rho[ii - 1] + uu[ii - 1] / c2) + 0.0;
        s_be[t] = mom_be(pb, ii);
        s_ga[t] = mom_ga(pb, ii);
        s_rc[t] = redge[ii] * c2;
        s_vs[t] = vsh[ii - 1];
        s_ra[t] = rho[ii - 1];
      }
      __syncthreads();
      if (tid == 0) {
        for (int t = 0; t < cnt; t++) {
          const int ii = base + t;
          const double xl = pb.startx[0] + ii * dx;
          double A = (s_go[t] - gp) / dx;
          double B = gp - A * xl;
          double gavg = (A * s_be[t]) / s_ga[t] + B;
          const double V = s_vs[t], w = s_w[t];
          const double g = sqrt(s_vq[t] - ((2.0 * GNEWT) * (mp + V * (w * gavg))) / s_rc[t]);
          A = (g - gp) / dx;
          B = gp - A * xl;
          gavg = (A * s_be[t]) / s_ga[t] + B;
          const double m = mp + V * (w * gavg);
          gtov[ii] = g;
          mr[ii] = m;
          mstar += V * s_ra[t];
          gp = g;
          mp = m;
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      int done = 0;
      if (iter > 0 && fabs(mlast - mp) / mp < 1.0e-10) done = 1;
      if (!done) {
        mlast = mp;
        iter++;
        if (iter > MDS_GR_MAXPASS) {
          done = 1;
          pb.grst[2] += 1;            // did not converge: the reference exits here
        }
      }
      s_done = done;
      s_mstar = mstar;
      pb.grst[1] = iter;
    }
    __syncthreads();
    if (s_done) break;
  }

  // face gravity with the pressure and inertia terms (gravity.c:339-346)
  for (int ii = tid; ii <= n1; ii += NT) {
    double g = 0.0;
    if (ii > 0) {
      const int a = ii - 1, b = ii < n1 ? ii : ii - 1;
      const double pe = ii < n1 ? 0.5 * (pg[a] + pg[b]) : pg[a];
      const double re = ii < n1 ? 0.5 * (rho[a] + rho[b]) : rho[a];
      const double ue = ii < n1 ? 0.5 * (uu[a] + uu[b]) : uu[a];
      const double rg = redge[ii] * gtov[ii];
      g = ((((-GNEWT * (mr[ii] + (r34[ii] * (pe + 0.0)) / c2)) / (rg * rg)) * (re + (ue + pe) / c2)) / re) * drdx[ii];
    }
    gf[ii] = g;
  }
  __syncthreads();

  // the potential, integrated inwards from -G M_star / r_out (thread 0, the
  // face values staged through LDS a chunk at a time, from the top)
  double lp = 0.0;
  if (tid == 0) {
    lp = (-GNEWT * s_mstar) / redge[n1];
    le[n1] = lp;
  }
  for (int top = n1 - 1; top >= 0; top -= MDS_GR_CH) {
    const int cnt = min(MDS_GR_CH, top + 1);
    // s_go[t] = gf[top - t + 1], s_w[t] = gf[top - t]
    for (int t = tid; t < cnt; t += NT) {
      s_go[t] = gf[top - t + 1];
      s_w[t] = gf[top - t];
    }
    __syncthreads();
    if (tid == 0) {
      for (int t = 0; t < cnt; t++) {
        lp = lp + (0.5 * (s_w[t] + s_go[t])) * dx;
        le[top - t] = lp;
      }
    }
    __syncthreads();
  }

  // zone values: g and the potential at the volume centroid, then the lapses
  for (int i = tid; i < n1; i += NT) {
    const double xl = pb.startx[0] + i * dx;
    const double be = mom_be(pb, i), ga = mom_ga(pb, i);
    double A = (gf[i + 1] - gf[i]) / dx;
    double B = gf[i] - A * xl;
    gz[i] = (A * be) / ga + B;
    A = (le[i + 1] - le[i]) / dx;
    B = le[i] - A * xl;
    lc[i] = exp(((A * be) / ga + B) / c2);
  }
  __syncthreads();
  for (int i = tid; i <= n1; i += NT) le[i] = exp(le[i] / c2);
}
