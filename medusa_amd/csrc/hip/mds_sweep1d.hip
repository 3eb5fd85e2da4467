// mds_sweep1d.hip -- fused RK stage for batches of 1D columns (gfx950).
//
// One 64-lane wavefront owns 64 consecutive zones [c0-2, c0+62) of one
// problem and produces the update of the 60 zones [c0, c0+60):
//   lane loads its zone (boundary ghosts are mapped onto interior zones at
//   load time, bc.c:705-1421 semantics, so no ghost storage exists)
//   -> gamma-law EOS in registers (eos_gamma.c:45-57)
//   -> reconstruction of its own zone, neighbours by cross-lane shuffles
//      (once per zone instead of once per face as in step.c:500-547)
//   -> HLLC at its lower face (riemann.c:328-499)
//   -> flux divergence, sources and the RK update (step.c:327-478)
//   -> u -> p and store; wave-level min of the CFL limits, one atomicMin
//      per wave into the problem's device clock (step.c:560, 393-394).
// Reference loops replaced: prim_to_fluxes1 (step.c:482-567) + push
// (step.c:327-478) + reset_boundaries_gpu for 1D (bc.c:705-1030) +
// update_eos/update_eos_ghost (eos.c:4-32) + get_geom_src_terms
// (rad_fluid.c:195-214) + fixed_gravity (gravity.c:46-64).
#include "mds_internal.h"
#include "mds_physics.h"

namespace {

// Map a stencil zone index onto the interior zone that the reference's ghost
// fill would copy there (bc.c:727-1030 for direction 0).
MDS_DEV int bc_src(int c, int n, int lo, int hi, bool &flip)
{
  flip = false;
  if (c < 0) {
    if (c < -MDS_NG) return 0;
    if (lo == MDS_BC_REFLECT) { flip = true; return -1 - c; }
    if (lo == MDS_BC_PERIODIC) return c + n;
    return 0;
  }
  if (c >= n) {
    if (c >= n + MDS_NG) return n - 1;
    if (hi == MDS_BC_REFLECT) { flip = true; return 2 * n - 1 - c; }
    if (hi == MDS_BC_PERIODIC) return c - n;
    return n - 1;
  }
  return c;
}

template <class R>
MDS_DEV R ldg(const R *p) { return __builtin_nontemporal_load(p); }

}  // namespace

// ---------------------------------------------------------------------------
// FP64 stage kernel (reference operation order)
// ---------------------------------------------------------------------------
template <int NV, bool SPH>
__global__ __launch_bounds__(256) void sweep1d_x64(Sweep1DArgs a, StateArrays<double> sp,
                                                   StateArrays<double> sph, int stage)
{
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (wv >= a.nseg) return;
  const int2 sg = a.seg[wv];
  const DevProb &pb = a.prob[sg.x];
  DevClock *ck = &a.clock[sg.x];
  if (!ck->active) return;
  const double dt = ck->dt;
  const int n = pb.n[0];
  const int c = sg.y - MDS_NG + lane;
  const long long off = pb.off;
  const double gam = pb.gam;

  // stage input state: sim_p for the first stage, sim_ph for the second
  const StateArrays<double> &in = (stage == 0) ? sp : sph;
  bool flip;
  const int s = bc_src(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
  double rho = in.v[0][off + s];
  double uu = in.v[1][off + s];
  double v[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < NV; d++) v[d] = in.v[2 + d][off + s];
  if (flip) v[0] *= -1.0;
  const double press = (gam - 1.0) * uu;

  // reconstruction geometry of this zone
  int q = c + MDS_NG;
  q = q < 0 ? 0 : (q > n + 2 * MDS_NG - 1 ? n + 2 * MDS_NG - 1 : q);
  const double *mom = pb.mom[0];
  ReconGeo64 rg;
  rg.a0 = mom[3 * q + 0];
  rg.b0 = mom[3 * q + 1];
  rg.g0 = mom[3 * q + 2];
  rg.am = lane_from_below(rg.a0); rg.ap = lane_from_above(rg.a0);
  rg.bm = lane_from_below(rg.b0); rg.bp = lane_from_above(rg.b0);
  rg.gm = lane_from_below(rg.g0); rg.gp = lane_from_above(rg.g0);
  rg.xlo = pb.startx[0] + c * pb.dx[0];
  rg.xhi = pb.startx[0] + (c + 1) * pb.dx[0];
  recon_geo_x64(rg);

  // reconstruct rho, u, v^d, P (Gamma is uniform: para_recon returns it flat)
  double al[3 + NV], ar[3 + NV];
  {
    double qv[3 + NV];
    qv[0] = rho; qv[1] = uu;
    for (int d = 0; d < NV; d++) qv[2 + d] = v[d];
    qv[2 + NV] = press;
#pragma unroll
    for (int k = 0; k < 3 + NV; k++) {
      double qm = lane_from_below(qv[k]);
      double qp = lane_from_above(qv[k]);
      para_recon_x64(qm, qv[k], qp, rg, al[k], ar[k]);
    }
  }

  // face below this zone: left state = right edge of the zone below
  PState<double, NV> L, Rt;
  L.rho = lane_from_below(ar[0]);  Rt.rho = al[0];
  L.uu = lane_from_below(ar[1]);   Rt.uu = al[1];
  for (int d = 0; d < NV; d++) { L.v[d] = lane_from_below(ar[2 + d]); Rt.v[d] = al[2 + d]; }
  for (int d = NV; d < 3; d++) { L.v[d] = 0.0; Rt.v[d] = 0.0; }
  L.P = lane_from_below(ar[2 + NV]); Rt.P = al[2 + NV];
  L.G = gam; Rt.G = gam;

  FaceGeo<double> fg;
  const double *zg = nullptr;
  if (SPH) {
    zg = pb.zone + (size_t)q * MDS_ZG_COUNT;
    fg.area = zg[MDS_ZG_AREA0];
    fg.scale = zg[MDS_ZG_SCALEF];
    fg.gcov[0] = zg[MDS_ZG_GCOVF0];
    fg.gcov[1] = zg[MDS_ZG_GCOVF1];
    fg.gcov[2] = zg[MDS_ZG_GCOVF2];
  } else {
    fg.area = pb.area[0];
    fg.scale = 1.0;
    fg.gcov[0] = fg.gcov[1] = fg.gcov[2] = 1.0;
  }
  double F[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, sig = 1.0;
  int isshock = hllc_x64<NV, !SPH>(L, Rt, fg, 0, F, sig);
  const bool face_ok = (lane >= 2) && (lane <= 62) && (c >= 0) && (c <= n);
  // CFL limits, reduced per wavefront before the single division (min of
  // dx/|s| = dx/max|s|, min of 1/loc = 1/max loc; bit-identical)
  double smax = face_ok ? fabs(sig) : 0.0;
  double locmax = 0.0, srcnum = 1.0e300, lim_src = 1.0e300;
  unsigned long long nsh = __popcll(__ballot(face_ok && isshock && lane < 62));

  // flux through the upper face from the lane above
  double Fp[5];
  for (int k = 0; k < 2 + NV; k++) Fp[k] = lane_from_above(F[k]);

  const bool out_ok = (lane >= 2) && (lane <= 61) && (c >= 0) && (c < n);
  if (out_ok) {
    double gcov0[3] = {1.0, 1.0, 1.0}, gcon[3] = {1.0, 1.0, 1.0};
    if (SPH) {
      gcov0[0] = zg[MDS_ZG_GCOV00]; gcov0[1] = zg[MDS_ZG_GCOV01]; gcov0[2] = zg[MDS_ZG_GCOV02];
      gcon[0] = zg[MDS_ZG_GCON0]; gcon[1] = zg[MDS_ZG_GCON1]; gcon[2] = zg[MDS_ZG_GCON2];
    }
    // sources from the stage input state (rad_fluid.c:195-214, gravity.c:46-64)
    double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (SPH) {
      double vc[3], T[3][3];
      for (int d = 0; d < 3; d++) vc[d] = gcov0[d] * v[d];
      for (int l = 0; l < 3; l++) {
        for (int m = 0; m < 3; m++) T[l][m] = (rho * v[m]) * vc[l];
        T[l][l] += press;
      }
      // non-zero conn[l][m][n] in the reference's (l, n) accumulation order
      src[2] += T[0][0] * zg[MDS_ZG_C000];
      src[2] += T[1][1] * zg[MDS_ZG_C101];
      src[2] += T[2][2] * zg[MDS_ZG_C202];
      if (NV > 1) {
        src[3] += T[0][1] * zg[MDS_ZG_C011];
        src[3] += T[1][0] * zg[MDS_ZG_C110];
        src[3] += T[1][1] * zg[MDS_ZG_C111];
        src[3] += T[2][2] * zg[MDS_ZG_C212];
      }
      if (NV > 2) {
        src[4] += T[0][2] * zg[MDS_ZG_C022];
        src[4] += T[1][2] * zg[MDS_ZG_C122];
        src[4] += T[2][0] * zg[MDS_ZG_C220];
        src[4] += T[2][1] * zg[MDS_ZG_C221];
      }
    }
    if (pb.g0[0] != 0.0) {
      src[1] += rho * ((v[0] * pb.g0[0]) * gcov0[0]);
      src[2] += (rho * pb.g0[0]) * gcov0[0];
    }
    // U^n (sim_p) of this zone and its conserved form; the stage input's conserved form
    double rn = rho, un = uu, vn[3] = {v[0], v[1], v[2]};
    if (stage == 1) {
      rn = sp.v[0][off + c];
      un = sp.v[1][off + c];
      for (int d = 0; d < NV; d++) vn[d] = sp.v[2 + d][off + c];
    }
    double U[5], Uh[5];
    p_to_u_x64<NV, !SPH>(rn, un, vn, gcov0, U);
    p_to_u_x64<NV, !SPH>(rho, uu, v, gcov0, Uh);
    if (SPH || pb.g0[0] != 0.0) {
      double dsrc = dt_source_control_x64<NV>(un, vn, U, src);
      lim_src = (lim_src < dsrc) ? lim_src : dsrc;
    } else {
      // no sources: the three quotients share the denominator 1e-16 (step.c:212-219)
      double m0 = 0.02 * U[0], m1 = 0.02 * U[1], m2 = 2.0 * un;
      double m = (m0 < m1) ? m0 : m1;
      m = (m < m2) ? m : m2;
      srcnum = m;
    }

    const double vol = SPH ? zg[MDS_ZG_VOLUME] : pb.volume;
    const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
    for (int k = 0; k < 2 + NV; k++) {
      double divF = (0.0 + Fp[k]) - F[k];
      double ddt = (divF / vol) - src[k];
      U[k] = ((pf * U[k]) + (phf * Uh[k])) - ((ff * dt) * ddt);
    }
    double ro, uo, vo[3];
    u_to_p_x64<NV, !SPH>(U, gcon, ro, uo, vo);
    StateArrays<double> &out = (stage == 0) ? sph : sp;
    out.v[0][off + c] = ro;
    out.v[1][off + c] = uo;
    for (int d = 0; d < NV; d++) out.v[2 + d][off + c] = vo[d];

    if (stage == 1) {
      // estimate_dt zone bound (step.c:177-187): new velocity, sound speed of
      // the eos array as last filled (from this stage's input state)
      double cs = sqrt((gam * press) / rho);
      double sc0 = SPH ? zg[MDS_ZG_SCALE0] : 1.0;
      locmax = (fabs(vo[0]) + (SPH ? cs / sc0 : cs)) / pb.dx[0];
    }
  }
  double wsmax = wave_max(smax), wloc = wave_max(locmax), wsrc = wave_min(srcnum), wlim = wave_min(lim_src);
  if (lane == 0) {
    double lim = 1.0e300;
    if (wsmax > 0.0) lim = fmin(lim, pb.dx[0] / wsmax);
    if (wloc > 0.0) lim = fmin(lim, 1.0 / wloc);
    if (wsrc < 1.0e300) lim = fmin(lim, wsrc / (0.0 + 1.0e-16));
    lim = fmin(lim, wlim);
    if (lim < 1.0e300) atomicMin(&ck->mindt, dt_key(lim));
    if (nsh) atomicAdd(&a.counters[0], nsh);
  }
}

// ---------------------------------------------------------------------------
// FP32-bulk stage kernel
// ---------------------------------------------------------------------------
template <int NV, bool SPH>
__global__ __launch_bounds__(256) void sweep1d_f32(Sweep1DArgs a, StateArrays<float> sp,
                                                   StateArrays<float> sph, int stage)
{
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (wv >= a.nseg) return;
  const int2 sg = a.seg[wv];
  const DevProb &pb = a.prob[sg.x];
  DevClock *ck = &a.clock[sg.x];
  if (!ck->active) return;
  const double dt = ck->dt;
  const int n = pb.n[0];
  const int c = sg.y - MDS_NG + lane;
  const long long off = pb.off;
  const float gam = (float)pb.gam;

  const StateArrays<float> &in = (stage == 0) ? sp : sph;
  bool flip;
  const int s = bc_src(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
  float rho = in.v[0][off + s];
  float uu = in.v[1][off + s];
  float v[3] = {0.0f, 0.0f, 0.0f};
  for (int d = 0; d < NV; d++) v[d] = in.v[2 + d][off + s];
  if (flip) v[0] = -v[0];
  const float press = (gam - 1.0f) * uu;

  int q = c + MDS_NG;
  q = q < 0 ? 0 : (q > n + 2 * MDS_NG - 1 ? n + 2 * MDS_NG - 1 : q);
  const float *rc = pb.rc[0] + (size_t)q * MDS_RC_COUNT;
  ReconGeo32 rg;
  rg.clm = rc[MDS_RC_CLM]; rg.clp = rc[MDS_RC_CLP]; rg.crm = rc[MDS_RC_CRM]; rg.crp = rc[MDS_RC_CRP];
  rg.mu1 = rc[MDS_RC_MU1]; rg.kinv = rc[MDS_RC_KINV]; rg.inur = rc[MDS_RC_INUR]; rg.imu2 = rc[MDS_RC_IMU2];

  float al[3 + NV], ar[3 + NV];
  {
    float qv[3 + NV];
    qv[0] = rho; qv[1] = uu;
    for (int d = 0; d < NV; d++) qv[2 + d] = v[d];
    qv[2 + NV] = press;
#pragma unroll
    for (int k = 0; k < 3 + NV; k++) {
      float qm = lane_from_below(qv[k]);
      float qp = lane_from_above(qv[k]);
      para_recon_f32(qm, qv[k], qp, rg, al[k], ar[k]);
    }
  }
  PState<float, NV> L, Rt;
  L.rho = lane_from_below(ar[0]);  Rt.rho = al[0];
  L.uu = lane_from_below(ar[1]);   Rt.uu = al[1];
  for (int d = 0; d < NV; d++) { L.v[d] = lane_from_below(ar[2 + d]); Rt.v[d] = al[2 + d]; }
  for (int d = NV; d < 3; d++) { L.v[d] = 0.0f; Rt.v[d] = 0.0f; }
  L.P = lane_from_below(ar[2 + NV]); Rt.P = al[2 + NV];
  L.G = gam; Rt.G = gam;

  FaceGeo<float> fg;
  const double *zg = nullptr;
  double area;
  if (SPH) {
    zg = pb.zone + (size_t)q * MDS_ZG_COUNT;
    area = zg[MDS_ZG_AREA0];
    fg.area = (float)area;
    fg.scale = (float)zg[MDS_ZG_SCALEF];
    fg.gcov[0] = (float)zg[MDS_ZG_GCOVF0];
    fg.gcov[1] = (float)zg[MDS_ZG_GCOVF1];
    fg.gcov[2] = (float)zg[MDS_ZG_GCOVF2];
  } else {
    area = pb.area[0];
    fg.area = 1.0f; fg.scale = 1.0f; fg.gcov[0] = fg.gcov[1] = fg.gcov[2] = 1.0f;
  }
  float Ff[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, sig = 1.0f;
  int isshock = 0;
  if (area >= 1.0e-16) isshock = hllc_f32<NV, !SPH>(L, Rt, fg, 0, Ff, sig);
  else sig = 1.0e-30f;
  const bool face_ok = (lane >= 2) && (lane <= 62) && (c >= 0) && (c <= n);
  double smax = face_ok ? (double)fabsf(sig) : 0.0;
  double locmax = 0.0, lim_src = 1.0e300;
  unsigned long long nsh = __popcll(__ballot(face_ok && isshock && lane < 62));
  double F[5];
  for (int k = 0; k < 2 + NV; k++) F[k] = area * (double)Ff[k];
  double Fp[5];
  for (int k = 0; k < 2 + NV; k++) Fp[k] = lane_from_above(F[k]);

  const bool out_ok = (lane >= 2) && (lane <= 61) && (c >= 0) && (c < n);
  if (out_ok) {
    double gcov0[3] = {1.0, 1.0, 1.0}, gcon[3] = {1.0, 1.0, 1.0};
    if (SPH) {
      gcov0[0] = zg[MDS_ZG_GCOV00]; gcov0[1] = zg[MDS_ZG_GCOV01]; gcov0[2] = zg[MDS_ZG_GCOV02];
      gcon[0] = zg[MDS_ZG_GCON0]; gcon[1] = zg[MDS_ZG_GCON1]; gcon[2] = zg[MDS_ZG_GCON2];
    }
    const double drho = rho, duu = uu, dpress = press;
    double dv[3] = {v[0], v[1], v[2]};
    double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (SPH) {
      double vc[3], T[3][3];
      for (int d = 0; d < 3; d++) vc[d] = gcov0[d] * dv[d];
      for (int l = 0; l < 3; l++) {
        for (int m = 0; m < 3; m++) T[l][m] = (drho * dv[m]) * vc[l];
        T[l][l] += dpress;
      }
      src[2] += T[0][0] * zg[MDS_ZG_C000];
      src[2] += T[1][1] * zg[MDS_ZG_C101];
      src[2] += T[2][2] * zg[MDS_ZG_C202];
      if (NV > 1) {
        src[3] += T[0][1] * zg[MDS_ZG_C011];
        src[3] += T[1][0] * zg[MDS_ZG_C110];
        src[3] += T[1][1] * zg[MDS_ZG_C111];
        src[3] += T[2][2] * zg[MDS_ZG_C212];
      }
      if (NV > 2) {
        src[4] += T[0][2] * zg[MDS_ZG_C022];
        src[4] += T[1][2] * zg[MDS_ZG_C122];
        src[4] += T[2][0] * zg[MDS_ZG_C220];
        src[4] += T[2][1] * zg[MDS_ZG_C221];
      }
    }
    if (pb.g0[0] != 0.0) {
      src[1] += drho * ((dv[0] * pb.g0[0]) * gcov0[0]);
      src[2] += (drho * pb.g0[0]) * gcov0[0];
    }
    double rn = drho, un = duu, vn[3] = {dv[0], dv[1], dv[2]};
    if (stage == 1) {
      rn = sp.v[0][off + c];
      un = sp.v[1][off + c];
      for (int d = 0; d < NV; d++) vn[d] = sp.v[2 + d][off + c];
    }
    double U[5], Uh[5];
    p_to_u_x64<NV, !SPH>(rn, un, vn, gcov0, U);
    p_to_u_x64<NV, !SPH>(drho, duu, dv, gcov0, Uh);
    lim_src = dt_source_control_fast<NV>(un, vn, U, src);
    const double vol = SPH ? zg[MDS_ZG_VOLUME] : pb.volume;
    const double ivol = drcp(vol);
    const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
    for (int k = 0; k < 2 + NV; k++) {
      double ddt = (Fp[k] - F[k]) * ivol - src[k];
      U[k] = pf * U[k] + phf * Uh[k] - (ff * dt) * ddt;
    }
    double ro = U[0], uo, vo[3] = {0.0, 0.0, 0.0};
    {
      double ri = drcp(U[0]), vsq = 0.0;
      for (int d = 0; d < NV; d++) {
        double m = U[2 + d] * ri;           // covariant velocity
        vo[d] = SPH ? gcon[d] * m : m;
        vsq += m * vo[d];
      }
      uo = U[1] - 0.5 * ro * vsq;
    }
    StateArrays<float> &out = (stage == 0) ? sph : sp;
    out.v[0][off + c] = (float)ro;
    out.v[1][off + c] = (float)uo;
    for (int d = 0; d < NV; d++) out.v[2 + d][off + c] = (float)vo[d];
    if (stage == 1) {
      float cs = fsqrt((float)pb.gam * press * frcp(rho));
      double sc0 = SPH ? zg[MDS_ZG_SCALE0] : 1.0;
      locmax = (fabs(vo[0]) + (double)cs * (SPH ? drcp(sc0) : 1.0)) * (1.0 / pb.dx[0]);
    }
  }
  double wsmax = wave_max(smax), wloc = wave_max(locmax), wlim = wave_min(lim_src);
  if (lane == 0) {
    double lim = 1.0e300;
    if (wsmax > 0.0) lim = fmin(lim, pb.dx[0] / wsmax);
    if (wloc > 0.0) lim = fmin(lim, 1.0 / wloc);
    lim = fmin(lim, wlim);
    if (lim < 1.0e300) atomicMin(&ck->mindt, dt_key(lim));
    if (nsh) atomicAdd(&a.counters[0], nsh);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define MDS_1D_INSTANCES(NV, SPH)                                                           \
  template __global__ void sweep1d_x64<NV, SPH>(Sweep1DArgs, StateArrays<double>,           \
                                                StateArrays<double>, int);                  \
  template __global__ void sweep1d_f32<NV, SPH>(Sweep1DArgs, StateArrays<float>,            \
                                                StateArrays<float>, int);

extern "C" hipError_t mds_launch_sweep1d(int fp32, int nv, int sph, const Sweep1DArgs *a,
                                         const void *sp, const void *sph_, int stage,
                                         hipStream_t st)
{
  const int waves_per_block = 4;
  dim3 block(64 * waves_per_block);
  dim3 grid((a->nseg + waves_per_block - 1) / waves_per_block);
  if (!fp32) {
    const StateArrays<double> &p = *(const StateArrays<double> *)sp;
    const StateArrays<double> &h = *(const StateArrays<double> *)sph_;
    if (sph) {
      hipLaunchKernelGGL((sweep1d_x64<3, true>), grid, block, 0, st, *a, p, h, stage);
    } else if (nv == 1) {
      hipLaunchKernelGGL((sweep1d_x64<1, false>), grid, block, 0, st, *a, p, h, stage);
    } else {
      hipLaunchKernelGGL((sweep1d_x64<3, false>), grid, block, 0, st, *a, p, h, stage);
    }
  } else {
    const StateArrays<float> &p = *(const StateArrays<float> *)sp;
    const StateArrays<float> &h = *(const StateArrays<float> *)sph_;
    if (sph && nv == 1) {
      hipLaunchKernelGGL((sweep1d_f32<1, true>), grid, block, 0, st, *a, p, h, stage);
    } else if (sph) {
      hipLaunchKernelGGL((sweep1d_f32<3, true>), grid, block, 0, st, *a, p, h, stage);
    } else if (nv == 1) {
      hipLaunchKernelGGL((sweep1d_f32<1, false>), grid, block, 0, st, *a, p, h, stage);
    } else {
      hipLaunchKernelGGL((sweep1d_f32<3, false>), grid, block, 0, st, *a, p, h, stage);
    }
  }
  return hipGetLastError();
}
