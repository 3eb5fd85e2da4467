// mds_sweep1d.hip -- RK stages for batches of 1D columns (gfx950).
//
// Each 64-lane wavefront owns one contiguous run of zones of one problem;
// lane l holds zone c0 - H + l:
//   lane loads its zone (boundary ghosts are mapped onto interior zones at
//   load time, bc.c:705-1421 semantics, so no ghost storage exists)
//   -> gamma-law EOS in registers (eos_gamma.c:45-57)
//   -> reconstruction of its own zone, neighbours by DPP lane shifts
//      (once per zone instead of once per face as in step.c:500-547)
//   -> HLLC at its lower face (riemann.c:328-499)
//   -> flux divergence, sources and the RK update (step.c:327-478)
//   -> u -> p; wave-level reductions of the CFL limits, one atomicMin per
//      wave into the problem's device clock (step.c:560, 393-394).
// Two kernel families:
//   sweep1d_*   one RK stage per launch, H = 2, 60 zones per wave
//               (p -> ph, then p, ph -> p; any boundary codes)
//   sweep1d2_*  both RK stages of a step in one launch, H = 4, 56 zones per
//               wave: stage 0 leaves 60 valid lanes, the boundary ghosts of
//               the intermediate state are refilled from them by lane
//               permutes, stage 1 produces the 56 owned zones.  The state is
//               read and written once per step (in -> out; the caller swaps
//               the buffers).  Used when no problem is periodic in x.
// Both families run the same per-lane stage code, so they give identical
// results.
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

MDS_DEV int clamp_q(int c, int n)
{
  int q = c + MDS_NG;
  return q < 0 ? 0 : (q > n + 2 * MDS_NG - 1 ? n + 2 * MDS_NG - 1 : q);
}

// CFL and source limits gathered by one lane over the stages it runs.
struct Lim1D {
  double smax = 0.0;        // max |signal speed| over faces
  double locmax = 0.0;      // max zone bound (|v| + cs/scale) / dx
  double srcnum = 1.0e300;  // Cartesian without sources: min numerator over 1e-16
  double lim_src = 1.0e300; // min dt_source_control
  unsigned nsh = 0;         // shock-tagged owned faces (wave-uniform)
};

// Lane roles of one stage.
struct Roles {
  bool face;    // the lower face is valid (its CFL limit counts)
  bool count;   // the lower face is owned by this wave (shock counter)
  bool out;     // the zone update is valid
};

// Fold the wave's limits into the problem clock (lane 0 of the wave).
MDS_DEV void commit_limits(const DevProb &pb, DevClock *ck, unsigned long long *counters, int lane,
                           const Lim1D &lim)
{
  double wsmax = wave_max(lim.smax), wloc = wave_max(lim.locmax);
  double wsrc = wave_min(lim.srcnum), wlim = wave_min(lim.lim_src);
  if (lane == 0) {
    double l = 1.0e300;
    if (wsmax > 0.0) l = fmin(l, pb.dx[0] / wsmax);
    if (wloc > 0.0) l = fmin(l, 1.0 / wloc);
    if (wsrc < 1.0e300) l = fmin(l, wsrc / (0.0 + 1.0e-16));
    l = fmin(l, wlim);
    if (l < 1.0e300) atomicMin(&ck->mindt, dt_key(l));
    if (lim.nsh) atomicAdd(&counters[0], (unsigned long long)lim.nsh);
  }
}

// Refill the boundary-ghost lanes of the intermediate state from the lanes
// holding their source zones (the reference's ghost fill between the stages,
// step.c:99).  Interior lanes permute onto themselves.  Non-periodic codes
// only: then every source zone lies in this wave.
template <int NV, class R>
MDS_DEV void ghost_permute(int c, int n, int c0, int lo, int hi, R &rho, R &uu, R v[3])
{
  bool flip;
  const int s = bc_src(c, n, lo, hi, flip);
  int src_lane = s - c0 + 2 * MDS_NG;
  src_lane = src_lane < 0 ? 0 : (src_lane > 63 ? 63 : src_lane);
  rho = __shfl(rho, src_lane, 64);
  uu = __shfl(uu, src_lane, 64);
  for (int d = 0; d < NV; d++) v[d] = __shfl(v[d], src_lane, 64);
  if (flip) v[0] = -v[0];
}

template <int NV, class R>
MDS_DEV void load_prims(const StateArrays<R> &in, long long off, int s, bool flip, R &rho, R &uu, R v[3])
{
  rho = in.v[0][off + s];
  uu = in.v[1][off + s];
  for (int d = 0; d < 3; d++) v[d] = R(0);
  for (int d = 0; d < NV; d++) v[d] = in.v[2 + d][off + s];
  if (flip) v[0] = -v[0];
}

template <int NV, class R>
MDS_DEV void store_prims(StateArrays<R> &out, long long off, int c, R rho, R uu, const R v[3])
{
  out.v[0][off + c] = rho;
  out.v[1][off + c] = uu;
  for (int d = 0; d < NV; d++) out.v[2 + d][off + c] = v[d];
}

// ---------------------------------------------------------------------------
// FP64 (reference operation order)
// ---------------------------------------------------------------------------
struct Geo64 {
  ReconGeo64 rg;
  FaceGeo<double> fg;
  const double *zg;         // zone record (spherical)
  double vol;
  double gcov0[3], gcon[3];
};

template <bool SPH>
MDS_DEV void load_geo64(const DevProb &pb, int c, int n, Geo64 &g)
{
  const int q = clamp_q(c, n);
  const double *mom = pb.mom[0];
  ReconGeo64 &rg = g.rg;
  rg.a0 = mom[3 * q + 0];
  rg.b0 = mom[3 * q + 1];
  rg.g0 = mom[3 * q + 2];
  rg.am = lane_from_below(rg.a0); rg.ap = lane_from_above(rg.a0);
  rg.bm = lane_from_below(rg.b0); rg.bp = lane_from_above(rg.b0);
  rg.gm = lane_from_below(rg.g0); rg.gp = lane_from_above(rg.g0);
  rg.xlo = pb.startx[0] + c * pb.dx[0];
  rg.xhi = pb.startx[0] + (c + 1) * pb.dx[0];
  recon_geo_x64(rg);
  if (SPH) {
    const double *zg = pb.zone + (size_t)q * MDS_ZG_COUNT;
    g.zg = zg;
    g.fg.area = zg[MDS_ZG_AREA0];
    g.fg.scale = zg[MDS_ZG_SCALEF];
    g.fg.gcov[0] = zg[MDS_ZG_GCOVF0];
    g.fg.gcov[1] = zg[MDS_ZG_GCOVF1];
    g.fg.gcov[2] = zg[MDS_ZG_GCOVF2];
    g.vol = zg[MDS_ZG_VOLUME];
    g.gcov0[0] = zg[MDS_ZG_GCOV00]; g.gcov0[1] = zg[MDS_ZG_GCOV01]; g.gcov0[2] = zg[MDS_ZG_GCOV02];
    g.gcon[0] = zg[MDS_ZG_GCON0]; g.gcon[1] = zg[MDS_ZG_GCON1]; g.gcon[2] = zg[MDS_ZG_GCON2];
  } else {
    g.zg = nullptr;
    g.fg.area = pb.area[0];
    g.fg.scale = 1.0;
    g.fg.gcov[0] = g.fg.gcov[1] = g.fg.gcov[2] = 1.0;
    g.vol = pb.volume;
    for (int d = 0; d < 3; d++) { g.gcov0[d] = 1.0; g.gcon[d] = 1.0; }
  }
}

// One RK stage of one lane.  (rho, uu, v): stage input at this lane's zone;
// U0: conserved form of U^n (sim_p) at this zone, (un, vn) its primitives.
// Outputs (ro, uo, vo) are the input where the update is not valid.
template <int NV, bool SPH>
MDS_DEV void stage_x64(const DevProb &pb, const Geo64 &g, int stage, double dt, double rho, double uu,
                       const double v[3], const double U0[5], double un, const double vn[3], Roles r,
                       double &ro, double &uo, double vo[3], Lim1D &lim)
{
  const double gam = pb.gam;
  const double press = (gam - 1.0) * uu;

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
      para_recon_x64(qm, qv[k], qp, g.rg, al[k], ar[k]);
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

  double F[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, sig = 1.0;
  int isshock = hllc_x64<NV, !SPH>(L, Rt, g.fg, 0, F, sig);
  // CFL limits are reduced per wavefront before the single division
  // (min of dx/|s| = dx/max|s|, min of 1/loc = 1/max loc; bit-identical)
  if (r.face) lim.smax = fmax(lim.smax, fabs(sig));
  lim.nsh += __popcll(__ballot(r.count && isshock));

  // flux through the upper face from the lane above
  double Fp[5];
  for (int k = 0; k < 2 + NV; k++) Fp[k] = lane_from_above(F[k]);

  ro = rho; uo = uu;
  for (int d = 0; d < 3; d++) vo[d] = v[d];
  if (!r.out) return;
  const double *zg = g.zg;
  // sources from the stage input state (rad_fluid.c:195-214, gravity.c:46-64)
  double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (SPH) {
    double vc[3], T[3][3];
    for (int d = 0; d < 3; d++) vc[d] = g.gcov0[d] * v[d];
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
    src[1] += rho * ((v[0] * pb.g0[0]) * g.gcov0[0]);
    src[2] += (rho * pb.g0[0]) * g.gcov0[0];
  }
  double U[5], Uh[5];
  for (int k = 0; k < 5; k++) U[k] = U0[k];
  p_to_u_x64<NV, !SPH>(rho, uu, v, g.gcov0, Uh);
  if (SPH || pb.g0[0] != 0.0) {
    double dsrc = dt_source_control_x64<NV>(un, vn, U, src);
    lim.lim_src = (lim.lim_src < dsrc) ? lim.lim_src : dsrc;
  } else {
    // no sources: the three quotients share the denominator 1e-16 (step.c:212-219)
    double m0 = 0.02 * U[0], m1 = 0.02 * U[1], m2 = 2.0 * un;
    double m = (m0 < m1) ? m0 : m1;
    m = (m < m2) ? m : m2;
    lim.srcnum = (lim.srcnum < m) ? lim.srcnum : m;
  }
  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  for (int k = 0; k < 2 + NV; k++) {
    double divF = (0.0 + Fp[k]) - F[k];
    double ddt = (divF / g.vol) - src[k];
    U[k] = ((pf * U[k]) + (phf * Uh[k])) - ((ff * dt) * ddt);
  }
  u_to_p_x64<NV, !SPH>(U, g.gcon, ro, uo, vo);
  if (stage == 1) {
    // estimate_dt zone bound (step.c:177-187): new velocity, sound speed of
    // the eos array as last filled (from this stage's input state)
    double cs = sqrt((gam * press) / rho);
    double sc0 = SPH ? zg[MDS_ZG_SCALE0] : 1.0;
    double loc = (fabs(vo[0]) + (SPH ? cs / sc0 : cs)) / pb.dx[0];
    lim.locmax = fmax(lim.locmax, loc);
  }
}

}  // namespace

// FP64 single-stage kernel: stage 0 reads p, writes ph; stage 1 reads p, ph, writes p.
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

  const StateArrays<double> &in = (stage == 0) ? sp : sph;
  bool flip;
  const int s = bc_src(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
  double rho, uu, v[3];
  load_prims<NV>(in, off, s, flip, rho, uu, v);
  Geo64 g;
  load_geo64<SPH>(pb, c, n, g);

  Roles r;
  r.face = (lane >= 2) && (lane <= 62) && (c >= 0) && (c <= n);
  r.count = r.face && (lane <= 61 || c == n);
  r.out = (lane >= 2) && (lane <= 61) && (c >= 0) && (c < n);
  // U^n of this zone: the stage input itself in stage 0, sim_p in stage 1
  double rn = rho, un = uu, vn[3] = {v[0], v[1], v[2]};
  if (stage == 1 && r.out) load_prims<NV>(sp, off, c, false, rn, un, vn);
  double U0[5];
  p_to_u_x64<NV, !SPH>(rn, un, vn, g.gcov0, U0);

  Lim1D lim;
  double ro, uo, vo[3];
  stage_x64<NV, SPH>(pb, g, stage, dt, rho, uu, v, U0, un, vn, r, ro, uo, vo, lim);
  if (r.out) {
    StateArrays<double> &out = (stage == 0) ? sph : sp;
    store_prims<NV>(out, off, c, ro, uo, vo);
  }
  commit_limits(pb, ck, a.counters, lane, lim);
}

// FP64 two-stage kernel: reads in (U^n), writes out (U^{n+1}).
template <int NV, bool SPH>
__global__ __launch_bounds__(256) void sweep1d2_x64(Sweep1DArgs a, StateArrays<double> in,
                                                    StateArrays<double> out)
{
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (wv >= a.nseg) return;
  const int2 sg = a.seg[wv];
  const DevProb &pb = a.prob[sg.x];
  DevClock *ck = &a.clock[sg.x];
  const int n = pb.n[0];
  const int c0 = sg.y;
  const int c = c0 - 2 * MDS_NG + lane;
  const long long off = pb.off;
  const bool own = (lane >= 2 * MDS_NG) && (lane < 2 * MDS_NG + MDS_SEG1D2) && (c < n);
  if (!ck->active) {
    // frozen problem: carry its state over to the output buffer
    if (own) {
      double rho, uu, v[3];
      load_prims<NV>(in, off, c, false, rho, uu, v);
      store_prims<NV>(out, off, c, rho, uu, v);
    }
    return;
  }
  const double dt = ck->dt;
  bool flip;
  const int s = bc_src(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
  double rho, uu, v[3];
  load_prims<NV>(in, off, s, flip, rho, uu, v);
  Geo64 g;
  load_geo64<SPH>(pb, c, n, g);
  double U0[5];
  p_to_u_x64<NV, !SPH>(rho, uu, v, g.gcov0, U0);

  const bool inside_f = (c >= 0) && (c <= n), inside_z = (c >= 0) && (c < n);
  const bool owned_face = inside_f && (lane >= 2 * MDS_NG) && (lane < 2 * MDS_NG + MDS_SEG1D2 || c == n);
  Lim1D lim;
  // stage 0: U^n -> U* on lanes 2..61
  const Roles r0{inside_f && lane >= 2 && lane <= 62, owned_face, inside_z && lane >= 2 && lane <= 61};
  double rh, uh, vh[3];
  stage_x64<NV, SPH>(pb, g, 0, dt, rho, uu, v, U0, uu, v, r0, rh, uh, vh, lim);
  ghost_permute<NV>(c, n, c0, pb.bc_lo[0], pb.bc_hi[0], rh, uh, vh);
  // stage 1: U^n, U* -> U^{n+1} on lanes 4..59
  const Roles r1{inside_f && lane >= 4 && lane <= 60, owned_face, inside_z && lane >= 4 && lane <= 59};
  double ro, uo, vo[3];
  stage_x64<NV, SPH>(pb, g, 1, dt, rh, uh, vh, U0, uu, v, r1, ro, uo, vo, lim);
  if (own) store_prims<NV>(out, off, c, ro, uo, vo);
  commit_limits(pb, ck, a.counters, lane, lim);
}

// ---------------------------------------------------------------------------
// FP32-bulk
// ---------------------------------------------------------------------------
namespace {

struct Geo32 {
  ReconGeo32 rg;
  double area;              // lower face (FP64)
  FaceGeo<float> fg;        // FP32 face metric for the Riemann solve (unit area)
  double vol, gcov0, gcon0, cp;
  float c000, isc0;
};

template <bool SPH>
MDS_DEV void load_geo32(const DevProb &pb, int c, int n, Geo32 &g)
{
  const int q = clamp_q(c, n);
  const float *rc;
  g.fg.area = 1.0f;
  g.fg.gcov[1] = g.fg.gcov[2] = 1.0f;
  if (SPH) {
    const ZoneF32 &z = pb.z32[q];
    rc = z.rc;
    g.area = z.area;
    g.fg.scale = z.scalef;
    g.fg.gcov[0] = z.gcovf0;
    g.vol = z.vol;
    g.gcov0 = z.gcov0;
    g.gcon0 = z.gcon0;
    g.cp = z.cp;
    g.c000 = z.c000;
    g.isc0 = z.isc0;
  } else {
    rc = pb.rc[0] + (size_t)q * MDS_RC_COUNT;
    g.area = pb.area[0];
    g.fg.scale = 1.0f;
    g.fg.gcov[0] = 1.0f;
    g.vol = pb.volume;
    g.gcov0 = g.gcon0 = 1.0;
    g.cp = 0.0;
    g.c000 = 0.0f;
    g.isc0 = 1.0f;
  }
  g.rg.clm = rc[MDS_RC_CLM]; g.rg.clp = rc[MDS_RC_CLP]; g.rg.crm = rc[MDS_RC_CRM]; g.rg.crp = rc[MDS_RC_CRP];
  g.rg.mu1 = rc[MDS_RC_MU1]; g.rg.kinv = rc[MDS_RC_KINV]; g.rg.inur = rc[MDS_RC_INUR]; g.rg.imu2 = rc[MDS_RC_IMU2];
}

// U^n in conserved form (FP64) and its primitives as doubles.
template <int NV, bool SPH>
MDS_DEV void conserved_f32(const Geo32 &g, float rho, float uu, const float v[3], double U0[5], double &un,
                           double vn[3])
{
  un = uu;
  for (int d = 0; d < 3; d++) vn[d] = v[d];
  const double gcov[3] = {g.gcov0, 1.0, 1.0};
  p_to_u_x64<NV, !SPH>((double)rho, un, vn, gcov, U0);
}

// One FP32-bulk RK stage of one lane: reconstruction and Riemann fluxes in
// FP32 (per unit area), area-weighted fluxes, sources, the update and u -> p
// in FP64.  (rho, uu, v): stage input; U0: conserved U^n; (un, vn) its prims.
template <int NV, bool SPH>
MDS_DEV void stage_f32(const DevProb &pb, const Geo32 &g, int stage, double dt, float rho, float uu,
                       const float v[3], const double U0[5], double un, const double vn[3], Roles r,
                       float &ro, float &uo, float vo[3], Lim1D &lim)
{
  const float gam = (float)pb.gam;
  const float press = (gam - 1.0f) * uu;
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
      para_recon_f32(qm, qv[k], qp, g.rg, al[k], ar[k]);
    }
  }
  PState<float, NV> L, Rt;
  L.rho = lane_from_below(ar[0]);  Rt.rho = al[0];
  L.uu = lane_from_below(ar[1]);   Rt.uu = al[1];
  for (int d = 0; d < NV; d++) { L.v[d] = lane_from_below(ar[2 + d]); Rt.v[d] = al[2 + d]; }
  for (int d = NV; d < 3; d++) { L.v[d] = 0.0f; Rt.v[d] = 0.0f; }
  L.P = lane_from_below(ar[2 + NV]); Rt.P = al[2 + NV];
  L.G = gam; Rt.G = gam;

  float Ff[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, sig = 1.0f;
  int isshock = 0;
  if (g.area >= 1.0e-16) isshock = hllc_f32<NV, !SPH>(L, Rt, g.fg, 0, Ff, sig);
  else sig = 1.0e-30f;
  if (r.face) lim.smax = fmax(lim.smax, (double)fabsf(sig));
  lim.nsh += __popcll(__ballot(r.count && isshock));
  double F[5], Fp[5];
  for (int k = 0; k < 2 + NV; k++) F[k] = g.area * (double)Ff[k];
  for (int k = 0; k < 2 + NV; k++) Fp[k] = lane_from_above(F[k]);

  ro = rho; uo = uu;
  for (int d = 0; d < 3; d++) vo[d] = v[d];
  if (!r.out) return;
  const double drho = rho, dpress = press;
  double dv[3] = {v[0], v[1], v[2]};
  double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (SPH) {
    // radial column, v^1 only: T[0][0] conn000 + P (conn101 + conn202)
    src[2] = ((drho * dv[0]) * (g.gcov0 * dv[0])) * (double)g.c000 + dpress * g.cp;
  }
  if (pb.g0[0] != 0.0) {
    src[1] += drho * ((dv[0] * pb.g0[0]) * g.gcov0);
    src[2] += (drho * pb.g0[0]) * g.gcov0;
  }
  const double gcov[3] = {g.gcov0, 1.0, 1.0};
  double U[5], Uh[5];
  for (int k = 0; k < 5; k++) U[k] = U0[k];
  p_to_u_x64<NV, !SPH>(drho, (double)uu, dv, gcov, Uh);
  lim.lim_src = fmin(lim.lim_src, dt_source_control_fast<NV>(un, vn, U, src));
  const double ivol = drcp(g.vol);
  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  for (int k = 0; k < 2 + NV; k++) {
    double ddt = (Fp[k] - F[k]) * ivol - src[k];
    U[k] = pf * U[k] + phf * Uh[k] - (ff * dt) * ddt;
  }
  double vout[3] = {0.0, 0.0, 0.0}, vsq = 0.0;
  const double ri = drcp(U[0]);
  for (int d = 0; d < NV; d++) {
    double m = U[2 + d] * ri;           // covariant velocity
    vout[d] = SPH ? g.gcon0 * m : m;
    vsq += m * vout[d];
  }
  const double uout = U[1] - 0.5 * U[0] * vsq;
  ro = (float)U[0];
  uo = (float)uout;
  for (int d = 0; d < NV; d++) vo[d] = (float)vout[d];
  if (stage == 1) {
    float cs = fsqrt(gam * press * frcp(rho));
    double loc = (fabs(vout[0]) + (double)(cs * g.isc0)) * (1.0 / pb.dx[0]);
    lim.locmax = fmax(lim.locmax, loc);
  }
}

}  // namespace

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

  const StateArrays<float> &in = (stage == 0) ? sp : sph;
  bool flip;
  const int s = bc_src(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
  float rho, uu, v[3];
  load_prims<NV>(in, off, s, flip, rho, uu, v);
  Geo32 g;
  load_geo32<SPH>(pb, c, n, g);
  Roles r;
  r.face = (lane >= 2) && (lane <= 62) && (c >= 0) && (c <= n);
  r.count = r.face && (lane <= 61 || c == n);
  r.out = (lane >= 2) && (lane <= 61) && (c >= 0) && (c < n);
  float rn = rho, unf = uu, vnf[3] = {v[0], v[1], v[2]};
  if (stage == 1 && r.out) load_prims<NV>(sp, off, c, false, rn, unf, vnf);
  double U0[5], un, vn[3];
  conserved_f32<NV, SPH>(g, rn, unf, vnf, U0, un, vn);

  Lim1D lim;
  float ro, uo, vo[3];
  stage_f32<NV, SPH>(pb, g, stage, dt, rho, uu, v, U0, un, vn, r, ro, uo, vo, lim);
  if (r.out) {
    StateArrays<float> &out = (stage == 0) ? sph : sp;
    store_prims<NV>(out, off, c, ro, uo, vo);
  }
  commit_limits(pb, ck, a.counters, lane, lim);
}

template <int NV, bool SPH>
__global__ __launch_bounds__(256) void sweep1d2_f32(Sweep1DArgs a, StateArrays<float> in,
                                                    StateArrays<float> out)
{
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (wv >= a.nseg) return;
  const int2 sg = a.seg[wv];
  const DevProb &pb = a.prob[sg.x];
  DevClock *ck = &a.clock[sg.x];
  const int n = pb.n[0];
  const int c0 = sg.y;
  const int c = c0 - 2 * MDS_NG + lane;
  const long long off = pb.off;
  const bool own = (lane >= 2 * MDS_NG) && (lane < 2 * MDS_NG + MDS_SEG1D2) && (c < n);
  if (!ck->active) {
    if (own) {
      float rho, uu, v[3];
      load_prims<NV>(in, off, c, false, rho, uu, v);
      store_prims<NV>(out, off, c, rho, uu, v);
    }
    return;
  }
  const double dt = ck->dt;
  bool flip;
  const int s = bc_src(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
  float rho, uu, v[3];
  load_prims<NV>(in, off, s, flip, rho, uu, v);
  Geo32 g;
  load_geo32<SPH>(pb, c, n, g);
  double U0[5], un, vn[3];
  conserved_f32<NV, SPH>(g, rho, uu, v, U0, un, vn);

  const bool inside_f = (c >= 0) && (c <= n), inside_z = (c >= 0) && (c < n);
  const bool owned_face = inside_f && (lane >= 2 * MDS_NG) && (lane < 2 * MDS_NG + MDS_SEG1D2 || c == n);
  Lim1D lim;
  const Roles r0{inside_f && lane >= 2 && lane <= 62, owned_face, inside_z && lane >= 2 && lane <= 61};
  float rh, uh, vh[3];
  stage_f32<NV, SPH>(pb, g, 0, dt, rho, uu, v, U0, un, vn, r0, rh, uh, vh, lim);
  ghost_permute<NV>(c, n, c0, pb.bc_lo[0], pb.bc_hi[0], rh, uh, vh);
  const Roles r1{inside_f && lane >= 4 && lane <= 60, owned_face, inside_z && lane >= 4 && lane <= 59};
  float ro, uo, vo[3];
  stage_f32<NV, SPH>(pb, g, 1, dt, rh, uh, vh, U0, un, vn, r1, ro, uo, vo, lim);
  if (own) store_prims<NV>(out, off, c, ro, uo, vo);
  commit_limits(pb, ck, a.counters, lane, lim);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// stage 0/1: single-stage kernels on (p, ph); stage -1: the two-stage kernel
// with sp = input buffer and sph_ = output buffer (a->seg must then be the
// MDS_SEG1D2 segment table).
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
    if (stage < 0) {
      if (sph) hipLaunchKernelGGL((sweep1d2_x64<3, true>), grid, block, 0, st, *a, p, h);
      else if (nv == 1) hipLaunchKernelGGL((sweep1d2_x64<1, false>), grid, block, 0, st, *a, p, h);
      else hipLaunchKernelGGL((sweep1d2_x64<3, false>), grid, block, 0, st, *a, p, h);
    } else {
      if (sph) hipLaunchKernelGGL((sweep1d_x64<3, true>), grid, block, 0, st, *a, p, h, stage);
      else if (nv == 1) hipLaunchKernelGGL((sweep1d_x64<1, false>), grid, block, 0, st, *a, p, h, stage);
      else hipLaunchKernelGGL((sweep1d_x64<3, false>), grid, block, 0, st, *a, p, h, stage);
    }
  } else {
    const StateArrays<float> &p = *(const StateArrays<float> *)sp;
    const StateArrays<float> &h = *(const StateArrays<float> *)sph_;
    if (stage < 0) {
      if (sph) hipLaunchKernelGGL((sweep1d2_f32<1, true>), grid, block, 0, st, *a, p, h);
      else if (nv == 1) hipLaunchKernelGGL((sweep1d2_f32<1, false>), grid, block, 0, st, *a, p, h);
      else hipLaunchKernelGGL((sweep1d2_f32<3, false>), grid, block, 0, st, *a, p, h);
    } else {
      if (sph) hipLaunchKernelGGL((sweep1d_f32<1, true>), grid, block, 0, st, *a, p, h, stage);
      else if (nv == 1) hipLaunchKernelGGL((sweep1d_f32<1, false>), grid, block, 0, st, *a, p, h, stage);
      else hipLaunchKernelGGL((sweep1d_f32<3, false>), grid, block, 0, st, *a, p, h, stage);
    }
  }
  return hipGetLastError();
}
