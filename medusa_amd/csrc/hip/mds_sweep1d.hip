// mds_sweep1d.hip -- RK stages for batches of 1D columns (gfx950).
//
// Each 64-lane wavefront owns one contiguous run of zones of one problem;
// lane l holds zone c0 - H + l:
//   lane loads its zone (boundary ghosts are mapped onto interior zones at
//   load time, bc.c:705-1421 semantics, so no ghost storage exists)
//   -> EOS in registers: gamma law (eos_gamma.c:45-57) or the table
//      (eos_stuff.c:507-650, see mds_eos.h)
//   -> reconstruction of its own zone, neighbours by DPP lane shifts
//      (once per zone instead of once per face as in step.c:500-547)
//   -> HLLC at its lower face (riemann.c:328-499), Ye as a passive scalar
//   -> flux divergence, sources (geometric, fixed or monopole gravity) and
//      the RK update (step.c:327-478)
//   -> u -> p; wave-level reductions of the CFL limits, one atomicMin per
//      wave into the problem's device clock (step.c:560, 393-394).
// Two kernel families:
//   sweep1d_*   one RK stage per launch, H = 2, 60 zones per wave
//               (p -> ph, then p, ph -> p; any boundary codes; monopole
//               gravity, whose field comes from a separate per-stage kernel)
//   sweep1d2_*  both RK stages of a step in one launch, H = 4, 56 zones per
//               wave: stage 0 leaves 60 valid lanes, the boundary ghosts of
//               the intermediate state are refilled from them by lane
//               permutes, stage 1 produces the 56 owned zones.  The state is
//               read and written once per step (in -> out; the caller swaps
//               the buffers).  Used when no problem is periodic in x and
//               there is no monopole gravity.
// Both families run the same per-lane stage code, so they give identical
// results.
// Reference loops replaced: prim_to_fluxes1 (step.c:482-567) + push
// (step.c:327-478) + reset_boundaries_gpu for 1D (bc.c:705-1030) +
// update_eos/update_eos_ghost (eos.c:4-32) + get_geom_src_terms
// (rad_fluid.c:195-214) + fixed_gravity (gravity.c:46-64) + the source part
// of monopole_gravity_finish (gravity.c:432-498).
#include "mds_internal.h"
#include "mds_physics.h"
#include "mds_eos.h"
#include "mds_gravity.h"

namespace {

// Map a stencil zone index onto the interior zone that the reference's ghost
// fill would copy there (bc.c:727-1030 for direction 0).  flip gets the
// velocity components whose sign the fill changes (bit d: v^(d+1)): REFLECT
// v^1, the inner radius through the origin (SPHERICAL_ORIGIN, bc.c:774-801)
// v^1 and v^3.  In the general instances (GEN) a PROB or EXTRAP side has its
// own ghost layers: flip gets bit 8 (and bit 16 above the grid) and the layer
// index is returned.
template <bool GEN = false>
MDS_DEV int bc_src(int c, int n, int lo, int hi, int &flip)
{
  flip = 0;
  if (c < 0) {
    if (c < -MDS_NG) return 0;
    if constexpr (GEN)
      if (lo == MDS_BC_PROB || lo == MDS_BC_EXTRAP || lo == MDS_BC_DISK) { flip = 8; return c + MDS_NG; }
    if (lo == MDS_BC_REFLECT) { flip = 1; return -1 - c; }
    if (lo == MDS_BC_SPHERICAL_ORIGIN) { flip = 5; return -1 - c; }
    if (lo == MDS_BC_PERIODIC) return c + n;
    return 0;
  }
  if (c >= n) {
    if (c >= n + MDS_NG) return n - 1;
    if constexpr (GEN)
      if (hi == MDS_BC_PROB || hi == MDS_BC_EXTRAP) { flip = 8 | 16; return c - n; }
    if (hi == MDS_BC_REFLECT) { flip = 1; return 2 * n - 1 - c; }
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

// primitive record of one zone (rho, u, v^1..v^3, Ye)
template <class R>
struct Prim {
  R rho, uu, v[3], ye;
};

// CFL and source limits gathered by one lane over the stages it runs.
struct Lim1D {
  double smax = 0.0;        // max |signal speed| over faces
  double locmax = 0.0;      // max zone bound (|v| + cs/scale) / dx
  double srcnum = 1.0e300;  // Cartesian without sources: min numerator over 1e-16
  double lim_src = 1.0e300; // min dt_source_control
  unsigned nsh = 0;         // shock-tagged owned faces (wave-uniform)
  unsigned nbad = 0;        // owned zone updates with an unphysical state (wave-uniform)
  unsigned noff = 0;        // table EOS evaluations of owned zones off the table (wave-uniform)
};

// Lane roles of one stage.
struct Roles {
  bool face;    // the lower face is valid (its CFL limit counts)
  bool count;   // the lower face is owned by this wave (shock counter)
  bool out;     // the zone update is valid
  bool own;     // the zone is owned by this wave (health counts)
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
    if (lim.nbad) atomicAdd(&ck->nbad, (unsigned long long)lim.nbad);
    if (lim.noff) atomicAdd(&ck->noff, (unsigned long long)lim.noff);
  }
}

// Refill the boundary-ghost lanes of the intermediate state from the lanes
// holding their source zones (the reference's ghost fill between the stages,
// step.c:99).  Interior lanes permute onto themselves.  Non-periodic codes
// only: then every source zone lies in this wave.
template <int NV, int TAB, class R>
MDS_DEV void ghost_permute(int c, int n, int c0, int lo, int hi, Prim<R> &q)
{
  int flip;
  const int s = bc_src(c, n, lo, hi, flip);
  int src_lane = s - c0 + 2 * MDS_NG;
  src_lane = src_lane < 0 ? 0 : (src_lane > 63 ? 63 : src_lane);
  q.rho = __shfl(q.rho, src_lane, 64);
  q.uu = __shfl(q.uu, src_lane, 64);
  for (int d = 0; d < NV; d++) q.v[d] = __shfl(q.v[d], src_lane, 64);
  if (TAB) q.ye = __shfl(q.ye, src_lane, 64);
  if (flip & 1) q.v[0] = -q.v[0];
  if (NV > 2 && (flip & 4)) q.v[2] = -q.v[2];
}

// the same with Z zones per lane (zone c0 - 2 NG + Z lane + k in slot k):
// every slot of the source lane is fetched, the source zone's slot kept
template <int NV, int TAB, int Z, class R>
MDS_DEV void ghost_permute(const int (&cz)[Z], int n, int c0, int lo, int hi, Prim<R> (&q)[Z])
{
  if constexpr (Z == 1) {
    ghost_permute<NV, TAB>(cz[0], n, c0, lo, hi, q[0]);
  } else {
    int src_lane[Z], slot[Z], flip[Z];
#pragma unroll
    for (int k = 0; k < Z; k++) {
      const int s = bc_src(cz[k], n, lo, hi, flip[k]);
      int ps = s - c0 + 2 * MDS_NG;
      ps = ps < 0 ? 0 : (ps > 64 * Z - 1 ? 64 * Z - 1 : ps);
      src_lane[k] = ps / Z;
      slot[k] = ps % Z;
    }
    Prim<R> out[Z];
#pragma unroll
    for (int k = 0; k < Z; k++) {
#pragma unroll
      for (int m = 0; m < Z; m++) {
        Prim<R> v;
        v.rho = __shfl(q[m].rho, src_lane[k], 64);
        v.uu = __shfl(q[m].uu, src_lane[k], 64);
        for (int d = 0; d < NV; d++) v.v[d] = __shfl(q[m].v[d], src_lane[k], 64);
        v.ye = TAB ? __shfl(q[m].ye, src_lane[k], 64) : q[m].ye;
        if (m == 0 || slot[k] == m) out[k] = v;
      }
      for (int d = NV; d < 3; d++) out[k].v[d] = q[k].v[d];
      if (flip[k] & 1) out[k].v[0] = -out[k].v[0];
      if (NV > 2 && (flip[k] & 4)) out[k].v[2] = -out[k].v[2];
    }
#pragma unroll
    for (int k = 0; k < Z; k++) q[k] = out[k];
  }
}

template <int NV, int TAB, class R>
MDS_DEV Prim<R> load_prims(const StateArrays<R> &in, long long off, int s, int flip)
{
  Prim<R> q;
  q.rho = in.v[0][off + s];
  q.uu = in.v[1][off + s];
  for (int d = 0; d < 3; d++) q.v[d] = R(0);
  for (int d = 0; d < NV; d++) q.v[d] = in.v[2 + d][off + s];
  q.ye = TAB ? in.v[5][off + s] : R(0);
  if (flip & 1) q.v[0] = -q.v[0];
  if (NV > 2 && (flip & 4)) q.v[2] = -q.v[2];
  return q;
}

// the zone bc_src names: a state zone, or a ghost layer of a PROB/EXTRAP side
template <int NV, int TAB, bool GEN, class R>
MDS_DEV Prim<R> load_zone(const DevProb &pb, const StateArrays<R> &in, int s, int flip)
{
  if (GEN && (flip & 8)) {
    const double *e = pb.gs[0][(flip >> 4) & 1] + (size_t)s * MDS_SLAB_NV;   // one transverse position
    Prim<R> q;
    q.rho = (R)e[0];
    q.uu = (R)e[1];
    for (int d = 0; d < 3; d++) q.v[d] = d < NV ? (R)e[2 + d] : R(0);
    q.ye = TAB ? (R)e[5] : R(0);
    return q;
  }
  return load_prims<NV, TAB>(in, pb.off, s, flip);
}

template <int NV, int TAB, class R>
MDS_DEV void store_prims(StateArrays<R> &out, long long off, int c, const Prim<R> &q)
{
  out.v[0][off + c] = q.rho;
  out.v[1][off + c] = q.uu;
  for (int d = 0; d < NV; d++) out.v[2 + d][off + c] = q.v[d];
  if (TAB) out.v[5][off + c] = q.ye;
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
  rg.order = pb.order;
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

// conserved form (fluid.c:4-25), Ye in slot MDS_YI
template <int NV, bool SPH, int TAB>
MDS_DEV void conserved_x64(const Prim<double> &q, const double gcov[3], double U[6])
{
  p_to_u_x64<NV, !SPH>(q.rho, q.uu, q.v, gcov, U);
  U[MDS_YI] = TAB ? U[0] * q.ye : 0.0;
}

// One RK stage of one lane.  q: stage input at this lane's zone; U0: the
// conserved form of U^n (sim_p) at this zone, qn its primitives; gz: monopole
// gravity at the zone centre.  The output is the input where the update is
// not valid.
template <int NV, bool SPH, int TAB, int MONO, int GEN = 0>
MDS_DEV Prim<double> stage_x64(const DevProb &pb, const DevEos &eos, const Geo64 &g, int stage, double dt,
                               const Prim<double> &q, const double U0[6], const Prim<double> &qn, double gz,
                               Roles r, Lim1D &lim)
{
  const double gam = pb.gam;
  double press, gam1, cs = 0.0;
  if (TAB) {
    double lt;
    eos_fill_x64(eos, q.rho, q.uu, q.ye, press, gam1, cs, lt);
    lim.noff += (unsigned)__popcll(__ballot(r.own && eos_off_table(eos, q.rho, q.ye, lt)));
  } else {
    press = (gam - 1.0) * q.uu;
    gam1 = gam;
  }
  // reconstruct rho, u, v^d, (Ye), P, (Gamma_1); a uniform Gamma would
  // reconstruct flat (para_recon returns it unchanged)
  constexpr int NQ = 3 + NV + 2 * TAB;
  double al[NQ], ar[NQ];
  {
    double qv[NQ];
    qv[0] = q.rho; qv[1] = q.uu;
    for (int d = 0; d < NV; d++) qv[2 + d] = q.v[d];
    if (TAB) { qv[2 + NV] = q.ye; qv[3 + NV] = press; qv[4 + NV] = gam1; }
    else qv[2 + NV] = press;
    // hydro_interp_order 2 or 1: general instances only, one uniform branch
    // around the variable loop
    bool low = false;
    if constexpr (GEN != 0) low = g.rg.order != 3;
    if (!low) {
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        double qm = lane_from_below(qv[k]);
        double qp = lane_from_above(qv[k]);
        para_recon_x64(qm, qv[k], qp, g.rg, al[k], ar[k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        double qm = lane_from_below(qv[k]);
        double qp = lane_from_above(qv[k]);
        recon_x64(qm, qv[k], qp, g.rg, al[k], ar[k]);
      }
    }
  }
  // face below this zone: left state = right edge of the zone below
  PState<double, NV> L, Rt;
  L.rho = lane_from_below(ar[0]);  Rt.rho = al[0];
  L.uu = lane_from_below(ar[1]);   Rt.uu = al[1];
  for (int d = 0; d < NV; d++) { L.v[d] = lane_from_below(ar[2 + d]); Rt.v[d] = al[2 + d]; }
  for (int d = NV; d < 3; d++) { L.v[d] = 0.0; Rt.v[d] = 0.0; }
  if (TAB) {
    L.ye = lane_from_below(ar[2 + NV]); Rt.ye = al[2 + NV];
    L.P = lane_from_below(ar[3 + NV]);  Rt.P = al[3 + NV];
    L.G = lane_from_below(ar[4 + NV]);  Rt.G = al[4 + NV];
  } else {
    L.ye = Rt.ye = 0.0;
    L.P = lane_from_below(ar[2 + NV]); Rt.P = al[2 + NV];
    L.G = gam; Rt.G = gam;
  }

  double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, sig = 1.0;
  int isshock = hllc_x64<NV, !SPH, TAB>(L, Rt, g.fg, 0, F, sig);
  // CFL limits are reduced per wavefront before the single division
  // (min of dx/|s| = dx/max|s|, min of 1/loc = 1/max loc; bit-identical)
  if (r.face) lim.smax = fmax(lim.smax, fabs(sig));
  lim.nsh += __popcll(__ballot(r.count && isshock));

  // flux through the upper face from the lane above
  double Fp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < 2 + NV; k++) Fp[k] = lane_from_above(F[k]);
  if (TAB) Fp[MDS_YI] = lane_from_above(F[MDS_YI]);
  const double area_hi = MONO ? lane_from_above(g.fg.area) : 0.0;

  Prim<double> o = q;
  if (!r.out) return o;
  const double *zg = g.zg;
  // sources from the stage input state (rad_fluid.c:195-214, gravity.c:46-64)
  double src[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (SPH) {
    double vc[3], T[3][3];
    for (int d = 0; d < 3; d++) vc[d] = g.gcov0[d] * q.v[d];
    for (int l = 0; l < 3; l++) {
      for (int m = 0; m < 3; m++) T[l][m] = (q.rho * q.v[m]) * vc[l];
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
    src[1] += q.rho * ((q.v[0] * pb.g0[0]) * g.gcov0[0]);
    src[2] += (q.rho * pb.g0[0]) * g.gcov0[0];
  }
  if (MONO) {
    // gravity.c:475-486: rho g, and the mass flux per area, linear between
    // the faces, at the volume centroid times g
    const double xl = g.rg.xlo;
    double vr = Fp[0];
    vr /= (area_hi + 1.0e-16);
    src[2] += q.rho * gz;
    const double vl = F[0] / (g.fg.area + 1.0e-16);
    const double A = (vr - vl) / pb.dx[0];
    const double B = vl - A * xl;
    src[1] += ((A * g.rg.b0) / g.rg.g0 + B) * gz;
  }
  double U[6], Uh[6];
  for (int k = 0; k < 6; k++) U[k] = U0[k];
  conserved_x64<NV, SPH, TAB>(q, g.gcov0, Uh);
  if (SPH || pb.g0[0] != 0.0 || MONO) {
    double dsrc = dt_source_control_x64<NV>(qn.uu, qn.v, U, src);
    lim.lim_src = (lim.lim_src < dsrc) ? lim.lim_src : dsrc;
  } else {
    // no sources: the three quotients share the denominator 1e-16 (step.c:212-219)
    double m0 = 0.02 * U[0], m1 = 0.02 * U[1], m2 = 2.0 * qn.uu;
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
  if (TAB) {
    const int k = MDS_YI;
    double divF = (0.0 + Fp[k]) - F[k];
    double ddt = (divF / g.vol) - src[k];
    U[k] = ((pf * U[k]) + (phf * Uh[k])) - ((ff * dt) * ddt);
  }
  u_to_p_x64<NV, !SPH>(U, g.gcon, o.rho, o.uu, o.v);
  if (TAB) o.ye = U[MDS_YI] / U[0];
  if (stage == 1) {
    // estimate_dt zone bound (step.c:177-187): new velocity, sound speed of
    // the eos array as last filled (from this stage's input state)
    if (!TAB) cs = sqrt((gam * press) / q.rho);
    double sc0 = SPH ? zg[MDS_ZG_SCALE0] : 1.0;
    double loc = (fabs(o.v[0]) + (SPH ? cs / sc0 : cs)) / pb.dx[0];
    lim.locmax = fmax(lim.locmax, loc);
  }
  return o;
}

}  // namespace

// FP64 single-stage kernel: stage 0 reads p, writes ph; stage 1 reads p, ph, writes p.
template <int NV, bool SPH, int TAB, int MONO, int GEN>
__global__ __launch_bounds__(256) void sweep1d_x64(Sweep1DArgs a, StateArrays<double> sp,
                                                   StateArrays<double> sph, int stage)
{
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  // each wavefront walks the segment table (longest problems first) with the
  // grid's stride: one segment per wavefront, or, in a persistent grid, a
  // share of the table per resident wavefront
  for (int w = wv; w < a.nseg; w += nw) {
    const int2 sg = a.seg[w];
    const DevProb &pb = a.prob[sg.x];
    DevClock *ck = &a.clock[sg.x];
    if (!ck->active) continue;
    const double dt = ck->dt;
    const int n = pb.n[0];
    const int c = sg.y - MDS_NG + lane;
    const long long off = pb.off;

    const StateArrays<double> &in = (stage == 0) ? sp : sph;
    int flip;
    const int s = bc_src<GEN != 0>(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
    const Prim<double> q = load_zone<NV, TAB, GEN != 0>(pb, in, s, flip);
    Geo64 g;
    load_geo64<SPH>(pb, c, n, g);

    Roles r;
    r.face = (lane >= 2) && (lane <= 62) && (c >= 0) && (c <= n);
    r.count = r.face && (lane <= 61 || c == n);
    r.out = (lane >= 2) && (lane <= 61) && (c >= 0) && (c < n);
    r.own = r.out;
    // U^n of this zone: the stage input itself in stage 0, sim_p in stage 1
    Prim<double> qn = q;
    if (stage == 1 && r.out) qn = load_prims<NV, TAB>(sp, off, c, 0);
    double U0[6];
    conserved_x64<NV, SPH, TAB>(qn, g.gcov0, U0);
    const double gz = (MONO && r.out) ? a.grav[off + c] : 0.0;

    Lim1D lim;
    const Prim<double> o = stage_x64<NV, SPH, TAB, MONO, GEN>(pb, a.eos, g, stage, dt, q, U0, qn, gz, r, lim);
    lim.nbad += bad_state_lanes(r.out, o.rho, o.uu, !TAB);
    if (r.out) {
      StateArrays<double> &out = (stage == 0) ? sph : sp;
      store_prims<NV, TAB>(out, off, c, o);
    }
    commit_limits(pb, ck, a.counters, lane, lim);
  }
}

// FP64 two-stage kernel: reads in (U^n), writes out (U^{n+1}).
template <int NV, bool SPH, int TAB>
__global__ __launch_bounds__(256) void sweep1d2_x64(Sweep1DArgs a, StateArrays<double> in,
                                                    StateArrays<double> out)
{
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  // each wavefront walks the segment table (longest problems first) with the
  // grid's stride: one segment per wavefront, or, in a persistent grid, a
  // share of the table per resident wavefront
  for (int w = wv; w < a.nseg; w += nw) {
    const int2 sg = a.seg[w];
    const DevProb &pb = a.prob[sg.x];
    DevClock *ck = &a.clock[sg.x];
    const int n = pb.n[0];
    const int c0 = sg.y;
    const int c = c0 - 2 * MDS_NG + lane;
    const long long off = pb.off;
    const bool own = (lane >= 2 * MDS_NG) && (lane < 2 * MDS_NG + MDS_SEG1D2) && (c < n);
    if (!ck->active) {
      // frozen problem: carry its state over to the output buffer
      if (own) store_prims<NV, TAB>(out, off, c, load_prims<NV, TAB>(in, off, c, 0));
      continue;
    }
    const double dt = ck->dt;
    int flip;
    const int s = bc_src<false>(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
    const Prim<double> q = load_zone<NV, TAB, false>(pb, in, s, flip);
    Geo64 g;
    load_geo64<SPH>(pb, c, n, g);
    double U0[6];
    conserved_x64<NV, SPH, TAB>(q, g.gcov0, U0);

    const bool inside_f = (c >= 0) && (c <= n), inside_z = (c >= 0) && (c < n);
    const bool owned_face = inside_f && (lane >= 2 * MDS_NG) && (lane < 2 * MDS_NG + MDS_SEG1D2 || c == n);
    Lim1D lim;
    // stage 0: U^n -> U* on lanes 2..61
    const Roles r0{inside_f && lane >= 2 && lane <= 62, owned_face, inside_z && lane >= 2 && lane <= 61, own};
    Prim<double> h = stage_x64<NV, SPH, TAB, 0>(pb, a.eos, g, 0, dt, q, U0, q, 0.0, r0, lim);
    lim.nbad += bad_state_lanes(own, h.rho, h.uu, !TAB);
    ghost_permute<NV, TAB>(c, n, c0, pb.bc_lo[0], pb.bc_hi[0], h);
    // stage 1: U^n, U* -> U^{n+1} on lanes 4..59
    const Roles r1{inside_f && lane >= 4 && lane <= 60, owned_face, inside_z && lane >= 4 && lane <= 59, own};
    const Prim<double> o = stage_x64<NV, SPH, TAB, 0>(pb, a.eos, g, 1, dt, h, U0, q, 0.0, r1, lim);
    lim.nbad += bad_state_lanes(own, o.rho, o.uu, !TAB);
    if (own) store_prims<NV, TAB>(out, off, c, o);
    commit_limits(pb, ck, a.counters, lane, lim);
  }
}

// ---------------------------------------------------------------------------
// FP32-bulk
// ---------------------------------------------------------------------------
namespace {

struct Geo32 {
  ReconGeo32 rg;
  double area;              // lower face (FP64)
  FaceGeo<float> fg;        // FP32 face metric for the Riemann solve (unit area)
  double ivol, gcov0, cp;
  float c000, dcen;
  // spherical columns with three velocity components: the zone's transverse
  // gcov[0][1..2], gcon[1..2] gcov[0][0], and its connection coefficients
  // (MDS_TC_C* order) -- the problem's shared constants on a log-r column,
  // else the FP64 zone record
  double gc1, gc2, gi1, gi2;
  const double *zg;
  const double *tc;
};

// connection coefficient t (MDS_TC_C000 ..) of a rotating column's zone
MDS_DEV double conn_of(const Geo32 &g, int t)
{
  if (g.tc) return g.tc[t];
  const int map[MDS_TC_COUNT] = {0, 0, 0, 0, 0, 0, MDS_ZG_C000, MDS_ZG_C101, MDS_ZG_C202, MDS_ZG_C011, MDS_ZG_C110,
                                 MDS_ZG_C111, MDS_ZG_C212, MDS_ZG_C022, MDS_ZG_C122, MDS_ZG_C220, MDS_ZG_C221};
  return g.zg[map[t]];
}

template <bool SPH, int NV = 1>
MDS_DEV void load_geo32(const DevProb &pb, int c, int n, Geo32 &g)
{
  const int q = clamp_q(c, n);
  g.fg.area = 1.0f;
  g.fg.gcov[1] = g.fg.gcov[2] = 1.0f;
  g.zg = nullptr;
  g.tc = nullptr;
  g.gc1 = g.gc2 = g.gi1 = g.gi2 = 1.0;
  if (SPH) {
    const ZoneF32 &z = pb.z32[q];
    g.area = z.area;
    g.fg.scale = z.scalef;
    g.fg.gcov[0] = z.gcovf0;
    g.ivol = z.ivol;
    g.gcov0 = z.gcov0;
    g.cp = z.cp;
    g.c000 = z.c000;
    g.dcen = z.dcen;
    if (NV == 3) {
      // rotating columns: the transverse metric and connection coefficients,
      // shared by the problem on a log-r column (uniform branch), else from
      // the FP64 zone record (the radial-only path keeps its 48-byte record)
      if (pb.tcu) {
        g.tc = pb.tcu_c;
        g.gc1 = g.gcov0 * pb.tcu_c[MDS_TC_G1];
        g.gc2 = g.gcov0 * pb.tcu_c[MDS_TC_G2];
        g.gi1 = pb.tcu_c[MDS_TC_GI1];
        g.gi2 = pb.tcu_c[MDS_TC_GI2];
        g.fg.gcov[1] = g.fg.gcov[0] * (float)pb.tcu_c[MDS_TC_F1];
        g.fg.gcov[2] = g.fg.gcov[0] * (float)pb.tcu_c[MDS_TC_F2];
      } else {
        g.zg = pb.zone + (size_t)q * MDS_ZG_COUNT;
        g.gc1 = g.zg[MDS_ZG_GCOV01];
        g.gc2 = g.zg[MDS_ZG_GCOV02];
        g.gi1 = g.zg[MDS_ZG_GCON1] * g.gcov0;
        g.gi2 = g.zg[MDS_ZG_GCON2] * g.gcov0;
        g.fg.gcov[1] = (float)g.zg[MDS_ZG_GCOVF1];
        g.fg.gcov[2] = (float)g.zg[MDS_ZG_GCOVF2];
      }
    }
  } else {
    g.area = pb.area[0];
    g.fg.scale = 1.0f;
    g.fg.gcov[0] = 1.0f;
    g.ivol = pb.ivolume;
    g.gcov0 = 1.0;
    g.cp = 0.0;
    g.c000 = 0.0f;
    g.dcen = 0.5f * (float)pb.dx[0];
  }
  // one coefficient set per problem (wave-uniform: scalar loads) or per zone
  if (pb.rcu) {
    g.rg.clm = pb.rcu_c[MDS_RC_CLM]; g.rg.clp = pb.rcu_c[MDS_RC_CLP];
    g.rg.crm = pb.rcu_c[MDS_RC_CRM]; g.rg.crp = pb.rcu_c[MDS_RC_CRP];
    g.rg.mu1 = pb.rcu_c[MDS_RC_MU1]; g.rg.kinv = pb.rcu_c[MDS_RC_KINV];
    g.rg.inur = pb.rcu_c[MDS_RC_INUR]; g.rg.imu2 = pb.rcu_c[MDS_RC_IMU2];
  } else {
    const float *rc = pb.rc[0] + (size_t)q * MDS_RC_COUNT;
    g.rg.clm = rc[MDS_RC_CLM]; g.rg.clp = rc[MDS_RC_CLP]; g.rg.crm = rc[MDS_RC_CRM]; g.rg.crp = rc[MDS_RC_CRP];
    g.rg.mu1 = rc[MDS_RC_MU1]; g.rg.kinv = rc[MDS_RC_KINV]; g.rg.inur = rc[MDS_RC_INUR]; g.rg.imu2 = rc[MDS_RC_IMU2];
  }
}

// U^n in conserved form (FP64) and its primitives as doubles.
template <int NV, bool SPH, int TAB>
MDS_DEV void conserved_f32(const Geo32 &g, const Prim<float> &q, double U0[6], Prim<double> &qn)
{
  qn.rho = q.rho; qn.uu = q.uu; qn.ye = q.ye;
  for (int d = 0; d < 3; d++) qn.v[d] = q.v[d];
  const bool rot = SPH && NV == 3;
  const double gcov[3] = {g.gcov0, rot ? g.gc1 : 1.0, rot ? g.gc2 : 1.0};
  p_to_u_x64<NV, !SPH>(qn.rho, qn.uu, qn.v, gcov, U0);
  U0[MDS_YI] = TAB ? U0[0] * qn.ye : 0.0;
}

// CFL and source limits of one FP32-bulk lane.  The faces' |signal speed| /
// dx and the zone's estimate_dt bound are rates (1 / time), kept as their
// maximum; the source limits are times, kept as their minimum.  The lane's
// candidate min(1 / rate, dts) is formed once, when the wavefront reduces it:
// one reciprocal per lane instead of one per face and zone.
struct Lim32 {
  float rate = 0.0f;
  float dts = __builtin_huge_valf();
  unsigned nsh = 0;         // shock-tagged owned faces (wave-uniform)
  unsigned nbad = 0;        // owned zone updates with an unphysical state (wave-uniform)
  unsigned noff = 0;        // table EOS evaluations of owned zones off the table (wave-uniform)
};

MDS_DEV float wave_min_f32(float v)
{
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}

MDS_DEV void commit_limits32(DevClock *ck, unsigned long long *counters, int lane, const Lim32 &lim)
{
  const float w = wave_min_f32(fminf(frcp(lim.rate), lim.dts));
  if (lane == 0) {
    if (w < __builtin_huge_valf()) atomicMin(&ck->mindt, dt_key((double)w));
    if (lim.nsh) atomicAdd(&counters[0], (unsigned long long)lim.nsh);
    if (lim.nbad) atomicAdd(&ck->nbad, (unsigned long long)lim.nbad);
    if (lim.noff) atomicAdd(&ck->noff, (unsigned long long)lim.noff);
  }
}

// The update of one zone from its face fluxes (FP64, area-weighted): sources
// (geometric, fixed or monopole gravity), the RK combination, u -> p and the
// zone's CFL and source limits.
template <int NV, bool SPH, int TAB, int MONO>
MDS_DEV Prim<float> zone_update_f32(const DevProb &pb, const Geo32 &g, int stage, double dt, const Prim<float> &q,
                                    const double U0[6], const Prim<double> &qn, double gz, float press, double dpress,
                                    float cs, const double F[6], const double Fp[6], double area_hi, Lim32 &lim)
{
  const float gam = (float)pb.gam;
  Prim<float> o = q;
  const double drho = q.rho;
  double dv[3] = {q.v[0], q.v[1], q.v[2]};
  double src[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (SPH && NV == 3) {
    // rotating column: T[l][m] = rho v^m v_l + P delta, the 11 non-zero
    // conn[l][m][n] in the reference's (l, n) order (rad_fluid.c:195-214)
    const double gc[3] = {g.gcov0, g.gc1, g.gc2};
    double T[3][3];
    for (int l = 0; l < 3; l++) {
      for (int m = 0; m < 3; m++) T[l][m] = (drho * dv[m]) * (gc[l] * dv[l]);
      T[l][l] += dpress;
    }
    src[2] += T[0][0] * conn_of(g, MDS_TC_C000);
    src[2] += T[1][1] * conn_of(g, MDS_TC_C101);
    src[2] += T[2][2] * conn_of(g, MDS_TC_C202);
    src[3] += T[0][1] * conn_of(g, MDS_TC_C011);
    src[3] += T[1][0] * conn_of(g, MDS_TC_C110);
    src[3] += T[1][1] * conn_of(g, MDS_TC_C111);
    src[3] += T[2][2] * conn_of(g, MDS_TC_C212);
    src[4] += T[0][2] * conn_of(g, MDS_TC_C022);
    src[4] += T[1][2] * conn_of(g, MDS_TC_C122);
    src[4] += T[2][0] * conn_of(g, MDS_TC_C220);
    src[4] += T[2][1] * conn_of(g, MDS_TC_C221);
  } else if (SPH) {
    // radial column, v^1 only: T[0][0] conn000 + P (conn101 + conn202)
    src[2] = ((drho * dv[0]) * (g.gcov0 * dv[0])) * (double)g.c000 + dpress * g.cp;
  }
  if (pb.g0[0] != 0.0) {
    src[1] += drho * ((dv[0] * pb.g0[0]) * g.gcov0);
    src[2] += (drho * pb.g0[0]) * g.gcov0;
  }
  if (MONO) {
    // gravity.c:475-486, the linear mass flux per area evaluated at the
    // centroid offset from the lower face
    const double vl = F[0] / (g.area + 1.0e-16), vr = Fp[0] / (area_hi + 1.0e-16);
    src[2] += drho * gz;
    src[1] += (vl + ((vr - vl) * pb.idx[0]) * (double)g.dcen) * gz;
  }
  double U[6];
  for (int k = 0; k < 6; k++) U[k] = U0[k];
  {
    // dt_source_control (step.c:205-223) in FP32: there is no mass source
    // here, and the energy source is non-zero only with gravity (a uniform
    // branch per problem), so most zones spend one reciprocal on the limit
    const float s1 = fabsf((float)src[1]);
    float mom = 0.0f;
    for (int d = 0; d < NV; d++) mom += (float)qn.v[d] * (float)src[2 + d];
    float t = fminf(0.02f * (float)U[0], 0.02f * (float)U[1]) * 1.0e16f;
    if (MONO || pb.g0[0] != 0.0) t = fminf(0.02f * (float)U[0] * 1.0e16f, 0.02f * (float)U[1] * frcp(s1 + 1.0e-16f));
    t = fminf(t, (2.0f * (float)qn.uu) * frcp(fabsf(s1 - fabsf(mom)) + 1.0e-16f));
    lim.dts = fminf(lim.dts, t);
  }
  if (stage == 0) {
    // U* = U^n - dt L(U^n): the stage state is U^n itself, so the RK
    // combination 1 U^n + 0 U* needs no second conversion (same bits for
    // finite states)
    for (int k = 0; k < 2 + NV; k++) {
      double ddt = (Fp[k] - F[k]) * g.ivol - src[k];
      U[k] = U[k] - dt * ddt;
    }
    if (TAB) U[MDS_YI] = U[MDS_YI] - dt * ((Fp[MDS_YI] - F[MDS_YI]) * g.ivol);
  } else {
    const bool rot = SPH && NV == 3;
    const double gcov[3] = {g.gcov0, rot ? g.gc1 : 1.0, rot ? g.gc2 : 1.0};
    double Uh[6];
    p_to_u_x64<NV, !SPH>(drho, (double)q.uu, dv, gcov, Uh);
    Uh[MDS_YI] = TAB ? Uh[0] * (double)q.ye : 0.0;
    const double pf = 0.5, phf = 0.5, ff = 0.5;
    for (int k = 0; k < 2 + NV; k++) {
      double ddt = (Fp[k] - F[k]) * g.ivol - src[k];
      U[k] = pf * U[k] + phf * Uh[k] - (ff * dt) * ddt;
    }
    if (TAB) {
      const int k = MDS_YI;
      double ddt = (Fp[k] - F[k]) * g.ivol;
      U[k] = pf * U[k] + phf * Uh[k] - (ff * dt) * ddt;
    }
  }
  // u -> p (fluid.c:27-53): v^d = gcon[d] m_d / rho with gcon[0] = 1 / gcov0,
  // one reciprocal of rho gcov0 for the radial column
  double vout[3] = {0.0, 0.0, 0.0}, ke = 0.0;
  const double rg = drcp(SPH ? U[0] * g.gcov0 : U[0]);
  for (int d = 0; d < NV; d++) {
    // v^d = gcon[d] m_d / rho: gcon[0] = 1 / gcov0 is in rg; rg gcov0 = 1 / rho
    double f = (SPH && d == 0) ? rg : (SPH ? rg * g.gcov0 : rg);
    // rotating columns: gcon[d] / rho = rg gcov0 gcon[d] = rg gi_d
    if (SPH && NV == 3 && d > 0) f = rg * (d == 1 ? g.gi1 : g.gi2);
    vout[d] = U[2 + d] * f;
    ke += U[2 + d] * vout[d];
  }
  const double uout = U[1] - 0.5 * ke;
  o.rho = (float)U[0];
  o.uu = (float)uout;
  for (int d = 0; d < NV; d++) o.v[d] = (float)vout[d];
  if (TAB) o.ye = (float)(U[MDS_YI] * (SPH ? rg * g.gcov0 : rg));
  if (stage == 1) {
    // estimate_dt zone bound (step.c:177-187) with 1 / scale[0][0] = rsqrt(gcov0):
    // the gamma-law sound speed over the scale factor as one root
    float csi;
    if (!TAB) csi = fsqrt(gam * press * frcp(SPH ? q.rho * (float)g.gcov0 : q.rho));
    else csi = SPH ? cs * frsqrt((float)g.gcov0) : cs;
    const float loc = (fabsf((float)vout[0]) + csi) * (float)pb.idx[0];
    lim.rate = fmaxf(lim.rate, loc);
  }
  return o;
}

// One FP32-bulk RK stage of the Z consecutive zones of one lane: EOS table
// interpolation, reconstruction and Riemann fluxes in FP32 (per unit area),
// area-weighted fluxes, sources, the update and u -> p in FP64.  A zone's
// neighbours are the lane's own zones, or the edge zone of the next lane
// below or above by a DPP row shift; with Z = 2 a wavefront carries 128
// zones, and the shifts, the wave reductions and the per-lane set-up serve
// two zones each.  With the gamma law the face pressures are (gamma - 1)
// times the face energies: para_recon is homogeneous of degree one, so this
// is the reference's separate pressure reconstruction (reconstruct.c:246-266
// over P) up to rounding.
template <int NV, bool SPH, int TAB, int MONO, int Z>
MDS_DEV void stage_f32(const DevProb &pb, const DevEos &eos, const Geo32 (&g)[Z], int stage, double dt,
                       const Prim<float> (&q)[Z], const double (&U0)[Z][6], const Prim<double> (&qn)[Z],
                       const double (&gz)[Z], const Roles (&r)[Z], Lim32 &lim, Prim<float> (&o)[Z])
{
  const float gam = (float)pb.gam;
  constexpr int NQ = TAB ? 5 + NV : 2 + NV;
  float press[Z], cs[Z], qv[Z][NQ];
  double dpress[Z];
#pragma unroll
  for (int k = 0; k < Z; k++) {
    float gam1;
    cs[k] = 0.0f;
    if (TAB) {
      double lt;
      eos_fill_f32(eos, (double)q[k].rho, (double)q[k].uu, (double)q[k].ye, dpress[k], gam1, cs[k], lt);
      press[k] = (float)dpress[k];
      lim.noff += (unsigned)__popcll(__ballot(r[k].own && eos_off_table(eos, (double)q[k].rho, (double)q[k].ye, lt)));
    } else {
      press[k] = (gam - 1.0f) * q[k].uu;
      dpress[k] = press[k];
      gam1 = gam;
    }
    qv[k][0] = q[k].rho; qv[k][1] = q[k].uu;
    for (int d = 0; d < NV; d++) qv[k][2 + d] = q[k].v[d];
    if (TAB) { qv[k][2 + NV] = q[k].ye; qv[k][3 + NV] = press[k]; qv[k][4 + NV] = gam1; }
  }
  float al[Z][NQ], ar[Z][NQ];
#pragma unroll
  for (int j = 0; j < NQ; j++) {
    const float below = lane_from_below(qv[Z - 1][j]);
    const float above = lane_from_above(qv[0][j]);
#pragma unroll
    for (int k = 0; k < Z; k++)
      para_recon_f32(k > 0 ? qv[k - 1][j] : below, qv[k][j], k < Z - 1 ? qv[k + 1][j] : above, g[k].rg,
                     al[k][j], ar[k][j]);
  }
  double F[Z][6], Fp[Z][6], area_hi[Z];
#pragma unroll
  for (int k = 0; k < Z; k++) {
    // the lower face of zone k: the right state of the zone below, the left state of zone k
    float arl[NQ];
    for (int j = 0; j < NQ; j++) arl[j] = k > 0 ? ar[k - 1][j] : lane_from_below(ar[Z - 1][j]);
    PState<float, NV> L, Rt;
    L.rho = arl[0];  Rt.rho = al[k][0];
    L.uu = arl[1];   Rt.uu = al[k][1];
    for (int d = 0; d < NV; d++) { L.v[d] = arl[2 + d]; Rt.v[d] = al[k][2 + d]; }
    for (int d = NV; d < 3; d++) { L.v[d] = 0.0f; Rt.v[d] = 0.0f; }
    if (TAB) {
      L.ye = arl[2 + NV]; Rt.ye = al[k][2 + NV];
      L.P = arl[3 + NV];  Rt.P = al[k][3 + NV];
      L.G = arl[4 + NV];  Rt.G = al[k][4 + NV];
    } else {
      L.ye = Rt.ye = 0.0f;
      L.P = (gam - 1.0f) * L.uu; Rt.P = (gam - 1.0f) * Rt.uu;
      L.G = gam; Rt.G = gam;
    }
    float Ff[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, sig = 1.0f;
    int isshock = 0;
    if (g[k].area >= 1.0e-16) isshock = hllc_f32<NV, !SPH, TAB>(L, Rt, g[k].fg, 0, Ff, sig);
    else sig = 1.0e-30f;
    if (r[k].face) lim.rate = fmaxf(lim.rate, fabsf(sig) * (float)pb.idx[0]);
    lim.nsh += __popcll(__ballot(r[k].count && isshock));
    for (int v = 0; v < 6; v++) F[k][v] = 0.0;
    for (int v = 0; v < 2 + NV; v++) F[k][v] = g[k].area * (double)Ff[v];
    if (TAB) F[k][MDS_YI] = g[k].area * (double)Ff[MDS_YI];
  }
#pragma unroll
  for (int k = 0; k < Z; k++) {
    // the upper face of zone k: the lower face of the zone above
    for (int v = 0; v < 6; v++) Fp[k][v] = 0.0;
    for (int v = 0; v < 2 + NV; v++) Fp[k][v] = k < Z - 1 ? F[k + 1][v] : lane_from_above(F[0][v]);
    if (TAB) Fp[k][MDS_YI] = k < Z - 1 ? F[k + 1][MDS_YI] : lane_from_above(F[0][MDS_YI]);
    area_hi[k] = MONO ? (k < Z - 1 ? g[k + 1].area : lane_from_above(g[0].area)) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < Z; k++) {
    o[k] = q[k];
    if (r[k].out)
      o[k] = zone_update_f32<NV, SPH, TAB, MONO>(pb, g[k], stage, dt, q[k], U0[k], qn[k], gz[k], press[k], dpress[k],
                                                 cs[k], F[k], Fp[k], area_hi[k], lim);
  }
}

}  // namespace

template <int NV, bool SPH, int TAB, int MONO, int GEN>
__global__ __launch_bounds__(256) void sweep1d_f32(Sweep1DArgs a, StateArrays<float> sp,
                                                   StateArrays<float> sph, int stage)
{
  const int lane = threadIdx.x & 63;
  // wave-uniform (threadIdx.x / 64 is the same on every lane of a wave), so
  // the segment and problem records come in through scalar loads
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  // each wavefront walks the segment table (longest problems first) with the
  // grid's stride: one segment per wavefront, or, in a persistent grid, a
  // share of the table per resident wavefront
  for (int w = wv; w < a.nseg; w += nw) {
    const int2 sg = a.seg[w];
    const DevProb &pb = a.prob[sg.x];
    DevClock *ck = &a.clock[sg.x];
    if (!ck->active) continue;
    const double dt = ck->dt;
    const int n = pb.n[0];
    const int c = sg.y - MDS_NG + lane;
    const long long off = pb.off;

    const StateArrays<float> &in = (stage == 0) ? sp : sph;
    int flip;
    const int s = bc_src<GEN != 0>(c, n, pb.bc_lo[0], pb.bc_hi[0], flip);
    const Prim<float> q = load_zone<NV, TAB, GEN != 0>(pb, in, s, flip);
    Geo32 g[1];
    load_geo32<SPH, NV>(pb, c, n, g[0]);
    Roles r[1];
    r[0].face = (lane >= 2) && (lane <= 62) && (c >= 0) && (c <= n);
    r[0].count = r[0].face && (lane <= 61 || c == n);
    r[0].out = (lane >= 2) && (lane <= 61) && (c >= 0) && (c < n);
    r[0].own = r[0].out;
    Prim<float> qnf = q;
    if (stage == 1 && r[0].out) qnf = load_prims<NV, TAB>(sp, off, c, 0);
    double U0[1][6];
    Prim<double> qn[1];
    conserved_f32<NV, SPH, TAB>(g[0], qnf, U0[0], qn[0]);
    const double gz[1] = {(MONO && r[0].out) ? a.grav[off + c] : 0.0};

    Lim32 lim;
    const Prim<float> qa[1] = {q};
    Prim<float> o[1];
    stage_f32<NV, SPH, TAB, MONO, 1>(pb, a.eos, g, stage, dt, qa, U0, qn, gz, r, lim, o);
    lim.nbad += bad_state_lanes(r[0].out, o[0].rho, o[0].uu, !TAB);
    if (r[0].out) {
      StateArrays<float> &out = (stage == 0) ? sph : sp;
      store_prims<NV, TAB>(out, off, c, o[0]);
    }
    commit_limits32(ck, a.counters, lane, lim);
  }
}

// Z zones per lane: the wave holds zones c0 - 4 + p, p = Z lane + k in
// [0, 64 Z); stage 0 is valid for p in [2, 64 Z - 3], stage 1 for the owned
// p in [4, 64 Z - 5] (64 Z - 8 zones per segment).
template <int NV, bool SPH, int TAB, int Z>
__global__ __launch_bounds__(256) void sweep1d2_f32(Sweep1DArgs a, StateArrays<float> in,
                                                    StateArrays<float> out)
{
  constexpr int W = 64 * Z, SEG = W - 4 * MDS_NG;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  // each wavefront walks the segment table (longest problems first) with the
  // grid's stride: one segment per wavefront, or, in a persistent grid, a
  // share of the table per resident wavefront
  for (int w = wv; w < a.nseg; w += nw) {
    const int2 sg = a.seg[w];
    const DevProb &pb = a.prob[sg.x];
    DevClock *ck = &a.clock[sg.x];
    const int n = pb.n[0];
    const int c0 = sg.y;
    const long long off = pb.off;
    int cz[Z];
    bool own[Z];
#pragma unroll
    for (int k = 0; k < Z; k++) {
      const int p = Z * lane + k;
      cz[k] = c0 - 2 * MDS_NG + p;
      own[k] = (p >= 2 * MDS_NG) && (p < 2 * MDS_NG + SEG) && (cz[k] < n);
    }
    if (!ck->active) {
#pragma unroll
      for (int k = 0; k < Z; k++)
        if (own[k]) store_prims<NV, TAB>(out, off, cz[k], load_prims<NV, TAB>(in, off, cz[k], 0));
      continue;
    }
    const double dt = ck->dt;
    Prim<float> q[Z];
    Geo32 g[Z];
    double U0[Z][6];
    Prim<double> qn[Z];
    Roles r0[Z], r1[Z];
#pragma unroll
    for (int k = 0; k < Z; k++) {
      int flip;
      const int s = bc_src<false>(cz[k], n, pb.bc_lo[0], pb.bc_hi[0], flip);
      q[k] = load_zone<NV, TAB, false>(pb, in, s, flip);
      load_geo32<SPH, NV>(pb, cz[k], n, g[k]);
      conserved_f32<NV, SPH, TAB>(g[k], q[k], U0[k], qn[k]);
      const int p = Z * lane + k;
      const bool inside_f = (cz[k] >= 0) && (cz[k] <= n), inside_z = (cz[k] >= 0) && (cz[k] < n);
      const bool owned_face = inside_f && (p >= 2 * MDS_NG) && (p < 2 * MDS_NG + SEG || cz[k] == n);
      r0[k] = Roles{inside_f && p >= 2 && p <= W - 2, owned_face, inside_z && p >= 2 && p <= W - 3, own[k]};
      r1[k] = Roles{inside_f && p >= 4 && p <= W - 4, owned_face, inside_z && p >= 4 && p <= W - 5, own[k]};
    }
    const double gz[Z] = {};
    Lim32 lim;
    Prim<float> h[Z], o[Z];
    stage_f32<NV, SPH, TAB, 0, Z>(pb, a.eos, g, 0, dt, q, U0, qn, gz, r0, lim, h);
#pragma unroll
    for (int k = 0; k < Z; k++) lim.nbad += bad_state_lanes(own[k], h[k].rho, h[k].uu, !TAB);
    // boundary ghosts of the intermediate state, in the waves that hold some
    if (c0 - 2 * MDS_NG < 0 || c0 - 2 * MDS_NG + W > n) ghost_permute<NV, TAB, Z>(cz, n, c0, pb.bc_lo[0], pb.bc_hi[0], h);
    stage_f32<NV, SPH, TAB, 0, Z>(pb, a.eos, g, 1, dt, h, U0, qn, gz, r1, lim, o);
#pragma unroll
    for (int k = 0; k < Z; k++) {
      lim.nbad += bad_state_lanes(own[k], o[k].rho, o[k].uu, !TAB);
      if (own[k]) store_prims<NV, TAB>(out, off, cz[k], o[k]);
    }
    commit_limits32(ck, a.counters, lane, lim);
  }
}

// ---------------------------------------------------------------------------
// monopole gravity field (gravity.c:145-251 start, 420-450 finish), one
// workgroup per problem: shell masses V rho of the stage state, enclosed
// mass by a block scan, g = -G M / r^2 dr/dx at the faces, linear in x
// across each zone evaluated at the volume centroid beta/Gamma.
// ---------------------------------------------------------------------------
#define MDS_MONO_THREADS 256

template <class R>
__global__ __launch_bounds__(MDS_MONO_THREADS) void monopole_kernel(const DevProb *prob, StateArrays<R> st,
                                                                     double *grav)
{
  const DevProb &pb = prob[blockIdx.x];
  const int n = pb.n[0];
  const int t = threadIdx.x;
  const int chunk = (n + MDS_MONO_THREADS - 1) / MDS_MONO_THREADS;
  const int i0 = t * chunk < n ? t * chunk : n, i1 = (t + 1) * chunk < n ? (t + 1) * chunk : n;
  const long long off = pb.off;
  double part = 0.0;
  for (int i = i0; i < i1; i++)
    part = part + pb.zone[(size_t)(i + MDS_NG) * MDS_ZG_COUNT + MDS_ZG_VOLUME] * (double)st.v[0][off + i];
  // exclusive block scan of the chunk masses
  __shared__ double sm[MDS_MONO_THREADS];
  sm[t] = part;
  __syncthreads();
  for (int o = 1; o < MDS_MONO_THREADS; o <<= 1) {
    double v = t >= o ? sm[t - o] : 0.0;
    __syncthreads();
    sm[t] += v;
    __syncthreads();
  }
  double mr = sm[t] - part;
  const bool logr = pb.rcoord == MDS_RCOORD_LOG;
  const double GN = 6.67384e-8;
  auto edge = [&](int ii, double m) {
    if (ii == 0) return 0.0;
    const double xe = pb.startx[0] + ii * pb.dx[0];
    const double r = logr ? exp(xe) : fabs(xe);
    const double drdx = logr ? exp(xe) : 1.0;
    return ((-GN * m) / (r * r)) * drdx;
  };
  double ge = edge(i0, mr);
  for (int i = i0; i < i1; i++) {
    mr = mr + pb.zone[(size_t)(i + MDS_NG) * MDS_ZG_COUNT + MDS_ZG_VOLUME] * (double)st.v[0][off + i];
    const double gn = edge(i + 1, mr);
    const double xl = pb.startx[0] + i * pb.dx[0];
    const double A = (gn - ge) / pb.dx[0];
    const double B = ge - A * xl;
    const double be = pb.mom[0][3 * (i + MDS_NG) + 1], ga = pb.mom[0][3 * (i + MDS_NG) + 2];
    grav[off + i] = (A * be) / ga + B;
    ge = gn;
  }
}

// ---------------------------------------------------------------------------
// GR effective monopole field of spherical columns (gravity.c:142-375 with
// GR_MONOPOLE), one workgroup per problem: the shell sums V rho, V P, V u,
// V v^1 of the stage state (one zone per shell; P from the gamma law or the
// FP64 table EOS, as update_eos gives the reference), then the field
// (mds_gravity.h).  Problems whose clock has stopped keep their field.
// ---------------------------------------------------------------------------
template <class R, int TAB>
__global__ __launch_bounds__(MDS_MONO_THREADS) void monopole_gr_kernel(const DevProb *prob, const DevClock *clock,
                                                                        DevEos eos, StateArrays<R> st, double *grav)
{
  const DevProb &pb = prob[blockIdx.x];
  if (!clock[blockIdx.x].active) return;
  const int n = pb.n[0];
  const long long off = pb.off;
  double *rs = pb.gr + (size_t)MDS_GR_RHO * (n + 1), *ps = pb.gr + (size_t)MDS_GR_P * (n + 1);
  double *us = pb.gr + (size_t)MDS_GR_U * (n + 1), *vs = pb.gr + (size_t)MDS_GR_V * (n + 1);
  for (int i = threadIdx.x; i < n; i += MDS_MONO_THREADS) {
    const double vol = pb.zone[(size_t)(i + MDS_NG) * MDS_ZG_COUNT + MDS_ZG_VOLUME];
    const double rho = (double)st.v[0][off + i], u = (double)st.v[1][off + i], v1 = (double)st.v[2][off + i];
    double press;
    if (TAB) {
      double g1, cs, lt;
      eos_fill_x64(eos, rho, u, (double)st.v[5][off + i], press, g1, cs, lt);
    } else {
      press = (pb.gam - 1.0) * u;
    }
    rs[i] = vol * rho;
    ps[i] = vol * press;
    us[i] = vol * u;
    vs[i] = vol * v1;
  }
  __syncthreads();
  gr_monopole_block<MDS_MONO_THREADS>(pb, grav + off);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// mono: MDS_MONOPOLE_NEWTON or MDS_MONOPOLE_GR (then clock and eos are read;
// tab: the batch has a table EOS)
extern "C" hipError_t mds_launch_monopole(int fp32, int mono, int tab, int nprob, const DevProb *prob,
                                          const DevClock *clock, const DevEos *eos, const void *st, double *grav,
                                          hipStream_t s)
{
  const dim3 g(nprob), b(MDS_MONO_THREADS);
  if (mono == MDS_MONOPOLE_GR) {
    if (fp32) {
      const StateArrays<float> &a = *(const StateArrays<float> *)st;
      if (tab) hipLaunchKernelGGL((monopole_gr_kernel<float, 1>), g, b, 0, s, prob, clock, *eos, a, grav);
      else hipLaunchKernelGGL((monopole_gr_kernel<float, 0>), g, b, 0, s, prob, clock, *eos, a, grav);
    } else {
      const StateArrays<double> &a = *(const StateArrays<double> *)st;
      if (tab) hipLaunchKernelGGL((monopole_gr_kernel<double, 1>), g, b, 0, s, prob, clock, *eos, a, grav);
      else hipLaunchKernelGGL((monopole_gr_kernel<double, 0>), g, b, 0, s, prob, clock, *eos, a, grav);
    }
  } else if (fp32) {
    hipLaunchKernelGGL(monopole_kernel<float>, g, b, 0, s, prob, *(const StateArrays<float> *)st, grav);
  } else {
    hipLaunchKernelGGL(monopole_kernel<double>, g, b, 0, s, prob, *(const StateArrays<double> *)st, grav);
  }
  return hipGetLastError();
}

// workgroups of `block` threads of kernel k that one compute unit holds at once
template <class K>
static int resident_blocks(K k, int block)
{
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, block, 0) != hipSuccess || nb < 1) nb = 1;
  return nb;
}

// stage 0/1: single-stage kernels on (p, ph); stage -1: the two-stage kernel
// with sp = input buffer and sph_ = output buffer (a->seg must then be the
// MDS_SEG1D2 segment table, or the MDS_SEG1D2Z table when a->zpl is 2).  tab: tabulated EOS with Ye; mono: monopole
// gravity (single-stage kernels only, a->grav filled for the stage input).
extern "C" hipError_t mds_launch_sweep1d(int fp32, int nv, int sph, int tab, int mono, const Sweep1DArgs *a,
                                         const void *sp, const void *sph_, int stage, hipStream_t st)
{
  const int waves_per_block = 4;
  dim3 block(64 * waves_per_block);
  const int need = (a->nseg + waves_per_block - 1) / waves_per_block;
  if ((tab || mono) && !sph) return hipErrorInvalidValue;
  if (mono && stage < 0) return hipErrorInvalidValue;
  // one wavefront per segment, or a persistent grid: the workgroups the
  // device holds at once (occupancy of the kernel, once per instantiation)
#define MDS_GRID(K)                                                                                    \
  static const int nres = resident_blocks(K, 64 * waves_per_block);                                   \
  dim3 grid((unsigned)std::max(1, a->persist ? std::min(need, nres * std::max(1, a->ncu)) : need));
#define MDS_L1G(K, ...) do { MDS_GRID((K<__VA_ARGS__>)) hipLaunchKernelGGL((K<__VA_ARGS__>), grid, block, 0, st, *a, p, h, stage); } while (0)
// the general instances (PROB/EXTRAP ghost layers, hydro_interp_order < 3)
#define MDS_L1(K, ...) do { if (a->gen) MDS_L1G(K, __VA_ARGS__, 1); else MDS_L1G(K, __VA_ARGS__, 0); } while (0)
#define MDS_L2(K, ...) do { MDS_GRID((K<__VA_ARGS__>)) hipLaunchKernelGGL((K<__VA_ARGS__>), grid, block, 0, st, *a, p, h); } while (0)
  if (!fp32) {
    const StateArrays<double> &p = *(const StateArrays<double> *)sp;
    const StateArrays<double> &h = *(const StateArrays<double> *)sph_;
    if (stage < 0) {
      if (sph && tab) MDS_L2(sweep1d2_x64, 3, true, 1);
      else if (sph) MDS_L2(sweep1d2_x64, 3, true, 0);
      else if (nv == 1) MDS_L2(sweep1d2_x64, 1, false, 0);
      else MDS_L2(sweep1d2_x64, 3, false, 0);
    } else {
      if (sph && tab && mono) MDS_L1(sweep1d_x64, 3, true, 1, 1);
      else if (sph && tab) MDS_L1(sweep1d_x64, 3, true, 1, 0);
      else if (sph && mono) MDS_L1(sweep1d_x64, 3, true, 0, 1);
      else if (sph) MDS_L1(sweep1d_x64, 3, true, 0, 0);
      else if (nv == 1) MDS_L1(sweep1d_x64, 1, false, 0, 0);
      else MDS_L1(sweep1d_x64, 3, false, 0, 0);
    }
  } else {
    const StateArrays<float> &p = *(const StateArrays<float> *)sp;
    const StateArrays<float> &h = *(const StateArrays<float> *)sph_;
    // spherical columns carry v^1 (nv 1) or all three components (nv 3, rotating columns)
    if (stage < 0 && a->zpl == 2) {
      if (sph && nv == 3 && tab) MDS_L2(sweep1d2_f32, 3, true, 1, 2);
      else if (sph && nv == 3) MDS_L2(sweep1d2_f32, 3, true, 0, 2);
      else if (sph && tab) MDS_L2(sweep1d2_f32, 1, true, 1, 2);
      else if (sph) MDS_L2(sweep1d2_f32, 1, true, 0, 2);
      else if (nv == 1) MDS_L2(sweep1d2_f32, 1, false, 0, 2);
      else MDS_L2(sweep1d2_f32, 3, false, 0, 2);
    } else if (stage < 0) {
      if (sph && nv == 3 && tab) MDS_L2(sweep1d2_f32, 3, true, 1, 1);
      else if (sph && nv == 3) MDS_L2(sweep1d2_f32, 3, true, 0, 1);
      else if (sph && tab) MDS_L2(sweep1d2_f32, 1, true, 1, 1);
      else if (sph) MDS_L2(sweep1d2_f32, 1, true, 0, 1);
      else if (nv == 1) MDS_L2(sweep1d2_f32, 1, false, 0, 1);
      else MDS_L2(sweep1d2_f32, 3, false, 0, 1);
    } else if (sph && nv == 3) {
      if (tab && mono) MDS_L1(sweep1d_f32, 3, true, 1, 1);
      else if (tab) MDS_L1(sweep1d_f32, 3, true, 1, 0);
      else if (mono) MDS_L1(sweep1d_f32, 3, true, 0, 1);
      else MDS_L1(sweep1d_f32, 3, true, 0, 0);
    } else {
      if (sph && tab && mono) MDS_L1(sweep1d_f32, 1, true, 1, 1);
      else if (sph && tab) MDS_L1(sweep1d_f32, 1, true, 1, 0);
      else if (sph && mono) MDS_L1(sweep1d_f32, 1, true, 0, 1);
      else if (sph) MDS_L1(sweep1d_f32, 1, true, 0, 0);
      else if (nv == 1) MDS_L1(sweep1d_f32, 1, false, 0, 0);
      else MDS_L1(sweep1d_f32, 3, false, 0, 0);
    }
  }
#undef MDS_L1
#undef MDS_L1G
#undef MDS_L2
#undef MDS_GRID
  return hipGetLastError();
}
