// mds_sweep_common.h -- device helpers shared by the 2D and 3D Cartesian
// stage kernels: boundary index maps, zone records, reconstruction of a
// zone, Riemann solves, the x sweep of a row through DPP neighbours and the
// shock tags of the radial faces.
#pragma once
#include <type_traits>
#include "mds_internal.h"
#include "mds_physics.h"

namespace mds {

// Map a stencil index onto the interior zone the reference's ghost fill
// copies there (bc.c:705-1421).  In direction 2 REFLECT copies the first /
// last interior zone instead of mirroring (bc.c:1285-1300, 1360-1377).
MDS_DEV int bc_src2(int c, int n, int lo, int hi, bool &flip, bool dir2 = false)
{
  flip = false;
  if (c < 0) {
    if (c < -MDS_NG) return 0;
    if (lo == MDS_BC_REFLECT) { flip = true; return dir2 ? 0 : -1 - c; }
    if (lo == MDS_BC_PERIODIC) return c + n;
    return 0;
  }
  if (c >= n) {
    if (c >= n + MDS_NG) return n - 1;
    if (hi == MDS_BC_REFLECT) { flip = true; return dir2 ? n - 1 : 2 * n - 1 - c; }
    if (hi == MDS_BC_PERIODIC) return c - n;
    return n - 1;
  }
  return c;
}

template <class R, int NV>
struct Cell {
  R rho, uu, v[NV], P;
};

template <class R>
struct ReconGeoT;
template <>
struct ReconGeoT<double> { using type = ReconGeo64; };
template <>
struct ReconGeoT<float> { using type = ReconGeo32; };

template <class R>
MDS_DEV void recon(R qm, R q0, R qp, const typename ReconGeoT<R>::type &g, R &al, R &ar)
{
  if constexpr (std::is_same<R, double>::value) para_recon_x64(qm, q0, qp, g, al, ar);
  else para_recon_f32(qm, q0, qp, g, al, ar);
}

// reconstruction geometry of zone z (index without ghost offset) in direction d
template <class R>
MDS_DEV void load_rgeo(const DevProb &pb, int d, int z, typename ReconGeoT<R>::type &g)
{
  int nz = pb.n[d] + 2 * MDS_NG;
  int q = z + MDS_NG;
  q = q < 1 ? 1 : (q > nz - 2 ? nz - 2 : q);
  if constexpr (std::is_same<R, double>::value) {
    const double *m = pb.mom[d];
    g.am = m[3 * (q - 1) + 0]; g.bm = m[3 * (q - 1) + 1]; g.gm = m[3 * (q - 1) + 2];
    g.a0 = m[3 * q + 0];       g.b0 = m[3 * q + 1];       g.g0 = m[3 * q + 2];
    g.ap = m[3 * (q + 1) + 0]; g.bp = m[3 * (q + 1) + 1]; g.gp = m[3 * (q + 1) + 2];
    g.xlo = pb.startx[d] + z * pb.dx[d];
    g.xhi = pb.startx[d] + (z + 1) * pb.dx[d];
    recon_geo_x64(g);
    g.order = pb.order;
  } else {
    const float *r = pb.rc[d] + (size_t)q * MDS_RC_COUNT;
    g.clm = r[MDS_RC_CLM]; g.clp = r[MDS_RC_CLP]; g.crm = r[MDS_RC_CRM]; g.crp = r[MDS_RC_CRP];
    g.mu1 = r[MDS_RC_MU1]; g.kinv = r[MDS_RC_KINV]; g.inur = r[MDS_RC_INUR]; g.imu2 = r[MDS_RC_IMU2];
  }
}

// A stencil zone in a ghost layer of a PROB or EXTRAP side: its entries
// (entry v at e[v * T]).  The first such direction decides; the transverse
// position is clamped into the grid, so corner zones (which no output reads)
// stay in bounds.
MDS_DEV bool slab_entry(const DevProb &pb, int i, int j, int k, const double *&e, long long &T)
{
  const int c[3] = {i, j, k};
  for (int d = 0; d < pb.ndim; d++) {
    const int n = pb.n[d];
    int side, g;
    if (c[d] < 0 && c[d] >= -MDS_NG) { side = 0; g = c[d] + MDS_NG; }
    else if (c[d] >= n && c[d] < n + MDS_NG) { side = 1; g = c[d] - n; }
    else continue;
    const double *sl = pb.gs[d][side];
    if (!sl) continue;
    const int q0 = min(max(i, 0), pb.n[0] - 1), q1 = min(max(j, 0), pb.n[1] - 1), q2 = min(max(k, 0), pb.n[2] - 1);
    T = slab_T(pb, d);
    e = sl + (size_t)g * MDS_SLAB_NV * T + slab_t(pb, d, q0, q1, q2);
    return true;
  }
  return false;
}

template <class R, int NV, bool SLAB = false>
MDS_DEV void load_cell(const StateArrays<R> &in, const DevProb &pb, int i, int j, int k, R gm1, Cell<R, NV> &c)
{
  if constexpr (SLAB) {
    const double *e;
    long long T;
    if (slab_entry(pb, i, j, k, e, T)) {
      c.rho = (R)e[0];
      c.uu = (R)e[T];
      for (int d = 0; d < NV; d++) c.v[d] = (R)e[(2 + d) * T];
      c.P = gm1 * c.uu;
      return;
    }
  }
  bool fx, fy, fz = false;
  int si = bc_src2(i, pb.n[0], pb.bc_lo[0], pb.bc_hi[0], fx);
  int sj = bc_src2(j, pb.n[1], pb.bc_lo[1], pb.bc_hi[1], fy);
  int sk = 0;
  if (pb.ndim > 2) sk = bc_src2(k, pb.n[2], pb.bc_lo[2], pb.bc_hi[2], fz, true);
  long long idx = pb.off + ((long long)sk * pb.n[1] + sj) * pb.n[0] + si;
  c.rho = in.v[0][idx];
  c.uu = in.v[1][idx];
  for (int d = 0; d < NV; d++) c.v[d] = in.v[2 + d][idx];
  if (fx) c.v[0] = -c.v[0];
  if (fy && NV > 1) c.v[1] = -c.v[1];
  if (fz && NV > 2) c.v[2] = -c.v[2];
  c.P = gm1 * c.uu;
}

// internal energy of a zone only (P of a neighbour for the lane-63 shock tag)
template <class R, bool SLAB = false>
MDS_DEV R load_press(const StateArrays<R> &in, const DevProb &pb, int i, int j, int k, R gm1)
{
  if constexpr (SLAB) {
    const double *e;
    long long T;
    if (slab_entry(pb, i, j, k, e, T)) return gm1 * (R)e[T];
  }
  bool f;
  int si = bc_src2(i, pb.n[0], pb.bc_lo[0], pb.bc_hi[0], f);
  int sj = bc_src2(j, pb.n[1], pb.bc_lo[1], pb.bc_hi[1], f);
  int sk = 0;
  if (pb.ndim > 2) sk = bc_src2(k, pb.n[2], pb.bc_lo[2], pb.bc_hi[2], f, true);
  return gm1 * in.v[1][pb.off + ((long long)sk * pb.n[1] + sj) * pb.n[0] + si];
}

// Reconstruct all variables of a zone from its two neighbours.  The FP32
// path derives the internal-energy face values from the pressure ones: the
// reconstruction is positively homogeneous, so recon(u) = recon(P)/(gam-1)
// up to rounding.
// GEN instances (FP64) also serve hydro_interp_order 2 and 1, with one
// uniform branch for the whole zone; the plain instances run para_recon only.
template <class R, int NV, bool GEN = false>
MDS_DEV void recon_cell(const Cell<R, NV> &m, const Cell<R, NV> &c, const Cell<R, NV> &p,
                        const typename ReconGeoT<R>::type &g, R igm1, Cell<R, NV> &al, Cell<R, NV> &ar)
{
  if constexpr (std::is_same<R, double>::value && GEN) {
    if (g.order != 3) {
      recon_x64(m.rho, c.rho, p.rho, g, al.rho, ar.rho);
      for (int d = 0; d < NV; d++) recon_x64(m.v[d], c.v[d], p.v[d], g, al.v[d], ar.v[d]);
      recon_x64(m.P, c.P, p.P, g, al.P, ar.P);
      recon_x64(m.uu, c.uu, p.uu, g, al.uu, ar.uu);
      return;
    }
  }
  recon<R>(m.rho, c.rho, p.rho, g, al.rho, ar.rho);
  for (int d = 0; d < NV; d++) recon<R>(m.v[d], c.v[d], p.v[d], g, al.v[d], ar.v[d]);
  recon<R>(m.P, c.P, p.P, g, al.P, ar.P);
  if constexpr (std::is_same<R, double>::value) {
    recon<R>(m.uu, c.uu, p.uu, g, al.uu, ar.uu);
  } else {
    al.uu = al.P * igm1;
    ar.uu = ar.P * igm1;
  }
}

// Riemann solve in direction d; returns isshock.  FP64: area-weighted flux
// in reference order.  FP32: flux per unit area (the FP64 accumulate applies
// the area).  UNIG: the gamma law's wave-speed factors come from gk instead
// of being formed per face.  |signal speed| returned in ssig.
template <class R, int NV, bool UNIG = true>
MDS_DEV int solve(const Cell<R, NV> &l, const Cell<R, NV> &r, R gam, const GamSet &gk, double area, int d,
                  bool use_hlle, R F[5], R &ssig)
{
  PState<R, NV> L, Rt;
  L.rho = l.rho; L.uu = l.uu; L.P = l.P; L.G = gam;
  Rt.rho = r.rho; Rt.uu = r.uu; Rt.P = r.P; Rt.G = gam;
  for (int q = 0; q < 3; q++) { L.v[q] = q < NV ? l.v[q] : (R)0; Rt.v[q] = q < NV ? r.v[q] : (R)0; }
  int sh = 0;
  FaceGeo<R> g;
  g.scale = (R)1; g.gcov[0] = g.gcov[1] = g.gcov[2] = (R)1;
  if constexpr (std::is_same<R, double>::value) {
    g.area = area;
    double s = 0.0;
    if (use_hlle) hlle_x64<NV, true, 0, UNIG>(L, Rt, g, d, F, s, &gk.f64);
    else sh = hllc_x64<NV, true, 0, UNIG>(L, Rt, g, d, F, s, &gk.f64);
    ssig = fabs(s);
  } else {
    g.area = 1.0f;
    float s = 0.0f;
    if (use_hlle) hlle_f32<NV, true, 0, UNIG>(L, Rt, g, d, F, s, &gk.f32);
    else sh = hllc_f32<NV, true, 0, UNIG>(L, Rt, g, d, F, s, &gk.f32);
    ssig = fabsf(s);
  }
  return sh;
}

// x sweep of one row: per-lane flux through the lower x face, its shock tag
// and |signal speed|.  Lane 63 additionally needs P of the zone above it so
// that the tag of its own face is exact (pextra).
template <class R, int NV, bool SLAB = false, bool UNIG = true>
MDS_DEV void xsweep(const Cell<R, NV> &me, R pextra, const typename ReconGeoT<R>::type &rgx, R gam,
                    const GamSet &gk, R igm1, double area, R F[5], R &ssig, int &isshock)
{
  const int lane = threadIdx.x & 63;
  Cell<R, NV> mm, pp, al, ar;
  mm.rho = lane_from_below(me.rho); pp.rho = lane_from_above(me.rho);
  for (int d = 0; d < NV; d++) { mm.v[d] = lane_from_below(me.v[d]); pp.v[d] = lane_from_above(me.v[d]); }
  mm.P = lane_from_below(me.P); pp.P = lane_from_above(me.P);
  if (lane == 63) pp.P = pextra;
  if constexpr (std::is_same<R, double>::value) {
    mm.uu = lane_from_below(me.uu);
    pp.uu = lane_from_above(me.uu);
  } else {
    mm.uu = pp.uu = (R)0;
  }
  recon_cell<R, NV, SLAB>(mm, me, pp, rgx, igm1, al, ar);
  Cell<R, NV> L;
  L.rho = lane_from_below(ar.rho);
  L.uu = lane_from_below(ar.uu);
  for (int d = 0; d < NV; d++) L.v[d] = lane_from_below(ar.v[d]);
  L.P = lane_from_below(ar.P);
  isshock = solve<R, NV, UNIG>(L, al, gam, gk, area, 0, false, F, ssig);
}

// shock tags only (P reconstruction + the HLLC pressure-jump test, riemann.c:374-383)
template <class R, bool GEN = false>
MDS_DEV int xshock_only(R P, R pextra, const typename ReconGeoT<R>::type &rgx)
{
  const int lane = threadIdx.x & 63;
  R qm = lane_from_below(P), qp = lane_from_above(P);
  if (lane == 63) qp = pextra;
  R al, ar;
  if constexpr (std::is_same<R, double>::value && GEN) recon_x64(qm, P, qp, rgx, al, ar);
  else recon<R>(qm, P, qp, rgx, al, ar);
  R p_l = lane_from_below(ar), p_r = al;
  R p_max = p_l > p_r ? p_l : p_r, p_min = p_l > p_r ? p_r : p_l;
  if constexpr (std::is_same<R, double>::value) return ((p_max - p_min) / p_min > 0.5) ? 1 : 0;
  else return ((p_max - p_min) > 0.5f * p_min) ? 1 : 0;
}

// cell tag from face tags: zone l is tagged by faces l, l+1, l+2 (step.c:711-717)
MDS_DEV int zone_tag(int face_shock)
{
  int a1 = lane_from_above(face_shock);
  int a2 = lane_from_above(a1);
  return face_shock | a1 | a2;
}

}  // namespace mds
