// mds_sweep2d.hip -- fused RK stage for 2D Cartesian grids (gfx950).
//
// Decomposition: one 64-lane wavefront owns an x-strip of 64 zones
// [c0-2, c0+62) and marches in y over a chunk of rows [j0, j0+CH), producing
// the update of the 60 zones [c0, c0+60) of every row of the chunk.
//   x direction: each lane reconstructs its own zone; neighbours come from
//                cross-lane shuffles; the lane solves its lower x face.
//   y direction: a 5-row register window slides along y; each zone is
//                reconstructed once and each y face solved once.
// The x sweep runs one row ahead of the update so that the shock tags of the
// row above are known before the y face between them is solved (two-pass
// reading of step.c:711-755).  Boundary ghosts are mapped onto interior zones
// at load time (bc.c:705-1421); ghost zones count as shocked (their tags are
// never reset from malloc_rank1's 0, utils.c:171, fluid.c:82-89), so the
// y faces at j=0 and j=n2 use HLLE as in the reference.
// Replaces prim_to_fluxes2 (step.c:659-764), push (step.c:327-478) and the
// per-stage boundary/EOS/source passes for 2D Cartesian problems.
#include <type_traits>
#include "mds_internal.h"
#include "mds_physics.h"


namespace {

MDS_DEV int bc_src2(int c, int n, int lo, int hi, bool &flip)
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
  } else {
    const float *r = pb.rc[d] + (size_t)q * MDS_RC_COUNT;
    g.clm = r[MDS_RC_CLM]; g.clp = r[MDS_RC_CLP]; g.crm = r[MDS_RC_CRM]; g.crp = r[MDS_RC_CRP];
    g.mu1 = r[MDS_RC_MU1]; g.kinv = r[MDS_RC_KINV]; g.inur = r[MDS_RC_INUR]; g.imu2 = r[MDS_RC_IMU2];
  }
}

template <class R, int NV>
MDS_DEV void load_cell(const StateArrays<R> &in, const DevProb &pb, int i, int j, Cell<R, NV> &c)
{
  bool fx, fy;
  int si = bc_src2(i, pb.n[0], pb.bc_lo[0], pb.bc_hi[0], fx);
  int sj = bc_src2(j, pb.n[1], pb.bc_lo[1], pb.bc_hi[1], fy);
  long long idx = pb.off + (long long)sj * pb.n[0] + si;
  c.rho = in.v[0][idx];
  c.uu = in.v[1][idx];
  for (int d = 0; d < NV; d++) c.v[d] = in.v[2 + d][idx];
  if (fx) c.v[0] = -c.v[0];
  if (fy && NV > 1) c.v[1] = -c.v[1];
  c.P = ((R)pb.gam - (R)1.0) * c.uu;
}

// Riemann solve in direction d with the proper precision; returns isshock.
// F is area weighted and in double for both precisions.
template <class R, int NV>
MDS_DEV int solve(const Cell<R, NV> &l, const Cell<R, NV> &r, R gam, double area, int d, bool use_hlle,
                  double F[5], double &sig)
{
  PState<R, NV> L, Rt;
  L.rho = l.rho; L.uu = l.uu; L.P = l.P; L.G = gam;
  Rt.rho = r.rho; Rt.uu = r.uu; Rt.P = r.P; Rt.G = gam;
  for (int q = 0; q < 3; q++) { L.v[q] = q < NV ? l.v[q] : (R)0; Rt.v[q] = q < NV ? r.v[q] : (R)0; }
  int sh = 0;
  if constexpr (std::is_same<R, double>::value) {
    FaceGeo<double> g;
    g.area = area; g.scale = 1.0; g.gcov[0] = g.gcov[1] = g.gcov[2] = 1.0;
    if (use_hlle) hlle_x64<NV, true>(L, Rt, g, d, F, sig);
    else sh = hllc_x64<NV, true>(L, Rt, g, d, F, sig);
  } else {
    FaceGeo<float> g;
    g.area = 1.0f; g.scale = 1.0f; g.gcov[0] = g.gcov[1] = g.gcov[2] = 1.0f;
    float f[5] = {0, 0, 0, 0, 0}, s = 1.0f;
    if (use_hlle) hlle_f32<NV, true>(L, Rt, g, d, f, s);
    else sh = hllc_f32<NV, true>(L, Rt, g, d, f, s);
    for (int q = 0; q < 2 + NV; q++) F[q] = area * (double)f[q];
    sig = (double)s;
  }
  return sh;
}

// x sweep of one row: per-lane flux through the lower x face, and the shock
// tag of the face.  Lane 63 additionally needs P of the zone above it so
// that the tag of its own face is exact (pextra).
template <class R, int NV>
MDS_DEV void xsweep(const Cell<R, NV> &me, R pextra, const typename ReconGeoT<R>::type &rgx, R gam,
                    double area, double F[5], double &sig, int &isshock)
{
  const int lane = threadIdx.x & 63;
  Cell<R, NV> al, ar;
  R qm, qp;
  qm = lane_from_below(me.rho); qp = lane_from_above(me.rho); recon<R>(qm, me.rho, qp, rgx, al.rho, ar.rho);
  qm = lane_from_below(me.uu);  qp = lane_from_above(me.uu);  recon<R>(qm, me.uu, qp, rgx, al.uu, ar.uu);
  for (int d = 0; d < NV; d++) {
    qm = lane_from_below(me.v[d]); qp = lane_from_above(me.v[d]);
    recon<R>(qm, me.v[d], qp, rgx, al.v[d], ar.v[d]);
  }
  qm = lane_from_below(me.P); qp = lane_from_above(me.P);
  if (lane == 63) qp = pextra;
  recon<R>(qm, me.P, qp, rgx, al.P, ar.P);
  Cell<R, NV> L;
  L.rho = lane_from_below(ar.rho);
  L.uu = lane_from_below(ar.uu);
  for (int d = 0; d < NV; d++) L.v[d] = lane_from_below(ar.v[d]);
  L.P = lane_from_below(ar.P);
  isshock = solve<R, NV>(L, al, gam, area, 0, false, F, sig);
}

// shock tags only (P reconstruction + the HLLC pressure-jump test, riemann.c:374-383)
template <class R>
MDS_DEV int xshock_only(R P, R pextra, const typename ReconGeoT<R>::type &rgx)
{
  const int lane = threadIdx.x & 63;
  R qm = lane_from_below(P), qp = lane_from_above(P);
  if (lane == 63) qp = pextra;
  R al, ar;
  recon<R>(qm, P, qp, rgx, al, ar);
  R p_l = lane_from_below(ar), p_r = al;
  R p_max = p_l > p_r ? p_l : p_r, p_min = p_l > p_r ? p_r : p_l;
  return ((p_max - p_min) / p_min > (R)0.5) ? 1 : 0;
}

// cell tag from face tags: zone l is tagged by faces l, l+1, l+2 (step.c:711-717)
MDS_DEV int zone_tag(int face_shock)
{
  int a1 = __shfl_down(face_shock, 1, 64);
  int a2 = __shfl_down(face_shock, 2, 64);
  return face_shock | a1 | a2;
}

}  // namespace

template <class R, int NV>
__global__ __launch_bounds__(256) void sweep2d(Sweep2DArgs a, StateArrays<R> sp, StateArrays<R> sph, int stage)
{
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (wv >= a.nstrip * a.nchunk) return;
  const DevProb &pb = *a.prob;
  DevClock *ck = a.clock;
  if (!ck->active) return;
  const int strip = wv % a.nstrip, chunk = wv / a.nstrip;
  const int n1 = pb.n[0], n2 = pb.n[1];
  const int c0 = strip * MDS_SEG1D;
  const int i = c0 - MDS_NG + lane;
  const int j0 = chunk * a.chunk;
  const int j1 = min(j0 + a.chunk, n2);
  const double dt = ck->dt;
  const R gam = (R)pb.gam;
  const double area0 = pb.area[0], area1 = pb.area[1];
  const StateArrays<R> &in = (stage == 0) ? sp : sph;
  const bool col_out = (lane >= 2) && (lane <= 61) && (i >= 0) && (i < n1);
  const bool xface_ok = (lane >= 2) && (lane <= 62) && (i >= 0) && (i <= n1);

  typename ReconGeoT<R>::type rgx;
  load_rgeo<R>(pb, 0, i, rgx);

  double lim = 1.0e300;
  unsigned long long nshock = 0, nhlle = 0;

  // window W[k] = row j-2+k for the row j being updated
  Cell<R, NV> W[5];
  for (int k = 0; k < 4; k++) load_cell<R, NV>(in, pb, i, j0 - 2 + k, W[k]);
  auto pextra_of = [&](int j) -> R {
    // P of zone i+1 in row j, needed by lane 63 only
    R pe = (R)1;
    if (lane == 63) { Cell<R, NV> e; load_cell<R, NV>(in, pb, i + 1, j, e); pe = e.P; }
    return pe;
  };
  auto stage1_tag = [&](int j) -> int {
    if (stage == 1 && col_out) return a.flags[(long long)j * n1 + i];
    return 0;
  };

  // tags of row j0-1 (ghost rows are always tagged)
  int tag_prev;
  if (j0 - 1 < 0) tag_prev = 1;
  else tag_prev = zone_tag(xshock_only<R>(W[1].P, pextra_of(j0 - 1), rgx)) | stage1_tag(j0 - 1);

  // x sweep of row j0
  double Fx[5], sigx;
  int shx;
  xsweep<R, NV>(W[2], pextra_of(j0), rgx, gam, area0, Fx, sigx, shx);
  if (xface_ok) lim = fmin(lim, pb.dx[0] / fabs(sigx));
  if (xface_ok && lane < 62 && j0 < n2) nshock += shx;
  int tag_cur = zone_tag(shx) | stage1_tag(j0);
  double Fxp_cur[5], Fxm_cur[5];
  for (int q = 0; q < 2 + NV; q++) { Fxp_cur[q] = lane_from_above(Fx[q]); Fxm_cur[q] = Fx[q]; }

  // y reconstruction of rows j0-1 and j0, flux through face j0-1/2
  Cell<R, NV> yl_prev, yr_prev, yl_cur, yr_cur;
  {
    typename ReconGeoT<R>::type rg;
    load_rgeo<R>(pb, 1, j0 - 1, rg);
    recon<R>(W[0].rho, W[1].rho, W[2].rho, rg, yl_prev.rho, yr_prev.rho);
    recon<R>(W[0].uu, W[1].uu, W[2].uu, rg, yl_prev.uu, yr_prev.uu);
    for (int d = 0; d < NV; d++) recon<R>(W[0].v[d], W[1].v[d], W[2].v[d], rg, yl_prev.v[d], yr_prev.v[d]);
    recon<R>(W[0].P, W[1].P, W[2].P, rg, yl_prev.P, yr_prev.P);
    load_rgeo<R>(pb, 1, j0, rg);
    recon<R>(W[1].rho, W[2].rho, W[3].rho, rg, yl_cur.rho, yr_cur.rho);
    recon<R>(W[1].uu, W[2].uu, W[3].uu, rg, yl_cur.uu, yr_cur.uu);
    for (int d = 0; d < NV; d++) recon<R>(W[1].v[d], W[2].v[d], W[3].v[d], rg, yl_cur.v[d], yr_cur.v[d]);
    recon<R>(W[1].P, W[2].P, W[3].P, rg, yl_cur.P, yr_cur.P);
  }
  double Fym[5], sigy;
  {
    bool hl = (tag_prev | tag_cur) != 0;
    solve<R, NV>(yr_prev, yl_cur, gam, area1, 1, hl, Fym, sigy);
    if (col_out) {
      lim = fmin(lim, pb.dx[1] / fabs(sigy));
      if (j0 == 0) nhlle += hl;   // face j0 is also the last face of the chunk below
    }
  }

  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  for (int j = j0; j < j1; j++) {
    load_cell<R, NV>(in, pb, i, j + 2, W[4]);
    // x sweep (or tags only) of row j+1
    int tag_next;
    double Fxp_next[5], Fxm_next[5];
    if (j + 1 >= n2) {
      tag_next = 1;
    } else if (j + 1 < j1) {
      double F2[5], s2;
      int sh2;
      xsweep<R, NV>(W[3], pextra_of(j + 1), rgx, gam, area0, F2, s2, sh2);
      if (xface_ok) lim = fmin(lim, pb.dx[0] / fabs(s2));
      if (xface_ok && lane < 62) nshock += sh2;
      tag_next = zone_tag(sh2) | stage1_tag(j + 1);
      for (int q = 0; q < 2 + NV; q++) { Fxp_next[q] = lane_from_above(F2[q]); Fxm_next[q] = F2[q]; }
    } else {
      tag_next = zone_tag(xshock_only<R>(W[3].P, pextra_of(j + 1), rgx)) | stage1_tag(j + 1);
    }
    // y reconstruction of row j+1 and flux through face j+1/2
    Cell<R, NV> yl_next, yr_next;
    {
      typename ReconGeoT<R>::type rg;
      load_rgeo<R>(pb, 1, j + 1, rg);
      recon<R>(W[2].rho, W[3].rho, W[4].rho, rg, yl_next.rho, yr_next.rho);
      recon<R>(W[2].uu, W[3].uu, W[4].uu, rg, yl_next.uu, yr_next.uu);
      for (int d = 0; d < NV; d++) recon<R>(W[2].v[d], W[3].v[d], W[4].v[d], rg, yl_next.v[d], yr_next.v[d]);
      recon<R>(W[2].P, W[3].P, W[4].P, rg, yl_next.P, yr_next.P);
    }
    double Fyp[5], sy;
    bool hl = (tag_cur | tag_next) != 0;
    solve<R, NV>(yr_cur, yl_next, gam, area1, 1, hl, Fyp, sy);
    if (col_out) {
      lim = fmin(lim, pb.dx[1] / fabs(sy));
      nhlle += hl;
    }

    // update zone (i, j)
    if (col_out) {
      const long long idx = pb.off + (long long)j * n1 + i;
      const Cell<R, NV> &me = W[2];
      double rho = me.rho, uu = me.uu, press = me.P, v[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < NV; d++) v[d] = me.v[d];
      double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (pb.g0[0] != 0.0 || pb.g0[1] != 0.0) {
        double vs = 0.0;
        for (int d = 0; d < 3; d++) vs += v[d] * pb.g0[d];
        src[1] += rho * vs;
        for (int d = 0; d < 2 && d < NV; d++) src[2 + d] += rho * pb.g0[d];
      }
      double rn = rho, un = uu, vn[3] = {v[0], v[1], v[2]};
      if (stage == 1) {
        rn = sp.v[0][idx];
        un = sp.v[1][idx];
        for (int d = 0; d < NV; d++) vn[d] = sp.v[2 + d][idx];
      }
      const double one3[3] = {1.0, 1.0, 1.0};
      double U[5], Uh[5];
      p_to_u_x64<NV, true>(rn, un, vn, one3, U);
      p_to_u_x64<NV, true>(rho, uu, v, one3, Uh);
      double dsrc = dt_source_control_x64<NV>(un, vn, U, src);
      lim = fmin(lim, dsrc);
      for (int q = 0; q < 2 + NV; q++) {
        double divF = ((((0.0 + Fxp_cur[q]) - Fxm_cur[q]) + Fyp[q]) - Fym[q]);
        double ddt = (divF / pb.volume) - src[q];
        U[q] = ((pf * U[q]) + (phf * Uh[q])) - ((ff * dt) * ddt);
      }
      double ro, uo, vo[3];
      u_to_p_x64<NV, true>(U, one3, ro, uo, vo);
      StateArrays<R> &out = (stage == 0) ? sph : sp;
      out.v[0][idx] = (R)ro;
      out.v[1][idx] = (R)uo;
      for (int d = 0; d < NV; d++) out.v[2 + d][idx] = (R)vo[d];
      if (stage == 0) a.flags[(long long)j * n1 + i] = (unsigned char)(tag_cur & 1);
      if (stage == 1) {
        double cs = sqrt(((double)pb.gam * press) / rho);
        double loc = 0.0;
        loc += (fabs(vo[0]) + cs) / pb.dx[0];
        loc += (fabs(NV > 1 ? vo[1] : 0.0) + cs) / pb.dx[1];
        lim = fmin(lim, 1.0 / loc);
      }
    }

    // slide the window
    for (int k = 0; k < 4; k++) W[k] = W[k + 1];
    yl_cur = yl_next; yr_cur = yr_next;
    tag_cur = tag_next;
    for (int q = 0; q < 2 + NV; q++) { Fym[q] = Fyp[q]; Fxp_cur[q] = Fxp_next[q]; Fxm_cur[q] = Fxm_next[q]; }
  }

  double wmin = wave_min(lim);
  if (lane == 0 && wmin < 1.0e300) atomicMin(&ck->mindt, dt_key(wmin));
  // counters: sum over lanes
  for (int off = 32; off > 0; off >>= 1) {
    nshock += __shfl_xor(nshock, off, 64);
    nhlle += __shfl_xor(nhlle, off, 64);
  }
  if (lane == 0) {
    if (nshock) atomicAdd(&a.counters[0], nshock);
    if (nhlle) atomicAdd(&a.counters[1], nhlle);
  }
}

extern "C" hipError_t mds_launch_sweep2d(int fp32, int nv, const Sweep2DArgs *a, const void *sp,
                                         const void *sph, int stage, hipStream_t st)
{
  const int wpb = 4;
  int waves = a->nstrip * a->nchunk;
  dim3 block(64 * wpb), grid((waves + wpb - 1) / wpb);
  if (!fp32) {
    const StateArrays<double> &p = *(const StateArrays<double> *)sp;
    const StateArrays<double> &h = *(const StateArrays<double> *)sph;
    if (nv == 2) hipLaunchKernelGGL((sweep2d<double, 2>), grid, block, 0, st, *a, p, h, stage);
    else hipLaunchKernelGGL((sweep2d<double, 3>), grid, block, 0, st, *a, p, h, stage);
  } else {
    const StateArrays<float> &p = *(const StateArrays<float> *)sp;
    const StateArrays<float> &h = *(const StateArrays<float> *)sph;
    if (nv == 2) hipLaunchKernelGGL((sweep2d<float, 2>), grid, block, 0, st, *a, p, h, stage);
    else hipLaunchKernelGGL((sweep2d<float, 3>), grid, block, 0, st, *a, p, h, stage);
  }
  return hipGetLastError();
}
