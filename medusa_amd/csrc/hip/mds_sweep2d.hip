// mds_sweep2d.hip -- fused RK stage for 2D Cartesian grids (gfx950).
//
// Decomposition: one 64-lane wavefront owns an x-strip of 64 zones
// [c0-2, c0+62) and marches in y over a chunk of rows [j0, j0+ROWS), producing
// the update of the 60 zones [c0, c0+60) of every row of the chunk.
//   x direction: each lane reconstructs its own zone; neighbours come from
//                whole-wave DPP shifts; the lane solves its lower x face.
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
//
// CFL limits: min over faces of dx/|s| equals dx/max|s|, min over zones of
// 1/loc equals 1/max(loc), and with no source terms the three
// dt_source_control quotients share the denominator 1e-16 (step.c:212-219),
// so each wavefront reduces first and divides once -- bit-identical to the
// per-element divisions of the reference.
#include "mds_sweep_common.h"

using namespace mds;

template <class R, int NV, bool SRC, bool SLAB>
__global__ __launch_bounds__(256) void sweep2d(Sweep2DArgs a, StateArrays<R> sp, StateArrays<R> sph, int stage)
{
  constexpr bool X64 = std::is_same<R, double>::value;
  const int lane = threadIdx.x & 63;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (wv >= a.nstrip * a.nchunk) return;
  const DevProb &pb = *a.prob;
  DevClock *ck = a.clock;
  if (!ck->active) return;
  const unsigned long long seen = min_key_seen(&ck->mindt);
  const int strip = wv % a.nstrip, chunk = wv / a.nstrip;
  const int n1 = pb.n[0], n2 = pb.n[1];
  const int c0 = strip * MDS_SEG1D;
  const int i = c0 - MDS_NG + lane;
  const int j0 = chunk * a.chunk;
  const int j1 = min(j0 + a.chunk, n2);
  const double dt = ck->dt;
  const R gam = (R)pb.gam;
  const R gm1 = (R)(pb.gam - 1.0);
  const R igm1 = (R)(1.0 / (pb.gam - 1.0));
  // the wave-speed factors of Gamma are formed per face here (one set per
  // launch measured 0.8% slower in FP32 on a 2048^2 Kelvin-Helmholtz box)
  const GamSet gk{};
  const double area0 = pb.area[0], area1 = pb.area[1];
  const StateArrays<R> &in = (stage == 0) ? sp : sph;
  const bool col_out = (lane >= 2) && (lane <= 61) && (i >= 0) && (i < n1);
  const bool xface_ok = (lane >= 2) && (lane <= 62) && (i >= 0) && (i <= n1);

  typename ReconGeoT<R>::type rgx;
  load_rgeo<R>(pb, 0, i, rgx);

  // running maxima of |signal speed| (x, y faces), of the zone bound
  // 1/dt_zone and (no sources) the minimum dt_source_control numerator
  R smax_x = (R)0, smax_y = (R)0;
  double locmax = 0.0, srcnum = 1.0e300, lim_src = 1.0e300;
  unsigned long long nshock = 0, nhlle = 0, nbad = 0;

  // The lane's column and the row mapping's inputs, read once: the problem
  // record sits in global memory, and after the march's stores the compiler
  // would read its fields again for every row.  The column mapping (x) is
  // fixed per lane, the row mapping (y) is uniform.  The general instance
  // (PROB/EXTRAP ghost layers) keeps load_cell.
  bool fxi, fxp;
  const int si = bc_src2(i, n1, pb.bc_lo[0], pb.bc_hi[0], fxi);
  const int sip = bc_src2(i + 1, n1, pb.bc_lo[0], pb.bc_hi[0], fxp);
  const int blo1 = pb.bc_lo[1], bhi1 = pb.bc_hi[1];
  const long long poff = pb.off;
  auto load_row = [&](int j, Cell<R, NV> &c) {
    if constexpr (SLAB) {
      load_cell<R, NV, SLAB>(in, pb, i, j, 0, gm1, c);
    } else {
      bool fy;
      const int sj = bc_src2(j, n2, blo1, bhi1, fy);
      const long long idx = poff + (long long)sj * n1 + si;
      c.rho = in.v[0][idx];
      c.uu = in.v[1][idx];
      for (int d = 0; d < NV; d++) c.v[d] = in.v[2 + d][idx];
      if (fxi) c.v[0] = -c.v[0];
      if (fy && NV > 1) c.v[1] = -c.v[1];
      c.P = gm1 * c.uu;
    }
  };
  Cell<R, NV> W[5];
  for (int k = 0; k < 4; k++) load_row(j0 - 2 + k, W[k]);
  auto pextra_of = [&](int j) -> R {
    R pe = (R)1;
    if constexpr (SLAB) {
      if (lane == 63) pe = load_press<R, SLAB>(in, pb, i + 1, j, 0, gm1);
    } else if (lane == 63) {
      bool fy;
      const int sj = bc_src2(j, n2, blo1, bhi1, fy);
      pe = gm1 * in.v[1][poff + (long long)sj * n1 + sip];
    }
    return pe;
  };
  auto stage1_tag = [&](int j) -> int {
    if (stage == 1 && col_out) return a.flags[(long long)j * n1 + i];
    return 0;
  };

  int tag_prev;
  if (j0 - 1 < 0) tag_prev = 1;
  else tag_prev = zone_tag(xshock_only<R, SLAB>(W[1].P, pextra_of(j0 - 1), rgx)) | stage1_tag(j0 - 1);

  R Fx[5], sx;
  int shx;
  xsweep<R, NV, SLAB, false>(W[2], pextra_of(j0), rgx, gam, gk, igm1, area0, Fx, sx, shx);
  if (xface_ok) smax_x = sx > smax_x ? sx : smax_x;
  if (xface_ok && lane < 62) nshock += shx;
  int tag_cur = zone_tag(shx) | stage1_tag(j0);
  R Fxp_cur[5], Fxm_cur[5];
  for (int q = 0; q < 2 + NV; q++) { Fxp_cur[q] = lane_from_above(Fx[q]); Fxm_cur[q] = Fx[q]; }

  Cell<R, NV> yl_cur, yr_cur;
  R Fym[5];
  {
    typename ReconGeoT<R>::type rg;
    Cell<R, NV> yl_prev, yr_prev;
    load_rgeo<R>(pb, 1, j0 - 1, rg);
    recon_cell<R, NV, SLAB>(W[0], W[1], W[2], rg, igm1, yl_prev, yr_prev);
    load_rgeo<R>(pb, 1, j0, rg);
    recon_cell<R, NV, SLAB>(W[1], W[2], W[3], rg, igm1, yl_cur, yr_cur);
    bool hl = (tag_prev | tag_cur) != 0;
    R sy;
    solve<R, NV, false>(yr_prev, yl_cur, gam, gk, area1, 1, hl, Fym, sy);
    if (col_out) {
      smax_y = sy > smax_y ? sy : smax_y;
      if (j0 == 0) nhlle += hl;   // face j0 is also the last face of the chunk below
    }
  }

  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  const double ffdt = ff * dt;
  const double idx0 = 1.0 / pb.dx[0], idx1 = 1.0 / pb.dx[1];
  for (int j = j0; j < j1; j++) {
    load_row(j + 2, W[4]);
    int tag_next;
    R Fxp_next[5], Fxm_next[5];
    if (j + 1 >= n2) {
      tag_next = 1;
    } else if (j + 1 < j1) {
      R F2[5], s2;
      int sh2;
      xsweep<R, NV, SLAB, false>(W[3], pextra_of(j + 1), rgx, gam, gk, igm1, area0, F2, s2, sh2);
      if (xface_ok) smax_x = s2 > smax_x ? s2 : smax_x;
      if (xface_ok && lane < 62) nshock += sh2;
      tag_next = zone_tag(sh2) | stage1_tag(j + 1);
      for (int q = 0; q < 2 + NV; q++) { Fxp_next[q] = lane_from_above(F2[q]); Fxm_next[q] = F2[q]; }
    } else {
      tag_next = zone_tag(xshock_only<R, SLAB>(W[3].P, pextra_of(j + 1), rgx)) | stage1_tag(j + 1);
    }
    Cell<R, NV> yl_next, yr_next;
    {
      typename ReconGeoT<R>::type rg;
      load_rgeo<R>(pb, 1, j + 1, rg);
      recon_cell<R, NV, SLAB>(W[2], W[3], W[4], rg, igm1, yl_next, yr_next);
    }
    R Fyp[5], sy;
    bool hl = (tag_cur | tag_next) != 0;
    solve<R, NV, false>(yr_cur, yl_next, gam, gk, area1, 1, hl, Fyp, sy);
    if (col_out) {
      smax_y = sy > smax_y ? sy : smax_y;
      nhlle += hl;
    }

    bool bad_live = false;
    double bad_ro = 1.0, bad_uo = 0.0;
    if (col_out) {
      const long long idx = pb.off + (long long)j * n1 + i;
      const Cell<R, NV> &me = W[2];
      double rho = me.rho, uu = me.uu, press = me.P, v[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < NV; d++) v[d] = me.v[d];
      double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (SRC) {
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
      if (SRC) {
        double dsrc = dt_source_control_x64<NV>(un, vn, U, src);
        lim_src = (lim_src < dsrc) ? lim_src : dsrc;
      } else {
        // (0.02*u0, 0.02*u1, 2*uu) over the common denominator 1e-16
        double m0 = 0.02 * U[0], m1 = 0.02 * U[1], m2 = 2.0 * un;
        double m = (m0 < m1) ? m0 : m1;
        m = (m < m2) ? m : m2;
        srcnum = (srcnum < m) ? srcnum : m;
      }
      double ro, uo, vo[3];
      if constexpr (X64) {
        for (int q = 0; q < 2 + NV; q++) {
          double divF = ((((0.0 + Fxp_cur[q]) - Fxm_cur[q]) + Fyp[q]) - Fym[q]);
          double ddt = (divF / pb.volume) - src[q];
          U[q] = ((pf * U[q]) + (phf * Uh[q])) - ((ff * dt) * ddt);
        }
        u_to_p_x64<NV, true>(U, one3, ro, uo, vo);
      } else {
        for (int q = 0; q < 2 + NV; q++) {
          double divF = ((double)Fxp_cur[q] - (double)Fxm_cur[q]) * idx0 + ((double)Fyp[q] - (double)Fym[q]) * idx1;
          U[q] = pf * U[q] + phf * Uh[q] - ffdt * (divF - src[q]);
        }
        ro = U[0];
        double ri = drcp(U[0]);
        double vsq = 0.0;
        for (int d = 0; d < NV; d++) { vo[d] = U[2 + d] * ri; vsq += U[2 + d] * vo[d]; }
        uo = U[1] - 0.5 * vsq;
      }
      bad_ro = ro;
      bad_uo = uo;
      bad_live = true;
      StateArrays<R> &out = (stage == 0) ? sph : sp;
      out.v[0][idx] = (R)ro;
      out.v[1][idx] = (R)uo;
      for (int d = 0; d < NV; d++) out.v[2 + d][idx] = (R)vo[d];
      if (stage == 0) a.flags[(long long)j * n1 + i] = (unsigned char)(tag_cur & 1);
      if (stage == 1) {
        double loc;
        if constexpr (X64) {
          double cs = sqrt(((double)pb.gam * press) / rho);
          loc = 0.0;
          loc += (fabs(vo[0]) + cs) / pb.dx[0];
          loc += (fabs(NV > 1 ? vo[1] : 0.0) + cs) / pb.dx[1];
        } else {
          float cs = fsqrt((float)pb.gam * (float)press * frcp((float)rho));
          loc = (fabs(vo[0]) + cs) * idx0 + (fabs(NV > 1 ? vo[1] : 0.0) + cs) * idx1;
        }
        locmax = loc > locmax ? loc : locmax;
      }
    }

    nbad += bad_state_lanes(bad_live, bad_ro, bad_uo, true);
    for (int k = 0; k < 4; k++) W[k] = W[k + 1];
    yl_cur = yl_next; yr_cur = yr_next;
    tag_cur = tag_next;
    for (int q = 0; q < 2 + NV; q++) { Fym[q] = Fyp[q]; Fxp_cur[q] = Fxp_next[q]; Fxm_cur[q] = Fxm_next[q]; }
  }

  // one division per limit per wavefront
  double wsx = wave_max((double)smax_x), wsy = wave_max((double)smax_y), wloc = wave_max(locmax);
  double wsrc = wave_min(srcnum), wlim = wave_min(lim_src);
  if (lane == 0) {
    double lim = 1.0e300;
    if (wsx > 0.0) lim = fmin(lim, pb.dx[0] / wsx);
    if (wsy > 0.0) lim = fmin(lim, pb.dx[1] / wsy);
    if (wloc > 0.0) lim = fmin(lim, 1.0 / wloc);
    if (wsrc < 1.0e300) lim = fmin(lim, wsrc / (0.0 + 1.0e-16));
    lim = fmin(lim, wlim);
    if (lim < 1.0e300) fold_min_key(&ck->mindt, seen, lim);
  }
  for (int off = 32; off > 0; off >>= 1) {
    nshock += __shfl_xor(nshock, off, 64);
    nhlle += __shfl_xor(nhlle, off, 64);
  }
  if (lane == 0) {
    if (nshock) atomicAdd(&a.counters[0], nshock);
    if (nhlle) atomicAdd(&a.counters[1], nhlle);
    if (nbad) atomicAdd(&ck->nbad, nbad);
  }
}

extern "C" hipError_t mds_launch_sweep2d(int fp32, int nv, const Sweep2DArgs *a, const void *sp,
                                         const void *sph, int stage, hipStream_t st)
{
  const int wpb = 4;
  int waves = a->nstrip * a->nchunk;
  dim3 block(64 * wpb), grid((waves + wpb - 1) / wpb);
  bool src = a->src;
#define MDS_L2(R, NV, SRC)                                                                                \
  do {                                                                                                    \
    if (a->slab)                                                                                          \
      hipLaunchKernelGGL((sweep2d<R, NV, SRC, true>), grid, block, 0, st, *a, *(const StateArrays<R> *)sp, \
                         *(const StateArrays<R> *)sph, stage);                                            \
    else                                                                                                  \
      hipLaunchKernelGGL((sweep2d<R, NV, SRC, false>), grid, block, 0, st, *a, *(const StateArrays<R> *)sp, \
                         *(const StateArrays<R> *)sph, stage);                                            \
  } while (0)
  if (!fp32) {
    if (nv == 2) { if (src) MDS_L2(double, 2, true); else MDS_L2(double, 2, false); }
    else { if (src) MDS_L2(double, 3, true); else MDS_L2(double, 3, false); }
  } else {
    if (nv == 2) { if (src) MDS_L2(float, 2, true); else MDS_L2(float, 2, false); }
    else { if (src) MDS_L2(float, 3, true); else MDS_L2(float, 3, false); }
  }
#undef MDS_L2
  return hipGetLastError();
}
