// mds_sweep3d.hip -- fused RK stage for 3D Cartesian grids (gfx950).
//
// Decomposition: a workgroup of TY+4 wavefronts owns an x-strip of 64 zones
// [c0-2, c0+62) of the rows [j0-2, j0+TY+2) and marches in z over a chunk of
// planes [k0, k0+KC), producing the update of the 60 x TY zones
// [c0, c0+60) x [j0, j0+TY) of every plane of the chunk.
//   x: each lane reconstructs its own zone from DPP neighbours and solves
//      its lower x face (as in the 2D sweep);
//   z: the wavefronts of the output rows keep a 5-plane register window and
//      reconstruct / solve each zone and z face once;
//   y: per plane the zone records of all TY+4 rows are exchanged through LDS;
//      rows j0-1 .. j0+TY reconstruct in y, rows j0 .. j0+TY solve their
//      lower y face, and the output rows difference the y fluxes.
// Shock tags follow the two-pass reading of step.c:1196-1418: radial (x)
// faces tag zones i, i-1, i-2; a y face uses HLLE if zone (i,j,k) or
// (i,j-1,k) is tagged, a z face if (i,j,k) or (i,j,k-1) is; ghost zones are
// always tagged (never reset from malloc_rank1's 0, utils.c:171).
// Replaces prim_to_fluxes3 (step.c:1088-1430) on uniform grids, push
// (step.c:327-478), reset_boundaries_gpu (bc.c:705-1421), update_eos /
// update_eos_ghost (eos.c:4-32) and fixed_gravity (gravity.c:46-64).
// The polar-axis fix of step.c:1301-1401 belongs to spherical grids only.
#include "mds_sweep_common.h"

using namespace mds;

template <class R, int NV, bool SRC, int TY, bool SLAB = false>
__global__ __launch_bounds__((TY + 4) * 64) void sweep3d(Sweep3DArgs a, StateArrays<R> sp, StateArrays<R> sph,
                                                          int stage)
{
  constexpr bool X64 = std::is_same<R, double>::value;
  constexpr int NW = TY + 4;
  __shared__ R sCell[3 + NV][NW][64];
  __shared__ R sAr[3 + NV][NW][64];
  __shared__ R sFy[2 + NV][NW][64];
  __shared__ int sTag[NW][64];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const DevProb &pb = *a.prob;
  DevClock *ck = a.clock;
  if (!ck->active) return;                       // uniform over the launch
  const unsigned long long seen = min_key_seen(&ck->mindt);
  const int strip = blockIdx.x % a.nstrip;
  const int tile = (blockIdx.x / a.nstrip) % a.ntile;
  const int chunk = blockIdx.x / (a.nstrip * a.ntile);
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const int i = strip * MDS_SEG1D - MDS_NG + lane;
  const int j0 = tile * TY;
  const int j = j0 - 2 + w;
  const int k0 = chunk * a.chunk;
  const int k1 = min(k0 + a.chunk, n3);
  const double dt = ck->dt;
  const R gam = (R)pb.gam;
  const R gm1 = (R)(pb.gam - 1.0);
  const R igm1 = (R)(1.0 / (pb.gam - 1.0));
  const GamSet gk = gam_set_uniform(pb.gam);   // the gamma law's wave-speed factors
  const StateArrays<R> &in = (stage == 0) ? sp : sph;
  const bool zw = (w >= 2) && (w <= TY + 1);             // rows with a z window
  const bool row_in = (j >= 0) && (j < n2);
  const bool lane_out = (lane >= 2) && (lane <= 61) && (i >= 0) && (i < n1);
  const bool col_out = lane_out && zw && row_in;
  const bool xface_ok = (lane >= 2) && (lane <= 62) && (i >= 0) && (i <= n1) && zw && row_in;
  const bool yface = (w >= 2) && (w <= TY + 2) && (j <= n2) && lane_out;   // lower y face of row j
  const bool ycount = yface && (w <= TY + 1 || j == n2);

  typename ReconGeoT<R>::type rgx, rgy;
  load_rgeo<R>(pb, 0, i, rgx);
  load_rgeo<R>(pb, 1, j, rgy);

  R smax[3] = {(R)0, (R)0, (R)0};
  double locmax = 0.0, srcnum = 1.0e300, lim_src = 1.0e300;
  unsigned long long nshock = 0, nhlle = 0, nbad = 0;

  auto s1tag = [&](int kk) -> int {
    if (stage == 1 && lane_out && row_in && kk >= 0 && kk < n3)
      return a.flags[((long long)kk * n2 + j) * n1 + i];
    return 0;
  };
  auto pextra = [&](int kk) -> R {
    R pe = (R)1;
    if (lane == 63) pe = load_press<R, SLAB>(in, pb, i + 1, j, kk, gm1);
    return pe;
  };

  Cell<R, NV> Z[5];                 // zw rows: planes k-2 .. k+2
  R Fxm_cur[5], Fzm[5];            // x flux at the lower face; the upper face's comes from the lane above when used
  Cell<R, NV> zl_cur, zr_cur;
  int tag_cur = 1;
  if (zw) {
    for (int q = 0; q < 4; q++) load_cell<R, NV, SLAB>(in, pb, i, j, k0 - 2 + q, gm1, Z[q]);
    int tag_prev;
    if (!row_in || k0 - 1 < 0) tag_prev = 1;
    else tag_prev = zone_tag(xshock_only<R, SLAB>(Z[1].P, pextra(k0 - 1), rgx)) | s1tag(k0 - 1);
    R Fx[5], sx;
    int shx;
    xsweep<R, NV, SLAB>(Z[2], pextra(k0), rgx, gam, gk, igm1, pb.area[0], Fx, sx, shx);
    if (xface_ok) smax[0] = sx > smax[0] ? sx : smax[0];
    if (xface_ok && lane < 62) nshock += shx;
    tag_cur = row_in ? (zone_tag(shx) | s1tag(k0)) : 1;
    for (int q = 0; q < 2 + NV; q++) Fxm_cur[q] = Fx[q];
    typename ReconGeoT<R>::type rg;
    Cell<R, NV> zl_prev, zr_prev;
    load_rgeo<R>(pb, 2, k0 - 1, rg);
    recon_cell<R, NV, SLAB>(Z[0], Z[1], Z[2], rg, igm1, zl_prev, zr_prev);
    load_rgeo<R>(pb, 2, k0, rg);
    recon_cell<R, NV, SLAB>(Z[1], Z[2], Z[3], rg, igm1, zl_cur, zr_cur);
    bool hl = (tag_prev | tag_cur) != 0;
    R sz;
    solve<R, NV>(zr_prev, zl_cur, gam, gk, pb.area[2], 2, hl, Fzm, sz);
    if (col_out) {
      smax[2] = sz > smax[2] ? sz : smax[2];
      if (k0 == 0) nhlle += hl;
    }
  }

  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  const double ffdt = ff * dt;
  const double idx[3] = {1.0 / pb.dx[0], 1.0 / pb.dx[1], 1.0 / pb.dx[2]};

  for (int k = k0; k < k1; k++) {
    Cell<R, NV> me;
    R Fxm_next[5], Fzp[5];
    Cell<R, NV> zl_next, zr_next;
    int tag_next = 1, mytag = 1;
    if (zw) {
      load_cell<R, NV, SLAB>(in, pb, i, j, k + 2, gm1, Z[4]);
      if (!row_in || k + 1 >= n3) {
        tag_next = 1;
      } else if (k + 1 < k1) {
        R F2[5], s2;
        int sh2;
        xsweep<R, NV, SLAB>(Z[3], pextra(k + 1), rgx, gam, gk, igm1, pb.area[0], F2, s2, sh2);
        if (xface_ok) smax[0] = s2 > smax[0] ? s2 : smax[0];
        if (xface_ok && lane < 62) nshock += sh2;
        tag_next = zone_tag(sh2) | s1tag(k + 1);
        for (int q = 0; q < 2 + NV; q++) Fxm_next[q] = F2[q];
      } else {
        tag_next = zone_tag(xshock_only<R, SLAB>(Z[3].P, pextra(k + 1), rgx)) | s1tag(k + 1);
      }
      typename ReconGeoT<R>::type rg;
      load_rgeo<R>(pb, 2, k + 1, rg);
      recon_cell<R, NV, SLAB>(Z[2], Z[3], Z[4], rg, igm1, zl_next, zr_next);
      bool hl = (tag_cur | tag_next) != 0;
      R sz;
      solve<R, NV>(zr_cur, zl_next, gam, gk, pb.area[2], 2, hl, Fzp, sz);
      if (col_out) {
        smax[2] = sz > smax[2] ? sz : smax[2];
        nhlle += hl;
      }
      me = Z[2];
      mytag = tag_cur;
    } else {
      load_cell<R, NV, SLAB>(in, pb, i, j, k, gm1, me);
      if ((w == 1 || w == TY + 2) && row_in)
        mytag = zone_tag(xshock_only<R, SLAB>(me.P, pextra(k), rgx)) | s1tag(k);
      else
        mytag = 1;
    }

    // ---- y direction through LDS ----
    sCell[0][w][lane] = me.rho;
    sCell[1][w][lane] = me.uu;
    for (int d = 0; d < NV; d++) sCell[2 + d][w][lane] = me.v[d];
    sCell[2 + NV][w][lane] = me.P;
    sTag[w][lane] = mytag;
    __syncthreads();
    Cell<R, NV> yl;
    if (w >= 1 && w <= TY + 2) {
      Cell<R, NV> ym, yp, yr;
      ym.rho = sCell[0][w - 1][lane]; yp.rho = sCell[0][w + 1][lane];
      ym.uu = sCell[1][w - 1][lane];  yp.uu = sCell[1][w + 1][lane];
      for (int d = 0; d < NV; d++) { ym.v[d] = sCell[2 + d][w - 1][lane]; yp.v[d] = sCell[2 + d][w + 1][lane]; }
      ym.P = sCell[2 + NV][w - 1][lane]; yp.P = sCell[2 + NV][w + 1][lane];
      recon_cell<R, NV, SLAB>(ym, me, yp, rgy, igm1, yl, yr);
      sAr[0][w][lane] = yr.rho;
      sAr[1][w][lane] = yr.uu;
      for (int d = 0; d < NV; d++) sAr[2 + d][w][lane] = yr.v[d];
      sAr[2 + NV][w][lane] = yr.P;
    }
    __syncthreads();
    if (w >= 2 && w <= TY + 2) {
      Cell<R, NV> yL;
      yL.rho = sAr[0][w - 1][lane];
      yL.uu = sAr[1][w - 1][lane];
      for (int d = 0; d < NV; d++) yL.v[d] = sAr[2 + d][w - 1][lane];
      yL.P = sAr[2 + NV][w - 1][lane];
      bool hl = (sTag[w][lane] | sTag[w - 1][lane]) != 0;
      R Fy[5], sy;
      solve<R, NV>(yL, yl, gam, gk, pb.area[1], 1, hl, Fy, sy);
      for (int q = 0; q < 2 + NV; q++) sFy[q][w][lane] = Fy[q];
      if (yface) smax[1] = sy > smax[1] ? sy : smax[1];
      if (ycount) nhlle += hl;
    }
    __syncthreads();

    // the upper x face of this zone is the lower face of the lane above
    // (a DPP shift, taken here where every lane of the wavefront is active)
    R Fxp_cur[5];
    for (int q = 0; q < 2 + NV; q++) Fxp_cur[q] = lane_from_above(Fxm_cur[q]);
    bool bad_live = false;
    double bad_ro = 1.0, bad_uo = 0.0;
    if (col_out) {
      const long long idz = pb.off + ((long long)k * n2 + j) * n1 + i;
      double rho = me.rho, uu = me.uu, press = me.P, v[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < NV; d++) v[d] = me.v[d];
      double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (SRC) {
        double vs = 0.0;
        for (int d = 0; d < 3; d++) vs += v[d] * pb.g0[d];
        src[1] += rho * vs;
        for (int d = 0; d < NV; d++) src[2 + d] += rho * pb.g0[d];
      }
      double rn = rho, un = uu, vn[3] = {v[0], v[1], v[2]};
      if (stage == 1) {
        rn = sp.v[0][idz];
        un = sp.v[1][idz];
        for (int d = 0; d < NV; d++) vn[d] = sp.v[2 + d][idz];
      }
      const double one3[3] = {1.0, 1.0, 1.0};
      double U[5], Uh[5];
      p_to_u_x64<NV, true>(rn, un, vn, one3, U);
      p_to_u_x64<NV, true>(rho, uu, v, one3, Uh);
      if (SRC) {
        double dsrc = X64 ? dt_source_control_x64<NV>(un, vn, U, src) : dt_source_control_fast<NV>(un, vn, U, src);
        lim_src = (lim_src < dsrc) ? lim_src : dsrc;
      } else {
        double m0 = 0.02 * U[0], m1 = 0.02 * U[1], m2 = 2.0 * un;
        double m = (m0 < m1) ? m0 : m1;
        m = (m < m2) ? m : m2;
        srcnum = (srcnum < m) ? srcnum : m;
      }
      double ro, uo, vo[3];
      if constexpr (X64) {
        for (int q = 0; q < 2 + NV; q++) {
          double divF = ((((0.0 + Fxp_cur[q]) - Fxm_cur[q]) + sFy[q][w + 1][lane]) - sFy[q][w][lane]);
          divF += Fzp[q] - Fzm[q];
          double ddt = (divF / pb.volume) - src[q];
          U[q] = ((pf * U[q]) + (phf * Uh[q])) - ((ff * dt) * ddt);
        }
        u_to_p_x64<NV, true>(U, one3, ro, uo, vo);
      } else {
        for (int q = 0; q < 2 + NV; q++) {
          double divF = ((double)Fxp_cur[q] - (double)Fxm_cur[q]) * idx[0]
                        + ((double)sFy[q][w + 1][lane] - (double)sFy[q][w][lane]) * idx[1]
                        + ((double)Fzp[q] - (double)Fzm[q]) * idx[2];
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
      out.v[0][idz] = (R)ro;
      out.v[1][idz] = (R)uo;
      for (int d = 0; d < NV; d++) out.v[2 + d][idz] = (R)vo[d];
      if (stage == 0) a.flags[idz - pb.off] = (unsigned char)(tag_cur & 1);
      if (stage == 1) {
        double loc;
        if constexpr (X64) {
          double cs = sqrt(((double)pb.gam * press) / rho);
          loc = 0.0;
          for (int d = 0; d < 3; d++) loc += (fabs(d < NV ? vo[d] : 0.0) + cs) / pb.dx[d];
        } else {
          float cs = fsqrt((float)pb.gam * (float)press * frcp((float)rho));
          loc = 0.0;
          for (int d = 0; d < 3; d++) loc += (fabs(d < NV ? vo[d] : 0.0) + cs) * idx[d];
        }
        locmax = loc > locmax ? loc : locmax;
      }
    }

    nbad += bad_state_lanes(bad_live, bad_ro, bad_uo, true);
    if (zw) {
      for (int q = 0; q < 4; q++) Z[q] = Z[q + 1];
      zl_cur = zl_next; zr_cur = zr_next;
      tag_cur = tag_next;
      for (int q = 0; q < 2 + NV; q++) { Fzm[q] = Fzp[q]; Fxm_cur[q] = Fxm_next[q]; }
    }
  }

  double ws[3], wloc = wave_max(locmax);
  for (int d = 0; d < 3; d++) ws[d] = wave_max((double)smax[d]);
  double wsrc = wave_min(srcnum), wlim = wave_min(lim_src);
  if (lane == 0) {
    double lim = 1.0e300;
    for (int d = 0; d < 3; d++)
      if (ws[d] > 0.0) lim = fmin(lim, pb.dx[d] / ws[d]);
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

// The same stage with TY wavefronts, one per output row: the four halo rows
// are loaded by the first two and the last two wavefronts beside their own
// rows, and the first and last wavefronts also reconstruct the two halo rows
// that have faces (rows j0-1 and j0+TY) and solve the top face.  No
// wavefront of the workgroup is idle in the x and z sweeps.
template <class R, int NV, bool SRC, int TY, bool SLAB = false>
__global__ __launch_bounds__(TY * 64) void sweep3dh(Sweep3DArgs a, StateArrays<R> sp, StateArrays<R> sph,
                                                    int stage)
{
  constexpr bool X64 = std::is_same<R, double>::value;
  constexpr int NW = TY + 4;
  __shared__ R sCell[3 + NV][NW][64];
  __shared__ R sAr[3 + NV][NW][64];
  __shared__ R sFy[2 + NV][NW][64];
  __shared__ int sTag[NW][64];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const DevProb &pb = *a.prob;
  DevClock *ck = a.clock;
  if (!ck->active) return;                       // uniform over the launch
  const unsigned long long seen = min_key_seen(&ck->mindt);
  const int strip = blockIdx.x % a.nstrip;
  const int tile = (blockIdx.x / a.nstrip) % a.ntile;
  const int chunk = blockIdx.x / (a.nstrip * a.ntile);
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const int i = strip * MDS_SEG1D - MDS_NG + lane;
  const int j0 = tile * TY;
  const int r = w + 2;                                   // this wavefront's LDS row
  const int j = j0 + w;
  // halo row of this wavefront (LDS row h, grid row jh), or none
  const int h = w < 2 ? w : (w >= TY - 2 ? w + 4 : -1);
  const int jh = j0 - 2 + h;
  const bool top = w == TY - 1;                          // also solves face j0 + TY
  const int k0 = chunk * a.chunk;
  const int k1 = min(k0 + a.chunk, n3);
  const double dt = ck->dt;
  const R gam = (R)pb.gam;
  const R gm1 = (R)(pb.gam - 1.0);
  const R igm1 = (R)(1.0 / (pb.gam - 1.0));
  const GamSet gk = gam_set_uniform(pb.gam);   // the gamma law's wave-speed factors
  const StateArrays<R> &in = (stage == 0) ? sp : sph;
  const bool zw = true;                                  // every wavefront has a z window
  const bool row_in = j < n2;
  const bool halo_in = (jh >= 0) && (jh < n2);
  const bool lane_out = (lane >= 2) && (lane <= 61) && (i >= 0) && (i < n1);
  const bool col_out = lane_out && row_in;
  const bool xface_ok = (lane >= 2) && (lane <= 62) && (i >= 0) && (i <= n1) && row_in;
  const bool yface = (j <= n2) && lane_out;              // lower y face of row j
  const bool ycount = yface;
  const bool ytop = top && (j0 + TY <= n2) && lane_out;  // the top face j0 + TY
  const bool ytopcount = ytop && (j0 + TY == n2);

  typename ReconGeoT<R>::type rgx, rgy, rgh;
  load_rgeo<R>(pb, 0, i, rgx);
  load_rgeo<R>(pb, 1, j, rgy);
  if (w == 0) load_rgeo<R>(pb, 1, j0 - 1, rgh);
  if (top) load_rgeo<R>(pb, 1, j0 + TY, rgh);

  R smax[3] = {(R)0, (R)0, (R)0};
  double locmax = 0.0, srcnum = 1.0e300, lim_src = 1.0e300;
  unsigned long long nshock = 0, nhlle = 0, nbad = 0;

  auto s1tag = [&](int kk) -> int {
    if (stage == 1 && lane_out && row_in && kk >= 0 && kk < n3)
      return a.flags[((long long)kk * n2 + j) * n1 + i];
    return 0;
  };
  auto pextra = [&](int kk) -> R {
    R pe = (R)1;
    if (lane == 63) pe = load_press<R, SLAB>(in, pb, i + 1, j, kk, gm1);
    return pe;
  };
  auto s1tag_h = [&](int kk) -> int {
    if (stage == 1 && lane_out && halo_in && kk >= 0 && kk < n3)
      return a.flags[((long long)kk * n2 + jh) * n1 + i];
    return 0;
  };
  auto pextra_h = [&](int kk) -> R {
    R pe = (R)1;
    if (lane == 63) pe = load_press<R, SLAB>(in, pb, i + 1, jh, kk, gm1);
    return pe;
  };

  Cell<R, NV> Z[5];                 // zw rows: planes k-2 .. k+2
  R Fxm_cur[5], Fzm[5];            // x flux at the lower face; the upper face's comes from the lane above when used
  Cell<R, NV> zl_cur, zr_cur;
  int tag_cur = 1;
  if (zw) {
    for (int q = 0; q < 4; q++) load_cell<R, NV, SLAB>(in, pb, i, j, k0 - 2 + q, gm1, Z[q]);
    int tag_prev;
    if (!row_in || k0 - 1 < 0) tag_prev = 1;
    else tag_prev = zone_tag(xshock_only<R, SLAB>(Z[1].P, pextra(k0 - 1), rgx)) | s1tag(k0 - 1);
    R Fx[5], sx;
    int shx;
    xsweep<R, NV, SLAB>(Z[2], pextra(k0), rgx, gam, gk, igm1, pb.area[0], Fx, sx, shx);
    if (xface_ok) smax[0] = sx > smax[0] ? sx : smax[0];
    if (xface_ok && lane < 62) nshock += shx;
    tag_cur = row_in ? (zone_tag(shx) | s1tag(k0)) : 1;
    for (int q = 0; q < 2 + NV; q++) Fxm_cur[q] = Fx[q];
    typename ReconGeoT<R>::type rg;
    Cell<R, NV> zl_prev, zr_prev;
    load_rgeo<R>(pb, 2, k0 - 1, rg);
    recon_cell<R, NV, SLAB>(Z[0], Z[1], Z[2], rg, igm1, zl_prev, zr_prev);
    load_rgeo<R>(pb, 2, k0, rg);
    recon_cell<R, NV, SLAB>(Z[1], Z[2], Z[3], rg, igm1, zl_cur, zr_cur);
    bool hl = (tag_prev | tag_cur) != 0;
    R sz;
    solve<R, NV>(zr_prev, zl_cur, gam, gk, pb.area[2], 2, hl, Fzm, sz);
    if (col_out) {
      smax[2] = sz > smax[2] ? sz : smax[2];
      if (k0 == 0) nhlle += hl;
    }
  }

  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  const double ffdt = ff * dt;
  const double idx[3] = {1.0 / pb.dx[0], 1.0 / pb.dx[1], 1.0 / pb.dx[2]};

  for (int k = k0; k < k1; k++) {
    Cell<R, NV> me;
    R Fxm_next[5], Fzp[5];
    Cell<R, NV> zl_next, zr_next;
    int tag_next = 1, mytag = 1;
    if (zw) {
      load_cell<R, NV, SLAB>(in, pb, i, j, k + 2, gm1, Z[4]);
      if (!row_in || k + 1 >= n3) {
        tag_next = 1;
      } else if (k + 1 < k1) {
        R F2[5], s2;
        int sh2;
        xsweep<R, NV, SLAB>(Z[3], pextra(k + 1), rgx, gam, gk, igm1, pb.area[0], F2, s2, sh2);
        if (xface_ok) smax[0] = s2 > smax[0] ? s2 : smax[0];
        if (xface_ok && lane < 62) nshock += sh2;
        tag_next = zone_tag(sh2) | s1tag(k + 1);
        for (int q = 0; q < 2 + NV; q++) Fxm_next[q] = F2[q];
      } else {
        tag_next = zone_tag(xshock_only<R, SLAB>(Z[3].P, pextra(k + 1), rgx)) | s1tag(k + 1);
      }
      typename ReconGeoT<R>::type rg;
      load_rgeo<R>(pb, 2, k + 1, rg);
      recon_cell<R, NV, SLAB>(Z[2], Z[3], Z[4], rg, igm1, zl_next, zr_next);
      bool hl = (tag_cur | tag_next) != 0;
      R sz;
      solve<R, NV>(zr_cur, zl_next, gam, gk, pb.area[2], 2, hl, Fzp, sz);
      if (col_out) {
        smax[2] = sz > smax[2] ? sz : smax[2];
        nhlle += hl;
      }
      me = Z[2];
      mytag = tag_cur;
    }
    // this wavefront's halo row (tags only for the two rows that have faces)
    Cell<R, NV> hc;
    int htag = 1;
    if (h >= 0) {
      load_cell<R, NV, SLAB>(in, pb, i, jh, k, gm1, hc);
      if ((h == 1 || h == TY + 2) && halo_in) htag = zone_tag(xshock_only<R, SLAB>(hc.P, pextra_h(k), rgx)) | s1tag_h(k);
    }

    // ---- y direction through LDS ----
    auto put = [&](R (*dst)[NW][64], int row, const Cell<R, NV> &c) {
      dst[0][row][lane] = c.rho;
      dst[1][row][lane] = c.uu;
      for (int d = 0; d < NV; d++) dst[2 + d][row][lane] = c.v[d];
      dst[2 + NV][row][lane] = c.P;
    };
    auto get = [&](R (*src)[NW][64], int row) -> Cell<R, NV> {
      Cell<R, NV> c;
      c.rho = src[0][row][lane];
      c.uu = src[1][row][lane];
      for (int d = 0; d < NV; d++) c.v[d] = src[2 + d][row][lane];
      c.P = src[2 + NV][row][lane];
      return c;
    };
    put(sCell, r, me);
    sTag[r][lane] = mytag;
    if (h >= 0) {
      put(sCell, h, hc);
      sTag[h][lane] = htag;
    }
    __syncthreads();
    Cell<R, NV> yl, yr, ylh;
    recon_cell<R, NV, SLAB>(get(sCell, r - 1), me, get(sCell, r + 1), rgy, igm1, yl, yr);
    put(sAr, r, yr);
    if (w == 0) {
      // row j0-1: its right state for face j0
      Cell<R, NV> a0, a1;
      recon_cell<R, NV, SLAB>(get(sCell, 0), get(sCell, 1), get(sCell, 2), rgh, igm1, a0, a1);
      put(sAr, 1, a1);
    }
    if (top) {
      // row j0+TY: its left state for face j0+TY
      Cell<R, NV> a1;
      recon_cell<R, NV, SLAB>(get(sCell, TY + 1), get(sCell, TY + 2), get(sCell, TY + 3), rgh, igm1, ylh, a1);
    }
    __syncthreads();
    {
      const bool hl = (sTag[r][lane] | sTag[r - 1][lane]) != 0;
      R Fy[5], sy;
      solve<R, NV>(get(sAr, r - 1), yl, gam, gk, pb.area[1], 1, hl, Fy, sy);
      for (int q = 0; q < 2 + NV; q++) sFy[q][r][lane] = Fy[q];
      if (yface) smax[1] = sy > smax[1] ? sy : smax[1];
      if (ycount) nhlle += hl;
    }
    if (top) {
      const bool hl = (sTag[TY + 2][lane] | sTag[TY + 1][lane]) != 0;
      R Fy[5], sy;
      solve<R, NV>(yr, ylh, gam, gk, pb.area[1], 1, hl, Fy, sy);
      for (int q = 0; q < 2 + NV; q++) sFy[q][TY + 2][lane] = Fy[q];
      if (ytop) smax[1] = sy > smax[1] ? sy : smax[1];
      if (ytopcount) nhlle += hl;
    }
    __syncthreads();

    // the upper x face of this zone is the lower face of the lane above
    // (a DPP shift, taken here where every lane of the wavefront is active)
    R Fxp_cur[5];
    for (int q = 0; q < 2 + NV; q++) Fxp_cur[q] = lane_from_above(Fxm_cur[q]);
    bool bad_live = false;
    double bad_ro = 1.0, bad_uo = 0.0;
    if (col_out) {
      const long long idz = pb.off + ((long long)k * n2 + j) * n1 + i;
      double rho = me.rho, uu = me.uu, press = me.P, v[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < NV; d++) v[d] = me.v[d];
      double src[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (SRC) {
        double vs = 0.0;
        for (int d = 0; d < 3; d++) vs += v[d] * pb.g0[d];
        src[1] += rho * vs;
        for (int d = 0; d < NV; d++) src[2 + d] += rho * pb.g0[d];
      }
      double rn = rho, un = uu, vn[3] = {v[0], v[1], v[2]};
      if (stage == 1) {
        rn = sp.v[0][idz];
        un = sp.v[1][idz];
        for (int d = 0; d < NV; d++) vn[d] = sp.v[2 + d][idz];
      }
      const double one3[3] = {1.0, 1.0, 1.0};
      double U[5], Uh[5];
      p_to_u_x64<NV, true>(rn, un, vn, one3, U);
      p_to_u_x64<NV, true>(rho, uu, v, one3, Uh);
      if (SRC) {
        double dsrc = X64 ? dt_source_control_x64<NV>(un, vn, U, src) : dt_source_control_fast<NV>(un, vn, U, src);
        lim_src = (lim_src < dsrc) ? lim_src : dsrc;
      } else {
        double m0 = 0.02 * U[0], m1 = 0.02 * U[1], m2 = 2.0 * un;
        double m = (m0 < m1) ? m0 : m1;
        m = (m < m2) ? m : m2;
        srcnum = (srcnum < m) ? srcnum : m;
      }
      double ro, uo, vo[3];
      if constexpr (X64) {
        for (int q = 0; q < 2 + NV; q++) {
          double divF = ((((0.0 + Fxp_cur[q]) - Fxm_cur[q]) + sFy[q][r + 1][lane]) - sFy[q][r][lane]);
          divF += Fzp[q] - Fzm[q];
          double ddt = (divF / pb.volume) - src[q];
          U[q] = ((pf * U[q]) + (phf * Uh[q])) - ((ff * dt) * ddt);
        }
        u_to_p_x64<NV, true>(U, one3, ro, uo, vo);
      } else {
        for (int q = 0; q < 2 + NV; q++) {
          double divF = ((double)Fxp_cur[q] - (double)Fxm_cur[q]) * idx[0]
                        + ((double)sFy[q][r + 1][lane] - (double)sFy[q][r][lane]) * idx[1]
                        + ((double)Fzp[q] - (double)Fzm[q]) * idx[2];
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
      out.v[0][idz] = (R)ro;
      out.v[1][idz] = (R)uo;
      for (int d = 0; d < NV; d++) out.v[2 + d][idz] = (R)vo[d];
      if (stage == 0) a.flags[idz - pb.off] = (unsigned char)(tag_cur & 1);
      if (stage == 1) {
        double loc;
        if constexpr (X64) {
          double cs = sqrt(((double)pb.gam * press) / rho);
          loc = 0.0;
          for (int d = 0; d < 3; d++) loc += (fabs(d < NV ? vo[d] : 0.0) + cs) / pb.dx[d];
        } else {
          float cs = fsqrt((float)pb.gam * (float)press * frcp((float)rho));
          loc = 0.0;
          for (int d = 0; d < 3; d++) loc += (fabs(d < NV ? vo[d] : 0.0) + cs) * idx[d];
        }
        locmax = loc > locmax ? loc : locmax;
      }
    }

    nbad += bad_state_lanes(bad_live, bad_ro, bad_uo, true);
    if (zw) {
      for (int q = 0; q < 4; q++) Z[q] = Z[q + 1];
      zl_cur = zl_next; zr_cur = zr_next;
      tag_cur = tag_next;
      for (int q = 0; q < 2 + NV; q++) { Fzm[q] = Fzp[q]; Fxm_cur[q] = Fxm_next[q]; }
    }
  }

  double ws[3], wloc = wave_max(locmax);
  for (int d = 0; d < 3; d++) ws[d] = wave_max((double)smax[d]);
  double wsrc = wave_min(srcnum), wlim = wave_min(lim_src);
  if (lane == 0) {
    double lim = 1.0e300;
    for (int d = 0; d < 3; d++)
      if (ws[d] > 0.0) lim = fmin(lim, pb.dx[d] / ws[d]);
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

extern "C" hipError_t mds_launch_sweep3d(int fp32, int ty, int hshare, const Sweep3DArgs *a, const void *sp,
                                         const void *sph, int stage, hipStream_t st)
{
  dim3 grid((unsigned)(a->nstrip * a->ntile * a->nchunk));
  bool src = a->src;
  // the general instances (PROB/EXTRAP ghost layers, hydro_interp_order < 3) have one shape: 4 rows, halo wavefronts
  if (a->slab && (ty != 4 || hshare)) return hipErrorInvalidValue;
#define MDS_L3(R, SRC, TY)                                                                           \
  do {                                                                                               \
    if (a->slab)                                                                                     \
      hipLaunchKernelGGL((sweep3d<R, 3, SRC, 4, true>), grid, dim3(8 * 64), 0, st, *a,              \
                         *(const StateArrays<R> *)sp, *(const StateArrays<R> *)sph, stage);          \
    else if (hshare)                                                                                 \
      hipLaunchKernelGGL((sweep3dh<R, 3, SRC, TY>), grid, dim3(TY * 64), 0, st, *a,                \
                         *(const StateArrays<R> *)sp, *(const StateArrays<R> *)sph, stage);          \
    else                                                                                             \
      hipLaunchKernelGGL((sweep3d<R, 3, SRC, TY>), grid, dim3((TY + 4) * 64), 0, st, *a,           \
                         *(const StateArrays<R> *)sp, *(const StateArrays<R> *)sph, stage);          \
  } while (0)
  if (!fp32) {
    if (ty == 4) { if (src) MDS_L3(double, true, 4); else MDS_L3(double, false, 4); }
    else if (ty == 8) { if (src) MDS_L3(double, true, 8); else MDS_L3(double, false, 8); }
    else { if (src) MDS_L3(double, true, 12); else MDS_L3(double, false, 12); }
  } else {
    if (ty == 4) { if (src) MDS_L3(float, true, 4); else MDS_L3(float, false, 4); }
    else if (ty == 8) { if (src) MDS_L3(float, true, 8); else MDS_L3(float, false, 8); }
    else { if (src) MDS_L3(float, true, 12); else MDS_L3(float, false, 12); }
  }
#undef MDS_L3
  return hipGetLastError();
}
