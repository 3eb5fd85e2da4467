// mds_sph3d.hip -- RK stage of a 3D spherical (r, theta, phi) grid (gfx950),
// uniform or dendritic.
//
// Direction-split, as the reference's prim_to_fluxes3 + push (step.c:1088-1430,
// 327-478): per stage
//   eos_kernel       tabulated EOS of every zone (eos.c:4-32; skipped for a
//                    gamma law, which is evaluated in place)
//   shell kernels    monopole gravity per radial shell (gravity.c:145-498)
//   flux<0>          radial faces: HLLC, shock tags on zones i, i-1, i-2
//   flux<1>, flux<2> theta and phi faces: HLLE next to a tagged zone, else
//                    HLLC (fluid.c:189-210); the phi faces of the last theta
//                    row use the polar-axis states of step.c:1299-1397
//   update_kernel    geometric and gravity sources, flux divergence, the RK
//                    update, u -> p and the dt limits (step.c:327-478)
// One thread per face or zone, i fastest.  Ghost zones are not stored: every
// stencil index outside the grid is mapped onto the zone the reference's
// ghost fill copies there (bc.c:705-1421): radial REFLECT/OUTFLOW, theta
// across the pole (SPHERICAL_ORIGIN: mirror in j, phi + pi, v2 and v3
// flipped) or REFLECT/OUTFLOW, phi PERIODIC.  Face fluxes are stored in the
// state precision: FP64 area weighted, or (FP32-bulk) the FP32 flux per unit
// area, which the update multiplies by the FP64 face area -- the same product
// the flux kernel would form, so storing it halves the flux traffic and
// changes no bit.
//
// Dendritic grids (utils.c:181-590) keep the (i, j, k) storage with
// level-local j < nj(i), k < nk(i, j).  The flux kernels above handle the
// faces whose 4-zone pencil stays on one level; the radial and theta faces
// whose pencils cross a level change are listed at set-up and done by
// flux_irr with the reference's prolongation (kslopes0/1, prolongate0/1,
// limited linear) and volume-weighted restriction (restrict0/1),
// step.c:783-1085.  The update gathers the finer faces of a coarser zone
// (the fan-in of step.c:374-433).
#include <type_traits>
#include "mds_internal.h"
#include "mds_physics.h"
#include "mds_eos.h"
#include "mds_gravity.h"
#include "../common/mds_mp.h"

namespace {

struct Sph3Args {
  DevProb pb;
  DevClock *clock;
  unsigned long long *counters;
  unsigned char *flags;       // 1: no shock tag, 0: tagged (interior zones)
  DevEos eos;
  void *F[3];                 // face fluxes [slot][face] per direction, slot 5 = Ye
  long long nf[3];            // faces per direction
  const void *P, *G, *CS;     // table EOS arrays (state precision)
  unsigned short *tg;         // each zone's temperature bracket of the last EOS call
  const double *grav;         // monopole field per radial shell, or null
  const double *mpsrc;        // multipole sources [6][zones], or null
  int upd;                    // update workgroup shape (update_zone_of)
};

MDS_DEV long long zid3(const DevProb &pb, int i, int j, int k)
{
  return (long long)i + (long long)pb.n[0] * ((long long)j + (long long)pb.n[1] * k);
}

MDS_DEV long long fid3(const DevProb &pb, int D, int i, int j, int k)
{
  const long long a = pb.n[0] + (D == 0), b = pb.n[1] + (D == 1);
  return (long long)i + a * ((long long)j + b * k);
}

// store the fluxes of face f in direction D (slots 0..nslot-1)
template <class R>
MDS_DEV void store_flux(const Sph3Args &a, int D, long long f, const double F[6], int nslot)
{
  R *p = (R *)a.F[D];
  for (int v = 0; v < nslot; v++) p[(size_t)v * a.nf[D] + f] = (R)F[v];
}

// Per-(i, j) geometry record.  On the device the records are stored field
// by field (zone[field][j + NG][i + NG]): a wavefront runs along i, so every
// field load is one contiguous 512-byte read instead of 64 cache lines.
struct ZRec {
  const double *base;
  size_t stride;
  MDS_DEV double operator[](int f) const { return base[(size_t)f * stride]; }
};

MDS_DEV ZRec rec3(const DevProb &pb, int i, int j)
{
  const int N1 = pb.n[0] + 2 * MDS_NG, N2 = pb.n[1] + 2 * MDS_NG;
  ZRec r;
  r.stride = (size_t)N1 * N2;
  r.base = pb.zone + (size_t)(j + MDS_NG) * N1 + (i + MDS_NG);
  return r;
}

// zone counts of radial zone i and of row (i, j) at their levels
MDS_DEV int nj_of(const DevProb &pb, int i) { return pb.n[1] >> dend_lj(pb, i); }
MDS_DEV int nk_of(const DevProb &pb, int i, int j) { return pb.n[2] >> dend_lk(pb, i, j); }

// index idx of level `from` expressed on level `to` (idx * 2^from / 2^to)
MDS_DEV int rescale(int idx, int from, int to) { return from >= to ? idx << (from - to) : idx >> (to - from); }

// volume moments (alpha, beta, Gamma per zone, zone -NG first) of direction D
// at refinement level lev, and the FP32 coefficients
MDS_DEV const double *mom_of(const DevProb &pb, int D, int lev)
{
  return (D == 0 || !pb.lj) ? pb.mom[D] : pb.lmom[D - 1] + (size_t)lev * (pb.n[D] + 2 * MDS_NG) * 3;
}
MDS_DEV const float *rc_of(const DevProb &pb, int D, int lev)
{
  return (D == 0 || !pb.lj) ? pb.rc[D] : pb.lrc[D - 1] + (size_t)lev * (pb.n[D] + 2 * MDS_NG) * MDS_RC_COUNT;
}

// Map a stencil zone that leaves the grid in direction D onto the zone the
// ghost fill copies there; sgn gets the velocity sign flips.  n is the zone
// count along D at the zone's level; nkm (theta across the pole only) the phi
// count of the mirrored row.  The inner radius through the origin
// (SPHERICAL_ORIGIN, bc.c:774-801) mirrors i, takes theta to pi - theta at
// the level of the mirrored shell and phi to phi + pi, flipping v^1 and v^3.
MDS_DEV void map_zone_n(const DevProb &pb, int D, int &i, int &j, int &k, double sgn[3], int n, int nkm)
{
  sgn[0] = sgn[1] = sgn[2] = 1.0;
  int *c = D == 0 ? &i : (D == 1 ? &j : &k);
  if (*c >= 0 && *c < n) return;
  const int side = *c >= n;
  const int code = side ? pb.bc_hi[D] : pb.bc_lo[D];
  const int ig = *c;
  if (code == MDS_BC_PERIODIC) {
    *c = side ? ig - n : ig + n;
  } else if (code == MDS_BC_REFLECT) {
    if (D == 2) *c = side ? n - 1 : 0;                 // bc.c:1285-1300: copy, not mirror
    else *c = side ? 2 * n - 1 - ig : -1 - ig;
    sgn[D] = -1.0;
  } else if (code == MDS_BC_SPHERICAL_ORIGIN && D == 1) {
    *c = side ? 2 * n - 1 - ig : -1 - ig;
    k = (k + nkm / 2) % nkm;
    sgn[1] = sgn[2] = -1.0;
  } else if (code == MDS_BC_PROB || code == MDS_BC_EXTRAP) {
    // outer radius with ghost layers of its own: the index stays in the
    // ghost shell and the caller reads the layer (slab_state)
  } else if (code == MDS_BC_SPHERICAL_ORIGIN && D == 0 && !side) {
    i = -1 - ig;
    j = nj_of(pb, i) - 1 - j;
    const int nk = nk_of(pb, i, j);
    k = (k + nk / 2) % nk;
    sgn[0] = sgn[2] = -1.0;
  } else {
    *c = side ? n - 1 : 0;
  }
}

// map_zone_n with the counts of the zone's level read from the level tables
MDS_DEV void map_zone(const DevProb &pb, int D, int &i, int &j, int &k, double sgn[3])
{
  const int n = D == 0 ? pb.n[0] : (D == 1 ? nj_of(pb, i) : nk_of(pb, i, j));
  int nkm = 1;
  if (D == 1 && (j < 0 || j >= n)) nkm = nk_of(pb, i, j < 0 ? -1 - j : 2 * n - 1 - j);
  map_zone_n(pb, D, i, j, k, sgn, n, nkm);
}

// Any stencil zone of a dendritic grid, corners included: phi wrap within its
// row, then the radial fill, then the theta fill -- the composition of the
// reference's fills (its corner ghosts are never filled).
MDS_DEV void map_any(const DevProb &pb, int &i, int &j, int &k, double sgn[3])
{
  sgn[0] = sgn[1] = sgn[2] = 1.0;
  const int nk = nk_of(pb, i, j);
  if (k < 0) k += nk;
  else if (k >= nk) k -= nk;
  double s[3];
  if (i < 0 || i >= pb.n[0]) {
    map_zone(pb, 0, i, j, k, s);
    for (int d = 0; d < 3; d++) sgn[d] *= s[d];
  }
  if (j < 0 || j >= nj_of(pb, i)) {
    map_zone(pb, 1, i, j, k, s);
    for (int d = 0; d < 3; d++) sgn[d] *= s[d];
  }
}

template <class R>
MDS_DEV void zone_state(const Sph3Args &a, const StateArrays<R> &in, long long z, int tab, double gam, double q[8])
{
  // q = rho, u, v1, v2, v3, Ye, P, Gamma
  q[0] = in.v[0][z];
  q[1] = in.v[1][z];
  for (int d = 0; d < 3; d++) q[2 + d] = in.v[2 + d][z];
  if (tab) {
    q[5] = in.v[5][z];
    q[6] = ((const R *)a.P)[z];
    q[7] = ((const R *)a.G)[z];
  } else {
    q[5] = 0.0;
    q[6] = (gam - 1.0) * q[1];
    q[7] = gam;
  }
}

// the pencil values of a zone of the outer ghost shells of a PROB or EXTRAP
// radius (i >= n[0]): its ghost layer, at the state's precision
template <class R>
MDS_DEV void slab_state(const DevProb &pb, int i, int j, int k, int tab, double q[8])
{
  const long long T = slab_T(pb, 0);
  const double *e = pb.gs[0][1] + (size_t)(i - pb.n[0]) * MDS_SLAB_NV * T + slab_t(pb, 0, i, j, k);
  for (int v = 0; v < 5; v++) q[v] = (double)(R)e[v * T];
  if (tab) {
    q[5] = (double)(R)e[5 * T];
    q[6] = (double)(R)e[6 * T];
    q[7] = (double)(R)e[7 * T];
  } else {
    q[5] = 0.0;
    q[6] = (pb.gam - 1.0) * q[1];
    q[7] = pb.gam;
  }
}

// the pencil values of any (possibly ghost or corner) zone
template <class R>
MDS_DEV void any_state(const Sph3Args &a, const StateArrays<R> &in, int i, int j, int k, int tab, double q[8])
{
  double sgn[3];
  map_any(a.pb, i, j, k, sgn);
  if (i >= a.pb.n[0]) slab_state<R>(a.pb, i, j, k, tab, q);
  else zone_state<R>(a, in, zid3(a.pb, i, j, k), tab, a.pb.gam, q);
  for (int d = 0; d < 3; d++) q[2 + d] *= sgn[d];
}

// Polar-axis states of the phi face at phi_k between zones k-1 (L) and k (R):
// velocities rotated through Cartesian components (step.c:1299-1397); dk is
// the row's phi coarsening
MDS_DEV void polar_rotate(const DevProb &pb, const ZRec &rc, int k, int dk, double vl[3], double vr[3])
{
  const double sth = rc[MDS_ZS_STH], cth = rc[MDS_ZS_CTH];
  double *vv[2] = {vl, vr};
  for (int s = 0; s < 2; s++) {
    const int kt = k - 1 + s;
    double ph = pb.startx[2] + (kt + 0.5) * dk * pb.dx[2];
    double sph = sin(ph), cph = cos(ph);
    const double lam[3][3] = {{sth * cph, cth * cph, -sph}, {sth * sph, cth * sph, cph}, {cth, -sth, 0.0}};
    double v[3];
    for (int l = 0; l < 3; l++) {
      v[l] = 0.0;
      for (int m = 0; m < 3; m++) v[l] += lam[l][m] * vv[s][m] * rc[MDS_ZS_SCALE0 + m];
    }
    ph = pb.startx[2] + (k * dk) * pb.dx[2];
    sph = sin(ph);
    cph = cos(ph);
    const double lf[3][3] = {{sth * cph, cth * cph, -sph}, {sth * sph, cth * sph, cph}, {cth, -sth, 0.0}};
    for (int l = 0; l < 3; l++) {
      double o = 0.0;
      for (int m = 0; m < 3; m++) o += lf[m][l] * v[m] / rc[MDS_ZS_SCALE0 + l];
      vv[s][l] = o;
    }
  }
}

MDS_DEV double minmod(double a, double b)
{
  // decs.h:629-630
  const double sa = a < 0.0 ? -1.0 : 1.0, sb = b < 0.0 ? -1.0 : 1.0;
  return (0.5 * (sa + sb)) * fmin(fabs(a), fabs(b));
}

// Reconstruction and Riemann solve of one face from its 4-zone pencil q
// (zones fd-2 .. fd+1 along D, level lev).  GREC: reconstruct Gamma_1 of a
// gamma law too (a restricted radial pencil carries v^3 in that slot).
// hlle: transverse face next to a tagged zone.  Returns isshock.  F is area
// weighted in FP64 and per unit area in FP32 (zero on faces of area < 1e-16).
template <int D>
MDS_DEV FaceGeo<double> face_geo(const ZRec &rc)
{
  FaceGeo<double> fg;
  if (D == 0) {
    fg.area = rc[MDS_ZS_AREA0];
    fg.scale = rc[MDS_ZS_SCALEF0];
    for (int d = 0; d < 3; d++) fg.gcov[d] = rc[MDS_ZS_GCOVF0 + d];
  } else if (D == 1) {
    fg.area = rc[MDS_ZS_AREA1];
    fg.scale = rc[MDS_ZS_SCALEF1];
    for (int d = 0; d < 3; d++) fg.gcov[d] = rc[MDS_ZS_GCOVF1 + d];
  } else {
    fg.area = rc[MDS_ZS_AREA2];
    fg.scale = rc[MDS_ZS_SCALE0 + 2];
    for (int d = 0; d < 3; d++) fg.gcov[d] = rc[MDS_ZS_GCOV0 + d];
  }
  return fg;
}

template <class R, int D, int TAB, int GREC, int LO = 0>
MDS_DEV int face_flux(const DevProb &pb, int lev, int fd, double q[4][8], bool polar, int fk, int dk,
                      const ZRec &rc, bool hlle, double F[6], double &sig)
{
  const FaceGeo<double> fg = face_geo<D>(rc);
  int isshock = 0;
  sig = 0.0;
  for (int v = 0; v < 6; v++) F[v] = 0.0;
  const int s = 1 << lev;
  if (sizeof(R) == 8) {
    // FP64: reference operation order
    const double *mom = mom_of(pb, D, lev);
    const int b0 = fd - 2 + MDS_NG;
    ReconGeo64 gl, gr;     // zones fd-1 (left of the face) and fd (right)
    gl.am = mom[3 * b0]; gl.bm = mom[3 * b0 + 1]; gl.gm = mom[3 * b0 + 2];
    gl.a0 = mom[3 * (b0 + 1)]; gl.b0 = mom[3 * (b0 + 1) + 1]; gl.g0 = mom[3 * (b0 + 1) + 2];
    gl.ap = mom[3 * (b0 + 2)]; gl.bp = mom[3 * (b0 + 2) + 1]; gl.gp = mom[3 * (b0 + 2) + 2];
    gr.am = gl.a0; gr.bm = gl.b0; gr.gm = gl.g0;
    gr.a0 = gl.ap; gr.b0 = gl.bp; gr.g0 = gl.gp;
    gr.ap = mom[3 * (b0 + 3)]; gr.bp = mom[3 * (b0 + 3) + 1]; gr.gp = mom[3 * (b0 + 3) + 2];
    gl.xlo = pb.startx[D] + ((fd - 1) * s) * pb.dx[D];
    gl.xhi = pb.startx[D] + (fd * s) * pb.dx[D];
    gr.xlo = gl.xhi;
    gr.xhi = pb.startx[D] + ((fd + 1) * s) * pb.dx[D];
    recon_geo_x64(gl);
    recon_geo_x64(gr);
    gl.order = gr.order = pb.order;
    double Lq[8], Rq[8];
    if (polar) {
#pragma unroll
      for (int v = 0; v < 8; v++) { Lq[v] = q[1][v]; Rq[v] = q[2][v]; }
    } else {
#pragma unroll
      for (int v = 0; v < 8; v++) {
        if (!TAB && (v == 5 || (v == 7 && !GREC))) { Lq[v] = q[1][v]; Rq[v] = q[2][v]; continue; }
        double al, ar, al2, ar2;
        if (LO) {   // hydro_interp_order 2 or 1
          recon_x64(q[0][v], q[1][v], q[2][v], gl, al, ar);
          recon_x64(q[1][v], q[2][v], q[3][v], gr, al2, ar2);
        } else {
          para_recon_x64(q[0][v], q[1][v], q[2][v], gl, al, ar);
          para_recon_x64(q[1][v], q[2][v], q[3][v], gr, al2, ar2);
        }
        Lq[v] = ar;
        Rq[v] = al2;
      }
    }
    if (polar) {
      double vl[3] = {q[1][2], q[1][3], q[1][4]}, vr[3] = {q[2][2], q[2][3], q[2][4]};
      polar_rotate(pb, rc, fk, dk, vl, vr);
      for (int d = 0; d < 3; d++) { Lq[2 + d] = vl[d]; Rq[2 + d] = vr[d]; }
    }
    PState<double, 3> L, Rt;
    L.rho = Lq[0]; L.uu = Lq[1]; L.v[0] = Lq[2]; L.v[1] = Lq[3]; L.v[2] = Lq[4]; L.ye = Lq[5]; L.P = Lq[6]; L.G = Lq[7];
    Rt.rho = Rq[0]; Rt.uu = Rq[1]; Rt.v[0] = Rq[2]; Rt.v[1] = Rq[3]; Rt.v[2] = Rq[4]; Rt.ye = Rq[5]; Rt.P = Rq[6];
    Rt.G = Rq[7];
    if (D == 0 || !hlle) isshock = hllc_x64<3, false, TAB>(L, Rt, fg, D, F, sig);
    else hlle_x64<3, false, TAB>(L, Rt, fg, D, F, sig);
  } else {
    // FP32-bulk: reconstruction and fluxes per unit area in FP32
    const float *rcf = rc_of(pb, D, lev);
    const int b0 = fd - 1 + MDS_NG;
    ReconGeo32 gl, gr;
    const float *cl = rcf + (size_t)b0 * MDS_RC_COUNT, *cr = rcf + (size_t)(b0 + 1) * MDS_RC_COUNT;
    gl.clm = cl[MDS_RC_CLM]; gl.clp = cl[MDS_RC_CLP]; gl.crm = cl[MDS_RC_CRM]; gl.crp = cl[MDS_RC_CRP];
    gl.mu1 = cl[MDS_RC_MU1]; gl.kinv = cl[MDS_RC_KINV]; gl.inur = cl[MDS_RC_INUR]; gl.imu2 = cl[MDS_RC_IMU2];
    gr.clm = cr[MDS_RC_CLM]; gr.clp = cr[MDS_RC_CLP]; gr.crm = cr[MDS_RC_CRM]; gr.crp = cr[MDS_RC_CRP];
    gr.mu1 = cr[MDS_RC_MU1]; gr.kinv = cr[MDS_RC_KINV]; gr.inur = cr[MDS_RC_INUR]; gr.imu2 = cr[MDS_RC_IMU2];
    float Lq[8], Rq[8];
    if (polar) {
#pragma unroll
      for (int v = 0; v < 8; v++) { Lq[v] = (float)q[1][v]; Rq[v] = (float)q[2][v]; }
    } else {
#pragma unroll
      for (int v = 0; v < 8; v++) {
        if (!TAB && (v == 5 || (v == 7 && !GREC))) { Lq[v] = (float)q[1][v]; Rq[v] = (float)q[2][v]; continue; }
        float al, ar, al2, ar2;
        para_recon_f32((float)q[0][v], (float)q[1][v], (float)q[2][v], gl, al, ar);
        para_recon_f32((float)q[1][v], (float)q[2][v], (float)q[3][v], gr, al2, ar2);
        Lq[v] = ar;
        Rq[v] = al2;
      }
    }
    if (polar) {
      double vl[3] = {q[1][2], q[1][3], q[1][4]}, vr[3] = {q[2][2], q[2][3], q[2][4]};
      polar_rotate(pb, rc, fk, dk, vl, vr);
      for (int d = 0; d < 3; d++) { Lq[2 + d] = (float)vl[d]; Rq[2 + d] = (float)vr[d]; }
    }
    PState<float, 3> L, Rt;
    L.rho = Lq[0]; L.uu = Lq[1]; L.v[0] = Lq[2]; L.v[1] = Lq[3]; L.v[2] = Lq[4]; L.ye = Lq[5]; L.P = Lq[6]; L.G = Lq[7];
    Rt.rho = Rq[0]; Rt.uu = Rq[1]; Rt.v[0] = Rq[2]; Rt.v[1] = Rq[3]; Rt.v[2] = Rq[4]; Rt.ye = Rq[5]; Rt.P = Rq[6];
    Rt.G = Rq[7];
    FaceGeo<float> g32;
    g32.area = 1.0f;
    g32.scale = (float)fg.scale;
    for (int d = 0; d < 3; d++) g32.gcov[d] = (float)fg.gcov[d];
    float Ff[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, s32 = 0.0f;
    if (fg.area >= 1.0e-16) {
      if (D == 0 || !hlle) isshock = hllc_f32<3, false, TAB>(L, Rt, g32, D, Ff, s32);
      else hlle_f32<3, false, TAB>(L, Rt, g32, D, Ff, s32);
      for (int v = 0; v < 6; v++) F[v] = (double)Ff[v];   // per unit area
      sig = s32;
    } else {
      sig = 1.0e-30;
    }
  }
  return isshock;
}

// PPM of one zone along D from its 3-zone stencil q3 (zones zd-1, zd, zd+1,
// level lev): the left and right face values of every pencil slot.  Equal,
// slot by slot, to the values face_flux forms for the two faces of the zone.
template <class R, int D, int TAB, int LO = 0>
MDS_DEV void zone_recon(const DevProb &pb, int lev, int zd, const double q3[3][8], R al[8], R ar[8])
{
  const int s = 1 << lev;
  if (sizeof(R) == 8) {
    const double *mom = mom_of(pb, D, lev);
    const int b0 = zd - 1 + MDS_NG;
    ReconGeo64 g;
    g.am = mom[3 * b0]; g.bm = mom[3 * b0 + 1]; g.gm = mom[3 * b0 + 2];
    g.a0 = mom[3 * (b0 + 1)]; g.b0 = mom[3 * (b0 + 1) + 1]; g.g0 = mom[3 * (b0 + 1) + 2];
    g.ap = mom[3 * (b0 + 2)]; g.bp = mom[3 * (b0 + 2) + 1]; g.gp = mom[3 * (b0 + 2) + 2];
    g.xlo = pb.startx[D] + (zd * s) * pb.dx[D];
    g.xhi = pb.startx[D] + ((zd + 1) * s) * pb.dx[D];
    recon_geo_x64(g);
    g.order = pb.order;
#pragma unroll
    for (int v = 0; v < 8; v++) {
      if (!TAB && (v == 5 || v == 7)) { al[v] = (R)q3[1][v]; ar[v] = (R)q3[1][v]; continue; }
      double l, r;
      if (LO) recon_x64(q3[0][v], q3[1][v], q3[2][v], g, l, r);     // hydro_interp_order 2 or 1
      else para_recon_x64(q3[0][v], q3[1][v], q3[2][v], g, l, r);
      al[v] = (R)l;
      ar[v] = (R)r;
    }
  } else {
    const float *c = rc_of(pb, D, lev) + (size_t)(zd + MDS_NG) * MDS_RC_COUNT;
    ReconGeo32 g;
    g.clm = c[MDS_RC_CLM]; g.clp = c[MDS_RC_CLP]; g.crm = c[MDS_RC_CRM]; g.crp = c[MDS_RC_CRP];
    g.mu1 = c[MDS_RC_MU1]; g.kinv = c[MDS_RC_KINV]; g.inur = c[MDS_RC_INUR]; g.imu2 = c[MDS_RC_IMU2];
#pragma unroll
    for (int v = 0; v < 8; v++) {
      if (!TAB && (v == 5 || v == 7)) { al[v] = (R)q3[1][v]; ar[v] = (R)q3[1][v]; continue; }
      float l, r;
      para_recon_f32((float)q3[0][v], (float)q3[1][v], (float)q3[2][v], g, l, r);
      al[v] = (R)l;
      ar[v] = (R)r;
    }
  }
}

// Riemann solve of one face from its reconstructed states (rho, u, v1..v3,
// Ye, P, Gamma_1) and its geometry (face_geo of the face's zone record).  F
// as in face_flux.
template <class R, int D, int TAB>
MDS_DEV int face_riemann(const R Lq[8], const R Rq[8], const FaceGeo<double> &fg, bool hlle, double F[6],
                         double &sig)
{
  int isshock = 0;
  sig = 0.0;
  for (int v = 0; v < 6; v++) F[v] = 0.0;
  PState<R, 3> L, Rt;
  L.rho = Lq[0]; L.uu = Lq[1]; L.v[0] = Lq[2]; L.v[1] = Lq[3]; L.v[2] = Lq[4]; L.ye = Lq[5]; L.P = Lq[6]; L.G = Lq[7];
  Rt.rho = Rq[0]; Rt.uu = Rq[1]; Rt.v[0] = Rq[2]; Rt.v[1] = Rq[3]; Rt.v[2] = Rq[4]; Rt.ye = Rq[5]; Rt.P = Rq[6];
  Rt.G = Rq[7];
  if (sizeof(R) == 8) {
    if (D == 0 || !hlle) isshock = hllc_x64<3, false, TAB>(*(PState<double, 3> *)&L, *(PState<double, 3> *)&Rt, fg, D, F, sig);
    else hlle_x64<3, false, TAB>(*(PState<double, 3> *)&L, *(PState<double, 3> *)&Rt, fg, D, F, sig);
  } else {
    FaceGeo<float> g32;
    g32.area = 1.0f;
    g32.scale = (float)fg.scale;
    for (int d = 0; d < 3; d++) g32.gcov[d] = (float)fg.gcov[d];
    float Ff[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, s32 = 0.0f;
    if (fg.area >= 1.0e-16) {
      if (D == 0 || !hlle) isshock = hllc_f32<3, false, TAB>(*(PState<float, 3> *)&L, *(PState<float, 3> *)&Rt, g32, D, Ff, s32);
      else hlle_f32<3, false, TAB>(*(PState<float, 3> *)&L, *(PState<float, 3> *)&Rt, g32, D, Ff, s32);
      for (int v = 0; v < 6; v++) F[v] = (double)Ff[v];   // per unit area
      sig = s32;
    } else {
      sig = 1.0e-30;
    }
  }
  return isshock;
}

// shock-tag test of an interior zone (ghosts and non-zones count as tagged)
MDS_DEV bool tagged(const Sph3Args &a, int i, int j, int k)
{
  const DevProb &pb = a.pb;
  if (i < 0 || j < 0 || k < 0 || i >= pb.n[0] || j >= nj_of(pb, i) || k >= nk_of(pb, i, j)) return true;
  return a.flags[zid3(pb, i, j, k)] == 0;
}

// tagged() for an i inside the grid, with the row counts supplied; the tag
// byte is read unconditionally (zone (i, 0, 0) for a ghost) so that both
// zones' reads of a face go out together
MDS_DEV bool tagged_n(const Sph3Args &a, int i, int j, int k, int nj, int nk)
{
  const bool in = j >= 0 && k >= 0 && j < nj && k < nk;
  const unsigned char f = a.flags[zid3(a.pb, i, in ? j : 0, in ? k : 0)];
  return !in || f == 0;
}

// radial shock tags on zones i, i-1, i-2 at their levels (step.c:1190-1198)
MDS_DEV void tag_radial(const Sph3Args &a, int i, int j, int k)
{
  const DevProb &pb = a.pb;
  for (int s = 0; s < 3; s++) {
    const int ii = i - s;
    if (ii < 0 || ii >= pb.n[0]) continue;
    const int jm = rescale(j, dend_lj(pb, i), dend_lj(pb, ii));
    const int km = rescale(k, dend_lk(pb, i, j), dend_lk(pb, ii, jm));
    if (jm < nj_of(pb, ii) && km < nk_of(pb, ii, jm)) a.flags[zid3(pb, ii, jm, km)] = 0;
  }
}

// per-wave CFL (min width/|s| over the wave's faces) and counters
MDS_DEV void face_dt(const Sph3Args &a, unsigned long long seen, bool ok, double width, double sig, bool shock,
                     bool hl, int D)
{
  DevClock *ck = a.clock;
  const double w = (ok && sig != 0.0) ? width / fabs(sig) : 1.0e300;
  const double wmin = wave_min(w);
  const unsigned long long nsh = __popcll(__ballot(ok && D == 0 && shock));
  const unsigned long long nhl = __popcll(__ballot(ok && D > 0 && hl));
  if ((threadIdx.x & 63) == 0) {
    if (wmin < 1.0e300) fold_min_key(&ck->mindt, seen, wmin);
    if (nsh) atomicAdd(&a.counters[0], nsh);
    if (nhl) atomicAdd(&a.counters[1], nhl);
  }
}

// kslopes0 (step.c:783-840) / kslopes1 (:842-896): row (ii, jj) evaluated
// with a limited phi slope at the phi centroid of index kt on level lt
template <class R, int TAB>
MDS_DEV void kslopes(const Sph3Args &a, const StateArrays<R> &in, int lt, int kt, int ii, int jj, double pp[8])
{
  const DevProb &pb = a.pb;
  const int l = dend_lk(pb, ii, jj);
  const double *mp = mom_of(pb, 2, lt) + 3 * (kt + MDS_NG);
  const double x2p = mp[1] / mp[2];
  const int kk = rescale(kt, lt, l);
  const double *m = mom_of(pb, 2, l) + 3 * (kk + MDS_NG);
  const double x20 = m[-2] / m[-1], x21 = m[1] / m[2], x22 = m[4] / m[5];
  const double idx2m = 1.0 / (x21 - x20), idx2p = 1.0 / (x22 - x21), idx2c = 1.0 / (x22 - x20);
  double a0[8], a1[8], a2[8];
  any_state<R>(a, in, ii, jj, kk - 1, TAB, a0);
  any_state<R>(a, in, ii, jj, kk, TAB, a1);
  any_state<R>(a, in, ii, jj, kk + 1, TAB, a2);
#pragma unroll
  for (int v = 0; v < 8; v++) {
    const double sm = (a1[v] - a0[v]) * idx2m, sp = (a2[v] - a1[v]) * idx2p, sc = (a2[v] - a0[v]) * idx2c;
    const double sg = minmod(2.0 * minmod(sm, sp), sc);
    pp[v] = a1[v] + sg * (x2p - x21);
  }
}

// radial pencil entry of level ii for the face of zone (i, j, k): restrict0
// (step.c:1012-1049, whose last entry averages the primitive at the eos
// record's Gamma index, i.e. v^3, as the reference does) or prolongate0
// (:905-951)
template <class R, int TAB>
MDS_DEV void radial_entry(const Sph3Args &a, const StateArrays<R> &in, int i, int j, int k, int ii, double pp[8])
{
  const DevProb &pb = a.pb;
  const int li = dend_lj(pb, i), lii = dend_lj(pb, ii), lk0 = dend_lk(pb, i, j);
  if (lii < li) {
    const int njp = 1 << (li - lii), jp = j * njp;
    double vol_sum = 0.0;
    for (int jj = jp; jj < jp + njp; jj++) {
      const int l = dend_lk(pb, ii, jj);
      const int nkp = l <= lk0 ? 1 << (lk0 - l) : 0;
      const double vol = rec3(pb, ii, jj)[MDS_ZS_VOLUME];
      for (int kk = 0; kk < nkp; kk++) vol_sum += vol;
    }
    vol_sum = 1.0 / vol_sum;
    for (int v = 0; v < 8; v++) pp[v] = 0.0;
    for (int jj = jp; jj < jp + njp; jj++) {
      const int l = dend_lk(pb, ii, jj);
      const int nkp = l <= lk0 ? 1 << (lk0 - l) : 0;
      const double vol_frac = rec3(pb, ii, jj)[MDS_ZS_VOLUME] * vol_sum;
      for (int kk = k * nkp; kk < k * nkp + nkp; kk++) {
        double q[8];
        any_state<R>(a, in, ii, jj, kk, TAB, q);
#pragma unroll
        for (int v = 0; v < 7; v++) pp[v] += q[v] * vol_frac;
        pp[7] += q[4] * vol_frac;
      }
    }
    return;
  }
  const int jj = rescale(j, li, lii);
  const double *mp = mom_of(pb, 1, li) + 3 * (j + MDS_NG);
  const double x1p = mp[1] / mp[2];
  double p0[8], p1[8], p2[8];
  kslopes<R, TAB>(a, in, lk0, k, ii, jj - 1, p0);
  kslopes<R, TAB>(a, in, lk0, k, ii, jj, p1);
  kslopes<R, TAB>(a, in, lk0, k, ii, jj + 1, p2);
  const double *m = mom_of(pb, 1, lii) + 3 * (jj + MDS_NG);
  const double x10 = m[-2] / m[-1], x11 = m[1] / m[2], x12 = m[4] / m[5];
  const double idx1m = 1.0 / (x11 - x10), idx1p = 1.0 / (x12 - x11), idx1c = 1.0 / (x12 - x10);
#pragma unroll
  for (int v = 0; v < 8; v++) {
    const double sm = (p1[v] - p0[v]) * idx1m, sp = (p2[v] - p1[v]) * idx1p, sc = (p2[v] - p0[v]) * idx1c;
    const double sg = minmod(2.0 * minmod(sm, sp), sc);
    pp[v] = p1[v] + sg * (x1p - x11);
  }
}

// theta pencil entry of row jj for the face (i, j) at phi index kf of level
// lf = the finer of rows j-1, j: restrict1 (step.c:1051-1085), a copy, or
// prolongate1 through kslopes1 (:962-1010)
template <class R, int TAB>
MDS_DEV void theta_entry(const Sph3Args &a, const StateArrays<R> &in, int i, int lf, int kf, int jj, double pp[8])
{
  const DevProb &pb = a.pb;
  const int lr = dend_lk(pb, i, jj);
  if (lr < lf) {
    const int nkp = 1 << (lf - lr), k0 = kf * nkp;
    const double vol = rec3(pb, i, jj)[MDS_ZS_VOLUME];
    double vol_sum = 0.0;
    for (int kk = 0; kk < nkp; kk++) vol_sum += vol;
    vol_sum = 1.0 / vol_sum;
    const double vol_frac = vol * vol_sum;
    for (int v = 0; v < 8; v++) pp[v] = 0.0;
    for (int kk = k0; kk < k0 + nkp; kk++) {
      double q[8];
      any_state<R>(a, in, i, jj, kk, TAB, q);
#pragma unroll
      for (int v = 0; v < 8; v++) pp[v] += q[v] * vol_frac;
    }
  } else if (lr == lf) {
    any_state<R>(a, in, i, jj, kf, TAB, pp);
  } else {
    kslopes<R, TAB>(a, in, lf, kf, i, jj, pp);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// per-step shock-tag reset and the table EOS
// ---------------------------------------------------------------------------
__global__ void sph3_reset_flags(unsigned char *flags, long long n)
{
  long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (z < n) flags[z] = 1;
}

template <class R>
__global__ __launch_bounds__(256) void sph3_eos_kernel(Sph3Args a, StateArrays<R> in)
{
  const DevProb &pb = a.pb;
  const long long nz = (long long)pb.n[0] * pb.n[1] * pb.n[2];
  const long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  if (pb.lj && !dend_valid(pb, (int)(z % pb.n[0]), (int)((z / pb.n[0]) % pb.n[1]), (int)(z / ((long long)pb.n[0] * pb.n[1]))))
    return;
  double press, lt;
  int jt = a.tg[z];                    // the bracket search starts from the last one
  if (sizeof(R) == 8) {
    double g1, cs;
    eos_fill_x64(a.eos, (double)in.v[0][z], (double)in.v[1][z], (double)in.v[5][z], press, g1, cs, lt, &jt);
    ((R *)a.P)[z] = (R)press;
    ((R *)a.G)[z] = (R)g1;
    ((R *)a.CS)[z] = (R)cs;
  } else {
    float g1, cs;
    eos_fill_f32(a.eos, (double)in.v[0][z], (double)in.v[1][z], (double)in.v[5][z], press, g1, cs, lt, &jt);
    ((R *)a.P)[z] = (R)press;
    ((R *)a.G)[z] = (R)g1;
    ((R *)a.CS)[z] = (R)cs;
  }
  a.tg[z] = (unsigned short)jt;
  // table lookups off the table: one atomic per wavefront from its first active lane
  const unsigned long long m = __ballot(eos_off_table(a.eos, (double)in.v[0][z], (double)in.v[5][z], lt));
  if (m && (threadIdx.x & 63) == (unsigned)__builtin_ctzll(__ballot(true)))
    atomicAdd(&a.clock->noff, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// monopole gravity on radial shells: masses, then enclosed mass and the
// zone-centred field (gravity.c:229-233, 420-450)
// ---------------------------------------------------------------------------
// Shell masses in two passes, both in a fixed order: thread i of wave w of
// block (x, y) sums V rho over rows r = (4 y + w) MDS_SHELL_ROWS + t of shell
// i (row r = (j, k) = (r % n2, r / n2): a wavefront reads 64 consecutive
// shells of one row, coalesced), the block adds its four waves in order into
// part[y][i]; the field kernel adds the parts of a shell in order.
template <class R>
__global__ __launch_bounds__(256) void sph3_shell_part(Sph3Args a, StateArrays<R> in, double *part)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int r0 = (blockIdx.y * 4 + w) * MDS_SHELL_ROWS;
  const int r1 = min(r0 + MDS_SHELL_ROWS, n2 * n3);
  double s = 0.0;
  if (i < n1) {
    const int nj = nj_of(pb, i);
#pragma unroll 8
    for (int r = r0; r < r1; r++) {
      // every (i, j < n2, k < n3) is inside the stored box: read, then select
      const int j = r % n2, k = r / n2;
      const double v = rec3(pb, i, j)[MDS_ZS_VOLUME];
      const double rho = (double)in.v[0][zid3(pb, i, j, k)];
      const bool zone = j < nj && k < nk_of(pb, i, j);
      s += zone ? v * rho : 0.0;
    }
  }
  __shared__ double sh[4][64];
  sh[w][lane] = s;
  __syncthreads();
  if (w == 0 && i < n1) part[(size_t)blockIdx.y * n1 + i] = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
}

// g at radial face i from the mass inside it: -G M / r^2 dr/dx (gravity.c:420-425)
MDS_DEV double face_gravity(const DevProb &pb, int i, double mr)
{
  const double GN = 6.67384e-8;
  const double xe = pb.startx[0] + i * pb.dx[0];
  double r, drdx;
  if (pb.rcoord == MDS_RCOORD_LOG) {
    r = exp(xe);
    drdx = exp(xe);
  } else if (pb.rcoord == MDS_RCOORD_SINH) {
    r = pb.rtrans * sinh(fabs(xe) / pb.rtrans);
    drdx = cosh(fabs(xe) / pb.rtrans);
  } else {
    r = fabs(xe);
    drdx = 1.0;
  }
  return ((-GN * mr) / (r * r)) * drdx;
}

// One block: each shell's mass from its parts, the enclosed masses by a
// sequential running sum (thread 0, over 1024 shells at a time in LDS), then
// every shell's field at its volume centroid in parallel.
__global__ __launch_bounds__(1024) void sph3_shell_field(Sph3Args a, const double *part, int npart, double *ms,
                                                         double *grav)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0];
  __shared__ double sm[1024];
  double mr = 0.0;                                   // thread 0: running enclosed mass
  for (int base = 0; base < n1; base += 1024) {
    const int i = base + (int)threadIdx.x;
    double m = 0.0;
    if (i < n1) {
#pragma unroll 16
      for (int c = 0; c < npart; c++) m += part[(size_t)c * n1 + i];
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int cnt = min(1024, n1 - base);
      for (int t = 0; t < cnt; t++) {
        mr = mr + sm[t];
        sm[t] = mr;
      }
    }
    __syncthreads();
    if (i < n1) ms[i] = sm[threadIdx.x];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n1; i += blockDim.x) {
    const double ge = i == 0 ? 0.0 : face_gravity(pb, i, ms[i - 1]);
    const double gn = face_gravity(pb, i + 1, ms[i]);
    const double xl = pb.startx[0] + i * pb.dx[0];
    const double A = (gn - ge) / pb.dx[0];
    const double B = ge - A * xl;
    const double be = pb.mom[0][3 * (i + MDS_NG) + 1], ga = pb.mom[0][3 * (i + MDS_NG) + 2];
    grav[i] = (A * be) / ga + B;
  }
}

// GR effective monopole (gravity.c:142-375 with GR_MONOPOLE): the shell sums
// V rho, V P, V u and V v^1 in the same two fixed-order passes as the shell
// masses (part[q][y][i], q = rho, P, u, v), P from the stage's table EOS
// arrays or the gamma law; then one block reduces the parts of each shell
// into the problem's GR block and evaluates the field (mds_gravity.h).
template <class R, int TAB>
__global__ __launch_bounds__(256) void sph3_shell_part_gr(Sph3Args a, StateArrays<R> in, double *part, int npart)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int r0 = (blockIdx.y * 4 + w) * MDS_SHELL_ROWS;
  const int r1 = min(r0 + MDS_SHELL_ROWS, n2 * n3);
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n1) {
    const int nj = nj_of(pb, i);
#pragma unroll 4
    for (int r = r0; r < r1; r++) {
      const int j = r % n2, k = r / n2;
      const long long z = zid3(pb, i, j, k);
      const double v = rec3(pb, i, j)[MDS_ZS_VOLUME];
      const double u = (double)in.v[1][z];
      const double press = TAB ? (double)((const R *)a.P)[z] : (pb.gam - 1.0) * u;
      const bool zone = j < nj && k < nk_of(pb, i, j);
      s[0] += zone ? v * (double)in.v[0][z] : 0.0;
      s[1] += zone ? v * press : 0.0;
      s[2] += zone ? v * u : 0.0;
      s[3] += zone ? v * (double)in.v[2][z] : 0.0;
    }
  }
  __shared__ double sh[4][4][64];
  for (int q = 0; q < 4; q++) sh[q][w][lane] = s[q];
  __syncthreads();
  if (w == 0 && i < n1)
    for (int q = 0; q < 4; q++)
      part[((size_t)q * npart + blockIdx.y) * n1 + i] = ((sh[q][0][lane] + sh[q][1][lane]) + sh[q][2][lane]) + sh[q][3][lane];
}

__global__ __launch_bounds__(1024) void sph3_shell_field_gr(Sph3Args a, const double *part, int npart, double *grav)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0];
  double *sums = pb.gr + (size_t)MDS_GR_RHO * (n1 + 1);     // rho, P, u, v arrays follow each other
  for (int i = threadIdx.x; i < n1; i += blockDim.x)
    for (int q = 0; q < 4; q++) {
      double m = 0.0;
      for (int c = 0; c < npart; c++) m += part[((size_t)q * npart + c) * n1 + i];
      sums[(size_t)q * (n1 + 1) + i] = m;
    }
  __syncthreads();
  gr_monopole_block<1024>(pb, grav);
}

// ---------------------------------------------------------------------------
// multipole gravity (gravity.c:499-1110, SPHERICAL_MULTIPOLE_GRAV): the
// l >= 1 terms of the potential on top of the monopole.  Grid-only factors
// (zone integrals of the harmonics, harmonics at the corner angles, radial
// weights and factors, centroid weights) come from host tables made in the
// restatement's operation order; the kernels form the moments of the stage
// state, their running sums, the potential at every zone corner and the
// corner-difference sources, which the update subtracts after the monopole's.
// ---------------------------------------------------------------------------
// Moment parts: thread i of wave w of block (x, y, z) sums, over rows
// r = (4 y + w) MDS_MP_ROWS + t of shell i (r = (j, k) = (r % n2, r / n2)),
// rho S[l,m](zone) times the shell's rc and rd for the pair z; the block adds
// its four waves in order.
template <class R>
__global__ __launch_bounds__(256) void sph3_mp_part(Sph3Args a, MpDev M, StateArrays<R> in)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const short2 lm = M.lms[blockIdx.z];
  const int L1 = M.lmax + 1;
  const int r0 = (blockIdx.y * 4 + w) * MDS_MP_ROWS;
  const int r1 = min(r0 + MDS_MP_ROWS, n2 * n3);
  double c = 0.0, d = 0.0;
  if (i < n1) {
    const int lj = dend_lj(pb, i), nj = n2 >> lj;
    const double wc = M.rc[i * L1 + lm.y], wd = M.rd[i * L1 + lm.y];
    for (int r = r0; r < r1; r++) {
      const int j = r % n2, k = r / n2;
      if (j >= nj) continue;
      const int lk = dend_lk(pb, i, j);
      if (k >= (n3 >> lk)) continue;
      const double S = M.S[(M.Soff[lj * (M.maxklev + 1) + lk] + (long long)j * (n3 >> lk) + k) * M.nlm + lm.x];
      const double rs = (double)in.v[0][zid3(pb, i, j, k)] * S;
      c += rs * wc;
      d += rs * wd;
    }
  }
  __shared__ double sh[2][4][64];
  sh[0][w][lane] = c;
  sh[1][w][lane] = d;
  __syncthreads();
  if (w == 0 && i < n1) {
    double *p = M.part + (((size_t)blockIdx.z * M.npart + blockIdx.y) * n1 + i) * 2;
    p[0] = ((sh[0][0][lane] + sh[0][1][lane]) + sh[0][2][lane]) + sh[0][3][lane];
    p[1] = ((sh[1][0][lane] + sh[1][1][lane]) + sh[1][2][lane]) + sh[1][3][lane];
  }
}

// each shell's moments from its parts, in part order
__global__ __launch_bounds__(256) void sph3_mp_sum(Sph3Args a, MpDev M)
{
  const int n1 = a.pb.n[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
  if (i >= n1) return;
  double c = 0.0, d = 0.0;
  for (int y = 0; y < M.npart; y++) {
    const double *p = M.part + (((size_t)z * M.npart + y) * n1 + i) * 2;
    c += p[0];
    d += p[1];
  }
  M.C[(size_t)z * n1 + i] = c;
  M.D[(size_t)z * n1 + i] = d;
}

// running sums: C over the shells inside, D over the shells outside (gravity.c:822-839)
__global__ void sph3_mp_scan(Sph3Args a, MpDev M)
{
  const int n1 = a.pb.n[0];
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= M.nlist) return;
  double *C = M.C + (size_t)z * n1, *D = M.D + (size_t)z * n1;
  for (int i = 1; i < n1; i++) C[i] += C[i - 1];
  for (int i = n1 - 2; i >= 0; i--) D[i] += D[i + 1];
}

// the potential at every zone corner (gravity.c:848-906), one thread per box corner
__global__ __launch_bounds__(256) void sph3_mp_phi(Sph3Args a, MpDev M)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)(n1 + 1) * (n2 + 1) * (n3 + 1)) return;
  const int ii = (int)(q % (n1 + 1)), jj = (int)((q / (n1 + 1)) % (n2 + 1)), kk = (int)(q / ((long long)(n1 + 1) * (n2 + 1)));
  const int lj = dend_lj(pb, ii);
  if (jj > (n2 >> lj)) return;
  const int lk = dend_lk(pb, ii, jj);
  if (pb.ndim == 3 ? kk > (n3 >> lk) : kk != 0) return;
  const int J = jj << lj, K = pb.ndim == 3 ? kk << lk : 0;
  const double *Y = M.Y + ((size_t)J * (n3 + 1) + K) * M.nlm;
  const int L1 = M.lmax + 1;
  double v = 0.0;
  const double rm = M.rm[ii];
  if (rm < 1.0e-16) {
  } else {
    int z = 1;                                   // pair index: the l = 0 pair comes first
    for (int l = 1; l <= M.lmax; l++) {
      const double Cl = M.Cf[ii * L1 + l], Dl = M.Df[ii * L1 + l];
      const int mm = l < M.mmax ? l : M.mmax;
      for (int m = -mm; m <= mm; m++, z++) {
        const double y = Y[MDS_MP_LM(l, m)];
        // no mass inside the inner radius of a grid that does not reach the origin
        if (ii > 0) v += (Cl * y) * M.C[(size_t)z * n1 + ii - 1];
        if (ii < n1) v += (Dl * y) * M.D[(size_t)z * n1 + ii];
      }
    }
  }
  M.phi[((size_t)ii * (n2 + 1) + jj) * (n3 + 1) + kk] = v;
}

// Corner-difference sources (gravity.c:919-1103): per zone the potential's
// difference along each direction over its eight corners, weighted by the
// volume centroid (1/2 in phi), times -rho / dx and -<rho v> / dx, with
// <rho v> from this stage's mass fluxes per area (the level rules of the
// theta faces included), weighted alike.
template <class R>
__global__ __launch_bounds__(256) void sph3_mp_src(Sph3Args a, MpDev M, StateArrays<R> in)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  const long long nz = (long long)n1 * n2 * n3;
  const long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  const int i = (int)(z % n1), j = (int)((z / n1) % n2), k = (int)(z / ((long long)n1 * n2));
  const int li = dend_lj(pb, i);
  if (j >= (n2 >> li)) return;
  const int lk0 = dend_lk(pb, i, j);
  if (k >= (n3 >> lk0)) return;
  const bool d3 = pb.ndim == 3;
  const double *phi = M.phi;
  auto corner = [&](int ii, int jj, int kk) { return phi[((size_t)ii * (n2 + 1) + jj) * (n3 + 1) + kk]; };
  const double wx[2] = {M.wx[2 * i], M.wx[2 * i + 1]};
  const double wy[2] = {M.wy[2 * (li * n2 + j)], M.wy[2 * (li * n2 + j) + 1]};
  double wz[2] = {0.5, 0.5};
  if (d3) { wz[0] = M.wz[2 * (lk0 * n3 + k)]; wz[1] = M.wz[2 * (lk0 * n3 + k) + 1]; }
  const double sgn[2] = {-1.0, 1.0};
  const double dx0 = pb.dx[0], dx1 = (double)(1 << li) * pb.dx[1], dx2 = (double)(1 << lk0) * pb.dx[2];
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int ai = 0; ai <= 1; ai++) {
    const int ip = i + ai, lip = dend_lj(pb, ip);
    for (int b = 0; b <= 1; b++) {
      const int jp = rescale(j + b, li, lip);
      for (int qk = 0; qk < (d3 ? 2 : 1); qk++) {
        const int kp = d3 ? rescale(k + qk, lk0, dend_lk(pb, ip, jp)) : 0;
        const double ph = corner(ip, jp, kp);
        const double w2 = d3 ? wz[qk] : 1.0;
        d0 += ((ph * sgn[ai]) * wy[b]) * w2;
        d1 += ((ph * wx[ai]) * sgn[b]) * w2;
        if (d3) d2 += ((ph * wx[ai]) * wy[b]) * sgn[qk];
      }
    }
  }
  auto mflux = [&](int D, long long f, double area) -> double {
    const R x = ((const R *)a.F[D])[f];                 // slot 0: mass
    if constexpr (sizeof(R) == 8) return x;
    else return area * (double)x;
  };
  const double rho = (double)in.v[0][z];
  // radial <rho v>: the lower face, and the next shell's faces over this zone
  const double ain = rec3(pb, i, j)[MDS_ZS_AREA0];
  const double vx0 = mflux(0, fid3(pb, 0, i, j, k), ain) / (ain + 1.0e-16);
  const int lo = dend_lj(pb, i + 1);
  const int njp = li > lo ? 1 << (li - lo) : 1, jp0 = rescale(j, li, lo);
  double dA1 = 0.0, vx1 = 0.0;
  for (int jj = jp0; jj < jp0 + njp; jj++) {
    const int l = dend_lk(pb, i + 1, jj);
    const int nkp = l <= lk0 ? 1 << (lk0 - l) : 0;
    const double ar = rec3(pb, i + 1, jj)[MDS_ZS_AREA0];
    for (int kk = k * nkp; kk < k * nkp + nkp; kk++) {
      dA1 += ar + 1.0e-16;
      vx1 += mflux(0, fid3(pb, 0, i + 1, jj, kk), ar);
    }
  }
  vx1 /= dA1;
  const double vx = (vx0 * wx[0]) + (vx1 * wx[1]);
  double *S = M.src;
  S[0 * nz + z] = (rho * d0) / dx0;
  S[1 * nz + z] = (vx * d0) / dx0;
  // theta <rho v> with the phi level rules of its faces (gravity.c:1030-1058)
  const int lkp = dend_lk(pb, i, j + 1), lkm = dend_lk(pb, i, j - 1);
  const double a1 = rec3(pb, i, j)[MDS_ZS_AREA1], a1p = rec3(pb, i, j + 1)[MDS_ZS_AREA1];
  double A1 = 0.0, vy1 = 0.0, A0 = 0.0, vy0 = 0.0;
  if (lkp > lk0) {
    A1 += (0.5 * a1p) + 1.0e-16;
    vy1 += mflux(1, fid3(pb, 1, i, j + 1, k), a1p);
  } else if (lkp < lk0) {
    A1 += (a1p + a1p) + 1.0e-16;
    vy1 += mflux(1, fid3(pb, 1, i, j + 1, 2 * k), a1p) + mflux(1, fid3(pb, 1, i, j + 1, 2 * k + 1), a1p);
  } else {
    A1 += a1p + 1.0e-16;
    vy1 += mflux(1, fid3(pb, 1, i, j + 1, k), a1p);
  }
  vy1 /= A1;
  if (lkm < lk0) {
    A0 += a1 + 1.0e-16;
    vy0 += mflux(1, fid3(pb, 1, i, j, 2 * k), a1) + mflux(1, fid3(pb, 1, i, j, 2 * k + 1), a1);
  } else {
    A0 += a1 + 1.0e-16;
    vy0 += mflux(1, fid3(pb, 1, i, j, k), a1);
  }
  vy0 /= A0;
  const double vy = (vy0 * wy[0]) + (vy1 * wy[1]);
  S[2 * nz + z] = (rho * d1) / dx1;
  S[3 * nz + z] = (vy * d1) / dx1;
  if (d3) {
    const double a2 = rec3(pb, i, j)[MDS_ZS_AREA2];
    const double vz0 = mflux(2, fid3(pb, 2, i, j, k), a2) / (a2 + 1.0e-16);
    const double vz1 = mflux(2, fid3(pb, 2, i, j, k + 1), a2) / (a2 + 1.0e-16);
    const double vz = (vz0 * wz[0]) + (vz1 * wz[1]);
    S[4 * nz + z] = (rho * d2) / dx2;
    S[5 * nz + z] = (vz * d2) / dx2;
  }
}

// ---------------------------------------------------------------------------
// face fluxes in direction D (on dendritic grids, the faces whose 4-zone
// pencil stays on one level).  Each thread reconstructs one zone (the PPM of
// a zone serves its two faces); the face between zones z-1 and z takes the
// right value of z-1 from the neighbouring thread: the next lane for radial
// faces (a wavefront runs along i, 63 faces per 64 zones), the LDS row of the
// previous wavefront for theta and phi faces (a 64 x 8 block of zones in
// (i, j) or (i, k), 7 rows of faces).
// ---------------------------------------------------------------------------
template <class R, int D, int TAB, int DEND, int LO = 0>
__global__ __launch_bounds__(512, 4) void sph3_flux(Sph3Args a, StateArrays<R> in)
{
  const DevProb &pb = a.pb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  DevClock *ck = a.clock;
  const bool active = ck->active;
  const unsigned long long seen = min_key_seen(&ck->mindt);
  const double gam = pb.gam;
  // this thread's zone (zi, zj, zk) and the face (fi, fj, fk) = zone, whose
  // left neighbour is the zone of the previous lane (D = 0) or row (D > 0)
  int zi, zj, zk;
  bool has_face;
  if (D == 0) {
    const int nseg = (pb.n[0] + 1 + 62) / 63;
    const long long wv = (long long)blockIdx.x * (blockDim.x >> 6) + w;
    const int seg = (int)(wv % nseg);
    const long long rest = wv / nseg;
    zj = (int)(rest % pb.n[1]);
    zk = (int)(rest / pb.n[1]);
    zi = seg * 63 + lane - 1;
    has_face = lane > 0 && zi <= pb.n[0] && zk < pb.n[2];
  } else if (D == 1) {
    zi = blockIdx.x * 64 + lane;
    zj = blockIdx.y * 7 + w - 1;
    zk = blockIdx.z;
    has_face = w > 0 && zi < pb.n[0] && zj <= pb.n[1];
  } else {
    zi = blockIdx.x * 64 + lane;
    zj = blockIdx.y;
    zk = blockIdx.z * 7 + w - 1;
    has_face = w > 0 && zi < pb.n[0] && zk <= pb.n[2];
  }
  // Level values of this thread, read together up front (dendritic grids):
  // D = 0: lj and lk of shells zi-2 .. zi+1 in row zj; D = 1: lj of shell zi
  // and lk of its rows zj-2 .. zj+1; D = 2: lj of shell zi and lk of row zj.
  // Indices are clamped to the ghost-extended tables.  These values stand in
  // for every level lookup of the face and of the zone stencil below (the
  // tables reflect in their ghost rows, so a mirrored row has the level of
  // the row it mirrors).
  int Lj[4] = {0, 0, 0, 0}, Lk[4] = {0, 0, 0, 0};
  if (DEND) {
    auto ci_of = [&](int i) { return i < -MDS_NG ? -MDS_NG : (i > pb.n[0] + MDS_NG - 1 ? pb.n[0] + MDS_NG - 1 : i); };
    auto cj_of = [&](int j) { return j < -MDS_NG ? -MDS_NG : (j > pb.n[1] + MDS_NG - 1 ? pb.n[1] + MDS_NG - 1 : j); };
    if (D == 0) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int ci = ci_of(zi - 2 + t);
        Lj[t] = dend_lj(pb, ci);
        Lk[t] = dend_lk(pb, ci, cj_of(zj));
      }
    } else {
      const int ci = ci_of(zi);
      Lj[0] = dend_lj(pb, ci);
      if (D == 1) {
#pragma unroll
        for (int t = 0; t < 4; t++) Lk[t] = dend_lk(pb, ci, cj_of(zj - 2 + t));
      } else {
        Lk[0] = dend_lk(pb, ci, cj_of(zj));
      }
    }
  }
  // the face (zi, zj, zk) is done here if its 4-zone pencil stays on one level
  int lev = 0, dk = 1;
  bool live = has_face;
  if (!DEND) {
    live = live && zi <= pb.n[0] && zj <= pb.n[1] && zk <= pb.n[2];
  } else if (D == 0) {
    const int lj = Lj[2], lk = Lk[2];
    live = live && zj < (pb.n[1] >> lj) && zk < (pb.n[2] >> lk);
#pragma unroll
    for (int t = 0; t < 4; t++) live = live && Lj[t] == lj && Lk[t] == lk;
  } else if (D == 1) {
    const int lk = Lk[2];
    live = live && zj <= (pb.n[1] >> Lj[0]) && zk < (pb.n[2] >> lk);
#pragma unroll
    for (int t = 0; t < 4; t++) live = live && Lk[t] == lk;
    if (live) lev = Lj[0];
  } else {
    live = live && zj < (pb.n[1] >> Lj[0]) && zk <= (pb.n[2] >> Lk[0]);
    if (live) { lev = Lk[0]; dk = 1 << Lk[0]; }
  }
  // this zone exists (interior, or the one ghost zone at either end along D
  // that a face needs); other threads read nothing and own no face
  bool zvalid;
  if (D == 0) {
    zvalid = zi >= -1 && zi <= pb.n[0] && zj < pb.n[1] && zk < pb.n[2];
  } else {
    zvalid = zi < pb.n[0] && zj < pb.n[1] + (D == 1) && zk < pb.n[2] + (D == 2) && zj >= -1 && zk >= -1;
  }
  if (DEND && zvalid) {
    if (D == 0) {
      const int t = zi < 0 ? 3 : (zi >= pb.n[0] ? 1 : 2);       // shell clamped into the grid
      zvalid = zj < (pb.n[1] >> Lj[t]) && zk < (pb.n[2] >> Lk[t]);
    } else if (D == 1) {
      const int nj = pb.n[1] >> Lj[0];
      const int t = zj < 0 ? 3 : (zj >= nj ? 1 : 2);             // row clamped into the level
      zvalid = zj <= nj && zk < (pb.n[2] >> Lk[t]);
    } else {
      zvalid = zj < (pb.n[1] >> Lj[0]) && zk <= (pb.n[2] >> Lk[0]);
    }
  }
  // Positions that are no zone of their level fill most of the rows nearest
  // the poles: a wavefront (radial) or workgroup (theta, phi: they meet at a
  // barrier below) without a zone has no face and leaves here.
  if (DEND) {
    if (D == 0) {
      if (__ballot(zvalid) == 0) return;
    } else {
      if (!__syncthreads_or(zvalid)) return;
    }
  }
  // level of the zone's own stencil (for the reconstruction)
  int zlev = 0;
  if (DEND && D == 1) zlev = Lj[0];
  if (DEND && D == 2 && zvalid) zlev = Lk[0];
  const int bi = zvalid ? zi : 0, bj = zvalid ? zj : 0, bk = zvalid ? zk : 0;
  double q3[3][8];
  // the zone's stencil; with PROB/EXTRAP ghost layers at the outer radius
  // (radial faces) a second copy of the loop reads them, so that the plain
  // loop stays one block
  auto stencil = [&](auto slab) {
#pragma unroll
    for (int s = 0; s < 3; s++) {
      int i = bi, j = bj, k = bk;
      int n = pb.n[0], nkm = 1;
      if (D == 0) {
        i += s - 1;
      } else if (D == 1) {
        j += s - 1;
        n = pb.n[1] >> Lj[0];
        nkm = pb.n[2] >> Lk[s + 1];
      } else {
        k += s - 1;
        n = pb.n[2] >> Lk[0];
      }
      double sgn[3];
      map_zone_n(pb, D, i, j, k, sgn, n, nkm);
      if (decltype(slab)::value && i >= pb.n[0]) slab_state<R>(pb, i, j, k, TAB, q3[s]);
      else zone_state<R>(a, in, zid3(pb, i, j, k), TAB, gam, q3[s]);
      for (int d = 0; d < 3; d++) q3[s][2 + d] *= sgn[d];
    }
  };
  if (zvalid) {
    if (D == 0 && pb.has_slab) stencil(std::true_type{});
    else stencil(std::false_type{});
  } else {
    for (int s = 0; s < 3; s++)
      for (int v = 0; v < 8; v++) q3[s][v] = 1.0;
  }
  // the face's geometry and transverse-shock switch, read with the stencil
  const int fi = live ? zi : 0, fj = live ? zj : 0, fk = live ? zk : 0;
  const ZRec rc = rec3(pb, fi, fj);
  const FaceGeo<double> fg = face_geo<D>(rc);
  bool hlle = false;
  if (D > 0) {
    // transverse_shock (fluid.c:189-210) on the zones either side of the face
    const int nj = pb.n[1] >> Lj[0];
    if (D == 1)
      hlle = tagged_n(a, fi, fj, fk, nj, pb.n[2] >> Lk[2]) | tagged_n(a, fi, fj - 1, fk, nj, pb.n[2] >> Lk[1]);
    else
      hlle = tagged_n(a, fi, fj, fk, nj, pb.n[2] >> Lk[0]) | tagged_n(a, fi, fj, fk - 1, nj, pb.n[2] >> Lk[0]);
  }
  const int zd = D == 0 ? bi : (D == 1 ? bj : bk);
  // polar-axis rows (last theta row): first-order states, exchanged as zone values
  const bool polar = (D == 2) && zvalid && (bj == (pb.n[1] >> Lj[0]) - 1);
  R al[8], ar[8];
  if (polar) {
    for (int v = 0; v < 8; v++) al[v] = ar[v] = (R)q3[1][v];
  } else {
    zone_recon<R, D, TAB, LO>(pb, zlev, zd, q3, al, ar);
  }
  // the left state of this thread's face
  R Lq[8];
  if constexpr (D == 0) {
    for (int v = 0; v < 8; v++) Lq[v] = __shfl_up(ar[v], 1, 64);
  } else {
    __shared__ R sh[8][8][64];
    for (int v = 0; v < 8; v++) sh[v][w][lane] = ar[v];
    __syncthreads();
    for (int v = 0; v < 8; v++) Lq[v] = w > 0 ? sh[v][w - 1][lane] : ar[v];
  }
  if (polar && live) {
    // step.c:1299-1397: velocities of zones k-1 and k through Cartesian components
    double vl[3] = {(double)Lq[2], (double)Lq[3], (double)Lq[4]}, vr[3] = {(double)al[2], (double)al[3], (double)al[4]};
    polar_rotate(pb, rc, fk, dk, vl, vr);
    for (int d = 0; d < 3; d++) { Lq[2 + d] = (R)vl[d]; al[2 + d] = (R)vr[d]; }
  }
  double sig, F[6];
  const int isshock = face_riemann<R, D, TAB>(Lq, al, fg, hlle, F, sig);
  if (live && active) {
    store_flux<R>(a, D, fid3(pb, D, fi, fj, fk), F, TAB ? 6 : 5);
    if (D == 0 && isshock) {
      if (DEND) tag_radial(a, fi, fj, fk);
      else
        for (int s = 0; s < 3; s++) {
          const int ii = fi - s;
          if (ii >= 0 && ii < pb.n[0] && fj < pb.n[1] && fk < pb.n[2]) a.flags[zid3(pb, ii, fj, fk)] = 0;
        }
    }
  }
  if (DEND) {
    // DJS(i) dx1 / |s| and DKS(i,j) dx2 / |s| (step.c:1263, 1419)
    const double width = (double)(D == 0 ? 1 : (D == 1 ? 1 << lev : dk)) * pb.dx[D];
    face_dt(a, seen, live && active, width, sig, isshock, hlle, D);
    return;
  }
  // CFL: min dx/|s| over faces = dx / max |s| (one division per wave)
  double smax = (live && active) ? fabs(sig) : 0.0;
  const double wmax = wave_max(smax);
  const unsigned long long nsh = __popcll(__ballot(live && active && D == 0 && isshock));
  const unsigned long long nhl = __popcll(__ballot(live && active && D > 0 && hlle));
  if (lane == 0) {
    if (wmax > 0.0) fold_min_key(&ck->mindt, seen, pb.dx[D] / wmax);
    if (nsh) atomicAdd(&a.counters[0], nsh);
    if (nhl) atomicAdd(&a.counters[1], nhl);
  }
}

// ---------------------------------------------------------------------------
// dendritic faces whose pencil crosses a level change: radial (D = 0) and
// theta (D = 1) faces from a list of (i, j, k, 0), k the face's phi index
// ---------------------------------------------------------------------------
// A workgroup takes 64 listed faces.  Its four wavefronts build one pencil
// entry each (entry s of face l in lane l of wavefront s): the entries are
// independent, and each is a chain of dependent reads (levels, then the zones
// of a restriction or a limited-slope prolongation).  The entries meet in LDS
// and the first wavefront solves the faces.
template <class R, int D, int TAB, int LO = 0>
__global__ __launch_bounds__(256) void sph3_flux_irr(Sph3Args a, StateArrays<R> in, const int4 *list, long long n)
{
  const DevProb &pb = a.pb;
  const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
  const long long t = (long long)blockIdx.x * 64 + lane;
  const bool live = t < n;
  const int4 e = list[live ? t : 0];
  const int fi = e.x, fj = e.y, fk = e.z;
  __shared__ double sq[4][8][64];
  {
    double qs[8];
    if (D == 0) {
      radial_entry<R, TAB>(a, in, fi, fj, fk, fi - 2 + s, qs);
    } else {
      const int lj = dend_lk(pb, fi, fj), lm = dend_lk(pb, fi, fj - 1);
      theta_entry<R, TAB>(a, in, fi, lj < lm ? lj : lm, fk, fj - 2 + s, qs);
    }
#pragma unroll
    for (int v = 0; v < 8; v++) sq[s][v][lane] = qs[v];
  }
  __syncthreads();
  if (s != 0) return;                                  // uniform per wavefront
  const bool active = a.clock->active;
  const unsigned long long seen = min_key_seen(&a.clock->mindt);
  double q[4][8];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 8; v++) q[u][v] = sq[u][v][lane];
  int lev = 0;
  bool hlle = false;
  if (D == 1) {
    const int lj = dend_lk(pb, fi, fj), lm = dend_lk(pb, fi, fj - 1);
    const int lf = lj < lm ? lj : lm;
    lev = dend_lj(pb, fi);
    // the zone of row j owning this sub-face, and transverse_shock (fluid.c:189-201)
    const int kz = fk >> (lj - lf);
    hlle = tagged(a, fi, fj, kz);
    if (lj >= lm)
      for (int kk = kz << (lj - lm); kk < (kz + 1) << (lj - lm); kk++) hlle = hlle || tagged(a, fi, fj - 1, kk);
  }
  const ZRec rc = rec3(pb, fi, fj);
  double sig, F[6];
  const int isshock = face_flux<R, D, TAB, 1, LO>(pb, lev, D == 0 ? fi : fj, q, false, fk, 1, rc, hlle, F, sig);
  if (live && active) {
    store_flux<R>(a, D, fid3(pb, D, fi, fj, fk), F, TAB ? 6 : 5);
    if (D == 0 && isshock) tag_radial(a, fi, fj, fk);
  }
  face_dt(a, seen, live && active, (double)(1 << lev) * pb.dx[D], sig, isshock, hlle, D);
}

// ---------------------------------------------------------------------------
// zone update: sources, flux divergence, RK combination, u -> p, dt limits
// ---------------------------------------------------------------------------
// Zones to threads, by the workgroup shape (MDS_TILE_UPD3S): 0, 256
// consecutive zones in storage order; 1, 64 consecutive i at four theta rows
// j; 2, 64 i at two rows and two phi positions; 3, 64 i at four phi
// positions k (shape 1 on grids with fewer than four phi zones).  In shapes
// 1-3 the wavefronts of a workgroup read neighbouring records and faces
// (the same (i, j) records and a phi face twice in shape 3), which then come
// from the compute unit's cache instead of once per zone from memory.
MDS_DEV void update_zone_of(const DevProb &pb, int shape, int &i, int &j, int &k, bool &inbox)
{
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  if (shape == 0) {
    const long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    inbox = z < (long long)n1 * n2 * n3;
    const long long zl = inbox ? z : 0;
    i = (int)(zl % n1);
    j = (int)((zl / n1) % n2);
    k = (int)(zl / ((long long)n1 * n2));
    return;
  }
  if (shape == 3 && n3 < 4) shape = 1;
  if (shape == 2 && n3 < 2) shape = 1;
  const int nti = (n1 + 63) >> 6, w = threadIdx.x >> 6;
  const int b = blockIdx.x, ti = b % nti, r = b / nti;
  i = (ti << 6) + (threadIdx.x & 63);
  if (shape == 3) {
    j = r % n2;
    k = ((r / n2) << 2) + w;
  } else if (shape == 2) {
    const int njg = (n2 + 1) >> 1;
    j = ((r % njg) << 1) + (w & 1);
    k = ((r / njg) << 1) + (w >> 1);
  } else {
    const int njg = (n2 + 3) >> 2;
    j = ((r % njg) << 2) + w;
    k = r / njg;
  }
  inbox = i < n1 && j < n2 && k < n3;
}

// workgroups of the update's grid for a shape (the same rules)
static unsigned update_blocks(const DevProb &pb, int shape)
{
  const long long n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2], nti = (n1 + 63) / 64;
  if (shape == 0) return (unsigned)((n1 * n2 * n3 + 255) / 256);
  if ((shape == 3 && n3 < 4) || (shape == 2 && n3 < 2)) shape = 1;
  if (shape == 3) return (unsigned)(nti * n2 * ((n3 + 3) / 4));
  if (shape == 2) return (unsigned)(nti * ((n2 + 1) / 2) * ((n3 + 1) / 2));
  return (unsigned)(nti * ((n2 + 3) / 4) * n3);
}

template <class R, int TAB, int MONO, int DEND>
__global__ __launch_bounds__(256, 3) void sph3_update(Sph3Args a, StateArrays<R> sp, StateArrays<R> sph, int stage)
{
  const DevProb &pb = a.pb;
  const int n1 = pb.n[0], n2 = pb.n[1], n3 = pb.n[2];
  int iz, jz, kz;
  bool inbox;
  update_zone_of(pb, a.upd, iz, jz, kz, inbox);
  const long long zl = inbox ? zid3(pb, iz, jz, kz) : 0;
  const int i0 = inbox ? iz : 0, j0 = inbox ? jz : 0, k0 = inbox ? kz : 0;
  // level relations with the neighbours (dendritic grids; all 0 otherwise),
  // read first from the small level tables; lq is the outer shell's level of
  // row j, its row when li == lo
  int li = 0, lo = 0, lk0 = 0, lkp = 0, lkm = 0, lq = 0;
  if (DEND) {
    const int N1 = n1 + 2 * MDS_NG;
    const unsigned char *lkr = pb.lk + (size_t)(j0 + MDS_NG) * N1 + (i0 + MDS_NG);
    li = pb.lj[i0 + MDS_NG];
    lo = pb.lj[i0 + 1 + MDS_NG];
    lk0 = lkr[0];
    lkp = lkr[N1];
    lkm = lkr[-N1];
    lq = lkr[1];
  }
  const bool live = inbox && (!DEND || (j0 < (n2 >> li) && k0 < (n3 >> lk0)));
  if (DEND && __ballot(live) == 0) return;          // a wavefront without a zone
  // Every other read of a zone whose neighbours share its levels (every zone
  // of a uniform grid) goes out in one batch, with no branch between the
  // reads: the state of both stages, the record and the six faces.  A box
  // position that is no zone of its level reads zone (0, 0, 0) instead, whose
  // lines all such positions share.  A zone at a level change then redoes its
  // flux sum with the reference's fan-in.
  const int i = live ? i0 : 0, j = live ? j0 : 0, k = live ? k0 : 0;
  const long long zr = live ? zl : 0;
  DevClock *ck = a.clock;
  const bool active = ck->active;
  const unsigned long long seen = min_key_seen(&ck->mindt);
  const double dt = ck->dt;
  const double gam = pb.gam;
  const StateArrays<R> &in = stage == 0 ? sp : sph;
  double q[8], qs[8];
  zone_state<R>(a, in, zr, TAB, gam, q);
  zone_state<R>(a, sp, zr, TAB, gam, qs);          // stage 1 needs it (the same data at stage 0)
  const double cs_tab = TAB ? (double)((const R *)a.CS)[zr] : 0.0;
  const ZRec rc = rec3(pb, i, j);
  double gcov[3], gcon[3], scl[3], conn[11];
  for (int d = 0; d < 3; d++) {
    gcov[d] = rc[MDS_ZS_GCOV0 + d];
    gcon[d] = rc[MDS_ZS_GCON + d];
    scl[d] = rc[MDS_ZS_SCALE0 + d];
  }
  {
    const int cf[11] = {MDS_ZS_C000, MDS_ZS_C101, MDS_ZS_C202, MDS_ZS_C011, MDS_ZS_C110, MDS_ZS_C111,
                        MDS_ZS_C212, MDS_ZS_C022, MDS_ZS_C122, MDS_ZS_C220, MDS_ZS_C221};
#pragma unroll
    for (int t = 0; t < 11; t++) conn[t] = rc[cf[t]];
  }
  const double vol = rc[MDS_ZS_VOLUME];
  const long long f0 = fid3(pb, 0, i, j, k);
  const long long f1 = fid3(pb, 1, i, j, k), f2 = fid3(pb, 2, i, j, k), f2p = fid3(pb, 2, i, j, k + 1);
  const long long fo = fid3(pb, 0, i + 1, j, k), ft = fid3(pb, 1, i, j + 1, k);
  // Face fluxes: FP64 ones carry their area; FP32 ones are per unit area and
  // take the FP64 area of the record the flux kernel used, (fi, fj) of the
  // face, here forming the same product that kernel would have stored.
  // The slot base is uniform and the face's byte offset is 32-bit (one slot's
  // array stays below 4 GiB, checked at set-up), so a load needs one VGPR of
  // address, shared by the slots of a face.
  auto load = [&](int D, int v, long long f) -> R {
    const char *base = (const char *)a.F[D] + (size_t)v * (size_t)a.nf[D] * sizeof(R);
    return *(const R *)(base + (unsigned)f * (unsigned)sizeof(R));
  };
  auto scaled = [&](R x, double area) -> double {
    if constexpr (sizeof(R) == 8) return x;
    else return area * (double)x;
  };
  auto flux = [&](int D, int v, long long f, double area) -> double { return scaled(load(D, v, f), area); };
  const double ain = rc[MDS_ZS_AREA0], a1 = rc[MDS_ZS_AREA1], a2 = rc[MDS_ZS_AREA2];
  const double a1p = sizeof(R) == 8 ? 0.0 : rec3(pb, i, j + 1)[MDS_ZS_AREA1];
  const double ao = (sizeof(R) == 4 || MONO) ? rec3(pb, i + 1, j)[MDS_ZS_AREA0] : 0.0;
  constexpr int NSLOT = TAB ? 6 : 5;
  R x[NSLOT][6];
#pragma unroll
  for (int v = 0; v < NSLOT; v++) {
    x[v][0] = load(0, v, fo);
    x[v][1] = load(0, v, f0);
    x[v][2] = load(1, v, ft);
    x[v][3] = load(1, v, f1);
    x[v][4] = load(2, v, f2p);
    x[v][5] = load(2, v, f2);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (stage == 0)
    for (int t = 0; t < 8; t++) qs[t] = q[t];
  const double *qn = qs;
  // geometric sources (rad_fluid.c:195-214): the 11 non-zero conn[l][m][n]
  // in the reference's (l, n) order per m
  double src[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double press = q[6];
  {
    double vc[3], T[3][3];
    for (int d = 0; d < 3; d++) vc[d] = gcov[d] * q[2 + d];
    for (int l = 0; l < 3; l++) {
      for (int m = 0; m < 3; m++) T[l][m] = (q[0] * q[2 + m]) * vc[l];
      T[l][l] += press;
    }
    src[2] += T[0][0] * conn[0];
    src[2] += T[1][1] * conn[1];
    src[2] += T[2][2] * conn[2];
    src[3] += T[0][1] * conn[3];
    src[3] += T[1][0] * conn[4];
    src[3] += T[1][1] * conn[5];
    src[3] += T[2][2] * conn[6];
    src[4] += T[0][2] * conn[7];
    src[4] += T[1][2] * conn[8];
    src[4] += T[2][0] * conn[9];
    src[4] += T[2][1] * conn[10];
  }
  // flux divergence per slot, in the reference's order (step.c:398-438):
  // outer radial faces, inner radial, outer theta, inner theta, phi
  double divF[NSLOT];
  double vr = 0.0, varea = 0.0;                 // outer mass flux and area (monopole)
#pragma unroll
  for (int v = 0; v < NSLOT; v++) {
    const double fro = scaled(x[v][0], ao);
    double d = 0.0;
    d += fro;
    d -= scaled(x[v][1], ain);
    d += scaled(x[v][2], a1p);
    d -= scaled(x[v][3], a1);
    d += scaled(x[v][4], a2) - scaled(x[v][5], a2);
    divF[v] = d;
    if (v == 0) { vr += fro; varea += ao + 1.0e-16; }
  }
  const bool simple = !DEND || (li == lo && lq == lk0 && lkp == lk0 && lkm == lk0);
  if (DEND && live && !simple) {
    // a zone at a level change: the finer outer radial faces (step.c:398-408)
    // and the half-area theta rules (:411-433)
    const int njp = li > lo ? 1 << (li - lo) : 1;
    const int jp0 = rescale(j, li, lo);
    vr = 0.0;
    varea = 0.0;
#pragma unroll
    for (int v = 0; v < NSLOT; v++) divF[v] = 0.0;
    for (int jj = jp0; jj < jp0 + njp; jj++) {
      const int l = dend_lk(pb, i + 1, jj);
      const int nkp = l <= lk0 ? 1 << (lk0 - l) : 0;
      const double ar = rec3(pb, i + 1, jj)[MDS_ZS_AREA0];
      for (int kk = k * nkp; kk < k * nkp + nkp; kk++) {
        const long long f = fid3(pb, 0, i + 1, jj, kk);
#pragma unroll
        for (int v = 0; v < NSLOT; v++) {
          const double fro = flux(0, v, f, ar);
          divF[v] += fro;
          if (v == 0) { vr += fro; varea += ar + 1.0e-16; }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NSLOT; v++) divF[v] -= scaled(x[v][1], ain);
    // outer theta face (step.c:411-423)
    if (lkp > lk0) {
#pragma unroll
      for (int v = 0; v < NSLOT; v++) divF[v] += 0.5 * scaled(x[v][2], a1p);
    } else if (lkp < lk0) {
      const long long fa = fid3(pb, 1, i, j + 1, 2 * k), fb = fid3(pb, 1, i, j + 1, 2 * k + 1);
#pragma unroll
      for (int v = 0; v < NSLOT; v++) divF[v] += flux(1, v, fa, a1p) + flux(1, v, fb, a1p);
    } else {
#pragma unroll
      for (int v = 0; v < NSLOT; v++) divF[v] += scaled(x[v][2], a1p);
    }
    // inner theta face (step.c:425-433)
    if (lkm < lk0) {
      const long long fa = fid3(pb, 1, i, j, 2 * k), fb = fid3(pb, 1, i, j, 2 * k + 1);
#pragma unroll
      for (int v = 0; v < NSLOT; v++) divF[v] -= 0.5 * flux(1, v, fa, a1) + 0.5 * flux(1, v, fb, a1);
    } else {
#pragma unroll
      for (int v = 0; v < NSLOT; v++) divF[v] -= scaled(x[v][3], a1);
    }
#pragma unroll
    for (int v = 0; v < NSLOT; v++) divF[v] += scaled(x[v][4], a2) - scaled(x[v][5], a2);
  }
  if (MONO) {
    // gravity.c:415-437: mass flux per area through the finer outer faces
    const double gz = a.grav[i];
    const double xl = pb.startx[0] + i * pb.dx[0];
    vr /= varea;
    src[2] += q[0] * gz;
    const double vl = scaled(x[0][1], ain) / (ain + 1.0e-16);
    const double A = (vr - vl) / pb.dx[0];
    const double B = vl - A * xl;
    const double be = pb.mom[0][3 * (i + MDS_NG) + 1], ga = pb.mom[0][3 * (i + MDS_NG) + 2];
    src[1] += ((A * be) / ga + B) * gz;
  }
  if (MONO == 2) {
    // multipole gravity after the monopole (gravity.c:1162-1165, 985-1103)
    const long long nz = (long long)n1 * n2 * n3;
    src[2] -= a.mpsrc[0 * nz + zr];
    src[1] -= a.mpsrc[1 * nz + zr];
    src[3] -= a.mpsrc[2 * nz + zr];
    src[1] -= a.mpsrc[3 * nz + zr];
    if (pb.ndim == 3) {
      src[4] -= a.mpsrc[4 * nz + zr];
      src[1] -= a.mpsrc[5 * nz + zr];
    }
  }
  double U[6], Uh[6];
  p_to_u_x64<3, false>(qn[0], qn[1], &qn[2], gcov, U);
  U[MDS_YI] = TAB ? U[0] * qn[5] : 0.0;
  // FP32-bulk stage 0 combines 1 U^n + 0 U* with U* = U^n itself: no second
  // conversion (the FP64 path keeps the reference's combination as written)
  const bool short0 = sizeof(R) == 4 && stage == 0;
  if (!short0) {
    p_to_u_x64<3, false>(q[0], q[1], &q[2], gcov, Uh);
    Uh[MDS_YI] = TAB ? Uh[0] * q[5] : 0.0;
  }
  const double lim_src = dt_source_control_x64<3>(qn[1], &qn[2], U, src);
  const double pf = stage == 0 ? 1.0 : 0.5, phf = stage == 0 ? 0.0 : 0.5, ff = stage == 0 ? 1.0 : 0.5;
  if (short0) {
#pragma unroll
    for (int v = 0; v < NSLOT; v++) U[v] = U[v] - dt * ((divF[v] / vol) - src[v]);
  } else {
#pragma unroll
    for (int v = 0; v < NSLOT; v++) {
      const double ddt = (divF[v] / vol) - src[v];
      U[v] = ((pf * U[v]) + (phf * Uh[v])) - ((ff * dt) * ddt);
    }
  }
  double ro, uo, vo[3];
  u_to_p_x64<3, false>(U, gcon, ro, uo, vo);
  double locmax = 0.0;
  if (stage == 1) {
    // estimate_dt zone bound (step.c:177-187) with the stage's sound speed
    double cs = TAB ? cs_tab : sqrt((gam * press) / q[0]);
    double loc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (d >= pb.ndim) break;                 // NDIM directions (2D grids have no phi term)
      double ldx = pb.dx[d];
      if (DEND && d == 1) ldx *= (double)(1 << li);
      if (DEND && d == 2) ldx *= (double)(1 << lk0);
      loc += (fabs(vo[d]) + (cs / scl[d])) / ldx;
    }
    locmax = loc;
  }
  if (live && active) {
    StateArrays<R> &out = stage == 0 ? sph : sp;
    out.v[0][zl] = (R)ro;
    out.v[1][zl] = (R)uo;
    for (int d = 0; d < 3; d++) out.v[2 + d][zl] = (R)vo[d];
    if (TAB) out.v[5][zl] = (R)(U[MDS_YI] / U[0]);
  }
  const double wl = wave_min((live && active) ? lim_src : 1.0e300);
  const double wm = wave_max((live && active) ? locmax : 0.0);
  const unsigned nbad = bad_state_lanes(live && active, ro, uo, !TAB);
  if ((threadIdx.x & 63) == 0) {
    double l = wl;
    if (wm > 0.0) l = fmin(l, 1.0 / wm);
    if (l < 1.0e300) fold_min_key(&ck->mindt, seen, l);
    if (nbad) atomicAdd(&ck->nbad, (unsigned long long)nbad);
  }
}

// ---------------------------------------------------------------------------
// launcher: one RK stage
// ---------------------------------------------------------------------------
// One RK stage.  Work that does not depend on the kernels beside it runs on
// a second stream (L.side), joined through events:
//   main  [e0] reset_flags, eos [e1] flux0 [e3] {e2} flux1, flux2 {e4} update
//   side  {e0} shell_part, shell_field {e1} flux_irr0 [e2] {e3} flux_irr1 [e4]
// ([e] records, {e} waits).  The shell masses read only rho; the listed
// level-change faces need the EOS arrays, and the theta ones the radial shock
// tags of both radial kernels; the update needs everything.  Both streams
// have the default priority.
template <class R, int DEND>
static hipError_t launch_stage(const Sph3Launch &L, const Sph3Args &a, const StateArrays<R> &sp,
                               const StateArrays<R> &sph, int stage, hipStream_t st)
{
  const DevProb &pb = a.pb;
  const StateArrays<R> &in = stage == 0 ? sp : sph;
  const long long nz = (long long)pb.n[0] * pb.n[1] * pb.n[2];
  const dim3 blk(256);
  const bool irr = DEND && (L.nirr[0] > 0 || L.nirr[1] > 0);
  const bool fork = L.mono || irr;
  hipStream_t s2 = fork ? L.side : st;
  hipError_t e = hipSuccess;
  auto rec = [&](int k, hipStream_t s) { if (fork && e == hipSuccess) e = hipEventRecord(L.ev[k], s); };
  auto wait = [&](hipStream_t s, int k) { if (fork && e == hipSuccess) e = hipStreamWaitEvent(s, L.ev[k], 0); };
  rec(0, st);
  wait(s2, 0);
  if (stage == 0)
    hipLaunchKernelGGL(sph3_reset_flags, dim3((unsigned)((nz + 255) / 256)), blk, 0, st, L.flags, nz);
  if (L.tab) hipLaunchKernelGGL(sph3_eos_kernel<R>, dim3((unsigned)((nz + 255) / 256)), blk, 0, st, a, in);
  const dim3 pgrid((unsigned)((pb.n[0] + 63) / 64), (unsigned)L.npart);
  if (L.mono == MDS_MONOPOLE_GR) {
    // the pressure sums read the table EOS arrays of this stage: after the
    // EOS kernel, on its stream
    hipStream_t sg = L.tab ? st : s2;
    if (L.tab) hipLaunchKernelGGL((sph3_shell_part_gr<R, 1>), pgrid, blk, 0, sg, a, in, L.part, L.npart);
    else hipLaunchKernelGGL((sph3_shell_part_gr<R, 0>), pgrid, blk, 0, sg, a, in, L.part, L.npart);
    hipLaunchKernelGGL(sph3_shell_field_gr, dim3(1), dim3(1024), 0, sg, a, L.part, L.npart, L.grav);
  } else if (L.mono) {
    hipLaunchKernelGGL(sph3_shell_part<R>, pgrid, blk, 0, s2, a, in, L.part);
    hipLaunchKernelGGL(sph3_shell_field, dim3(1), dim3(1024), 0, s2, a, L.part, L.npart, L.ms, L.grav);
  }
  rec(1, st);
  wait(s2, 1);
  const int nseg = (pb.n[0] + 1 + 62) / 63;
  const dim3 fgrid[3] = {dim3((unsigned)(((long long)nseg * pb.n[1] * pb.n[2] + 7) / 8)),
                         dim3((unsigned)((pb.n[0] + 63) / 64), (unsigned)((pb.n[1] + 1 + 6) / 7), (unsigned)pb.n[2]),
                         dim3((unsigned)((pb.n[0] + 63) / 64), (unsigned)pb.n[1], (unsigned)((pb.n[2] + 1 + 6) / 7))};
  const dim3 fblk(512);
  // hydro_interp_order 2 or 1 (FP64 only) runs its own instances, so the
  // para_recon kernels keep their register budget
  const bool lo = sizeof(R) == 8 && pb.order != 3;
  constexpr int LOW_R = sizeof(R) == 8 ? 1 : 0;     // FP32-bulk runs order 3 only
#define MDS_FLUX(D)                                                                                        \
  {                                                                                                        \
    if (lo) {                                                                                              \
      if (L.tab) hipLaunchKernelGGL((sph3_flux<R, D, 1, DEND, LOW_R>), fgrid[D], fblk, 0, st, a, in);     \
      else hipLaunchKernelGGL((sph3_flux<R, D, 0, DEND, LOW_R>), fgrid[D], fblk, 0, st, a, in);           \
    } else {                                                                                               \
      if (L.tab) hipLaunchKernelGGL((sph3_flux<R, D, 1, DEND>), fgrid[D], fblk, 0, st, a, in);            \
      else hipLaunchKernelGGL((sph3_flux<R, D, 0, DEND>), fgrid[D], fblk, 0, st, a, in);                  \
    }                                                                                                      \
  }
#define MDS_FLUX_IRR(D)                                                                                    \
  if (DEND && L.nirr[D] > 0) {                                                                             \
    const dim3 g((unsigned)((L.nirr[D] + 63) / 64));                                                      \
    if (lo) {                                                                                              \
      if (L.tab) hipLaunchKernelGGL((sph3_flux_irr<R, D, 1, LOW_R>), g, blk, 0, s2, a, in, L.irr[D], L.nirr[D]); \
      else hipLaunchKernelGGL((sph3_flux_irr<R, D, 0, LOW_R>), g, blk, 0, s2, a, in, L.irr[D], L.nirr[D]);       \
    } else {                                                                                               \
      if (L.tab) hipLaunchKernelGGL((sph3_flux_irr<R, D, 1>), g, blk, 0, s2, a, in, L.irr[D], L.nirr[D]);   \
      else hipLaunchKernelGGL((sph3_flux_irr<R, D, 0>), g, blk, 0, s2, a, in, L.irr[D], L.nirr[D]);         \
    }                                                                                                      \
  }
  MDS_FLUX_IRR(0)
  rec(2, s2);
  MDS_FLUX(0)
  rec(3, st);
  wait(st, 2);
  wait(s2, 3);
  MDS_FLUX_IRR(1)
  rec(4, s2);
  MDS_FLUX(1)
  if (pb.ndim == 3) MDS_FLUX(2)     // 2D grids have no phi faces (their flux slots stay zero)
  wait(st, 4);
#undef MDS_FLUX
#undef MDS_FLUX_IRR
  const dim3 grd((unsigned)((nz + 255) / 256));
  const dim3 gupd(update_blocks(pb, a.upd));
  if (L.mpl) {
    // multipole gravity: moments of the stage state, the corner potential and
    // the sources, after this stage's mass fluxes (all directions) are known
    const MpDev &M = L.mp;
    hipLaunchKernelGGL(sph3_mp_part<R>, dim3((unsigned)((pb.n[0] + 63) / 64), (unsigned)M.npart, (unsigned)M.nlist),
                       blk, 0, st, a, M, in);
    hipLaunchKernelGGL(sph3_mp_sum, dim3((unsigned)((pb.n[0] + 255) / 256), (unsigned)M.nlist), blk, 0, st, a, M);
    hipLaunchKernelGGL(sph3_mp_scan, dim3((unsigned)((M.nlist + 63) / 64)), dim3(64), 0, st, a, M);
    const long long nc = (long long)(pb.n[0] + 1) * (pb.n[1] + 1) * (pb.n[2] + 1);
    hipLaunchKernelGGL(sph3_mp_phi, dim3((unsigned)((nc + 255) / 256)), blk, 0, st, a, M);
    hipLaunchKernelGGL(sph3_mp_src<R>, grd, blk, 0, st, a, M, in);
    if (L.tab) hipLaunchKernelGGL((sph3_update<R, 1, 2, DEND>), gupd, blk, 0, st, a, sp, sph, stage);
    else hipLaunchKernelGGL((sph3_update<R, 0, 2, DEND>), gupd, blk, 0, st, a, sp, sph, stage);
  } else if (L.tab && L.mono) hipLaunchKernelGGL((sph3_update<R, 1, 1, DEND>), gupd, blk, 0, st, a, sp, sph, stage);
  else if (L.tab) hipLaunchKernelGGL((sph3_update<R, 1, 0, DEND>), gupd, blk, 0, st, a, sp, sph, stage);
  else if (L.mono) hipLaunchKernelGGL((sph3_update<R, 0, 1, DEND>), gupd, blk, 0, st, a, sp, sph, stage);
  else hipLaunchKernelGGL((sph3_update<R, 0, 0, DEND>), gupd, blk, 0, st, a, sp, sph, stage);
  return e;
}

template <class R>
static hipError_t stage_impl(const Sph3Launch &L, const DevProb &pb, DevClock *clock, unsigned long long *counters,
                             const DevEos &eos, const StateArrays<R> &sp, const StateArrays<R> &sph, int stage,
                             hipStream_t st)
{
  Sph3Args a;
  a.pb = pb;
  a.clock = clock;
  a.counters = counters;
  a.flags = L.flags;
  a.eos = eos;
  for (int d = 0; d < 3; d++) { a.F[d] = L.F[d]; a.nf[d] = L.nf[d]; }
  a.P = L.P; a.G = L.G; a.CS = L.CS;
  a.tg = L.tg;
  a.grav = L.grav;
  a.mpsrc = L.mpl ? L.mp.src : nullptr;
  a.upd = L.upd;
  const hipError_t e = L.dend ? launch_stage<R, 1>(L, a, sp, sph, stage, st) : launch_stage<R, 0>(L, a, sp, sph, stage, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

extern "C" hipError_t mds_launch_sph3d(const Sph3Launch *L, const DevProb *pb, DevClock *clock,
                                       unsigned long long *counters, const DevEos *eos, const void *sp,
                                       const void *sph, int stage, hipStream_t st)
{
  if (L->fp32)
    return stage_impl<float>(*L, *pb, clock, counters, *eos, *(const StateArrays<float> *)sp,
                             *(const StateArrays<float> *)sph, stage, st);
  return stage_impl<double>(*L, *pb, clock, counters, *eos, *(const StateArrays<double> *)sp,
                            *(const StateArrays<double> *)sph, stage, st);
}
