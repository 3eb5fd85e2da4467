// mds_internal.h -- device-side descriptors shared by the sweep kernels and
// the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "../../../include/mds_hydro.h"

// Compact per-zone geometry of a spherical column for the FP32-bulk path:
// the FP64 factors the update needs (inverse volume, lower-face area, zone
// metric, the combined pressure coefficient of the connection terms) and the
// FP32 factors of the Riemann solve and the gravity source.  gcon[0] =
// 1 / gcov0 and scale[0][0] = sqrt(gcov0) are recomputed where needed.  The
// reconstruction coefficients live beside it (DevProb::rc, or one set per
// problem, DevProb::rcu).  48 bytes, read with three 16-byte loads.
struct __align__(16) ZoneF32 {
  double ivol;             // 1 / zone_geom.volume
  double area;             // area[0] of the lower radial face
  double gcov0;            // zone-averaged gcov[0][0]
  double cp;               // conn[0][0][0] + conn[1][0][1] + conn[2][0][2]
  float c000;              // conn[0][0][0]
  float scalef, gcovf0;    // lower face: scale[1][0], gcov[1][0]
  float dcen;              // volume centroid beta/Gamma minus the lower face x
};
static_assert(sizeof(ZoneF32) == 48, "ZoneF32 layout");

// Tabulated EOS as the kernels see it (records of mds_eos_table in HBM).
struct DevEos {
  int nr, nt, ny;
  double lr0, lr1, lt0, lt1, ye0, ye1;
  double dlr, dlt, dye;          // node spacings, (last - first) / (n - 1)
  double rmin, rmax;             // 10^lr0, 10^lr1 (the table's density range)
  double escale, eshift, pscale; // table units: e / escale - eshift, exp(ln P) * pscale
  const float *e, *lnp, *g1;     // record bases (MDS_EOS_E, _LNP, _GAM1)
};

// One problem as the kernels see it.
// DevProb::tcu_c: gcov[0][1..2] / gcov[0][0], gcon[1..2] * gcov[0][0], face
// gcov[1][1..2] / gcov[1][0], then the 11 non-zero connection coefficients
enum {
  MDS_TC_G1 = 0, MDS_TC_G2, MDS_TC_GI1, MDS_TC_GI2, MDS_TC_F1, MDS_TC_F2,
  MDS_TC_C000, MDS_TC_C101, MDS_TC_C202, MDS_TC_C011, MDS_TC_C110, MDS_TC_C111, MDS_TC_C212,
  MDS_TC_C022, MDS_TC_C122, MDS_TC_C220, MDS_TC_C221, MDS_TC_COUNT
};

struct DevProb {
  long long off;           // first interior zone in the packed state arrays
  int n[3];                // interior zones
  int ndim;
  int bc_lo[3], bc_hi[3];
  int sph;                 // spherical 1D column
  int rcoord;              // radial map (MDS_RCOORD_*)
  double rtrans;           // MDS_RCOORD_SINH scale
  double gam;
  double g0[3];            // fixed gravity (contravariant)
  double startx[3], dx[3];
  double idx[3];           // 1 / dx
  double volume, area[3];  // Cartesian uniform geometry
  double ivolume;          // 1 / volume
  // direction-0 FP32 reconstruction coefficients shared by every zone (1:
  // rcu_c holds them) -- the moments of a log-r column or of a uniform
  // Cartesian grid are the same up to scale and shift in every zone
  int rcu;
  float rcu_c[MDS_RC_COUNT];
  // FP32-bulk spherical columns with three velocity components: the
  // transverse metric over the radial one and the connection coefficients
  // shared by every zone (1: tcu_c holds them, MDS_TC_*) -- constant on a
  // log-r column, whose geometry is self-similar; else read per zone
  int tcu;
  double tcu_c[MDS_TC_COUNT];
  const double *mom[3];    // per direction (n[d]+2NG)*3: alpha, beta, Gamma per zone
  const float *rc[3];      // per direction (n[d]+2NG)*MDS_RC_COUNT
  const double *zone;      // spherical columns: (n[0]+2NG) x MDS_ZG_COUNT records; 3D spherical:
                           // field-major, zone[field][j + NG][i + NG] (MDS_ZS_*)
  const ZoneF32 *z32;      // spherical, FP32-bulk: n[0]+2NG records
  // dendritic 3D spherical grids (null otherwise): log2 of the theta and phi
  // coarsening, ghost-extended: lj[i + NG], lk[(j + NG) * (n[0] + 2NG) + i + NG]
  const unsigned char *lj, *lk;
  // theta (0) and phi (1) moments per level l at l * (n + 2NG) * 3 (alpha,
  // beta, Gamma per zone) and FP32 coefficients at l * (n + 2NG) * MDS_RC_COUNT
  const double *lmom[2];
  const float *lrc[2];
  // Ghost layers of the PROB and EXTRAP sides (null on other sides): FP64,
  // gs[d][side][(g * MDS_SLAB_NV + v) * T + t] for ghost layer g (index
  // g - NG below the grid, n[d] + g above), entry v (rho, u, v1, v2, v3, Ye,
  // P, Gamma_1) and transverse position t (slab_t); T = slab_T(d).  PROB
  // layers hold the caller's values, EXTRAP layers are refilled from the
  // state before every stage; P and Gamma_1 are refilled for both.
  int has_slab;
  double *gs[3][2];
  int order;               // hydro_interp_order (3 para_recon, 2 lin_recon, 1 constant)
  // monopole gravity: 0 none, MDS_MONOPOLE_NEWTON, MDS_MONOPOLE_GR.  GR: gr
  // is this problem's block of MDS_GR_COUNT arrays of n[0] + 1 doubles, grst
  // its state ([0] Gamma set, [1] passes of the last call, [2] calls that did
  // not converge)
  int mono;
  double *gr;
  int *grst;
};

// arrays of the GR effective monopole block of a problem (n[0] + 1 doubles each)
enum {
  MDS_GR_REDGE = 0,        // face radius                   (host)
  MDS_GR_R34,              // (4 pi) r^3 at the faces       (host)
  MDS_GR_DRDX,             // dr/dx at the faces            (host)
  MDS_GR_VSH,              // shell volume 4 pi/3 d(r^3)    (host)
  MDS_GR_GTOV,             // TOV factor Gamma at the faces (kept between calls)
  MDS_GR_MR,               // TOV mass at the faces
  MDS_GR_LAPSE_E,          // lapse at the faces
  MDS_GR_LAPSE,            // lapse at the zone centres
  MDS_GR_GFACE,            // face gravity
  MDS_GR_RHO, MDS_GR_P, MDS_GR_U, MDS_GR_V,   // shell sums, then averages
  MDS_GR_COUNT
};

#define MDS_SLAB_NV 8

// host: the radial shell geometry of the GR effective monopole (mds_grid.c)
extern "C" int mds_gr_shell_geometry(const mds_grid *g, double *out);

// transverse positions of a direction-d ghost layer, and the index of the
// zone (i, j, k) among them (box positions, natural order)
__host__ __device__ __forceinline__ long long slab_T(const DevProb &pb, int d)
{
  return d == 0 ? (long long)pb.n[1] * pb.n[2] : (d == 1 ? (long long)pb.n[0] * pb.n[2] : (long long)pb.n[0] * pb.n[1]);
}
__host__ __device__ __forceinline__ long long slab_t(const DevProb &pb, int d, int i, int j, int k)
{
  return d == 0 ? j + (long long)pb.n[1] * k : (d == 1 ? i + (long long)pb.n[0] * k : i + (long long)pb.n[0] * j);
}

// level-local counts of a dendritic grid (uniform grids: level 0 everywhere)
__device__ __forceinline__ int dend_lj(const DevProb &pb, int i) { return pb.lj ? pb.lj[i + MDS_NG] : 0; }
__device__ __forceinline__ int dend_lk(const DevProb &pb, int i, int j)
{
  return pb.lk ? pb.lk[(size_t)(j + MDS_NG) * (pb.n[0] + 2 * MDS_NG) + (i + MDS_NG)] : 0;
}
__device__ __forceinline__ bool dend_valid(const DevProb &pb, int i, int j, int k)
{
  return !pb.lj || (j < (pb.n[1] >> dend_lj(pb, i)) && k < (pb.n[2] >> dend_lk(pb, i, j)));
}

// Device-resident clock of one problem (the reference globals t, dt, istep,
// min_dt and main.c's dtold).
struct DevClock {
  double t, dt, dtold, tmax;
  unsigned long long mindt;   // order-preserving bits of the running min_dt
  int istep, active;
  int status;                 // 1: dt fell below dtmin (main.c:229-238)
  int pad;
  // health counters since the clock was set (mds_update_eos): zone updates
  // that produced a non-finite or non-positive density, or a non-finite
  // internal energy (negative too with the gamma law, where it is the
  // pressure over gamma - 1); table EOS evaluations of a stage's own zones
  // outside the table in density, Ye or temperature (the reference prints
  // these and continues, eos_stuff.c:571-604)
  unsigned long long nbad, noff;
};

struct RunParams {
  double cfl, dt_init_safety, dt_grow_safety, dtmin, dtmax, initial_dt;
};

template <class R>
struct StateArrays {
  R *v[6];                 // rho, u, v1, v2, v3, Ye (packed over problems; Ye with a table EOS)
};

struct Sweep1DArgs {
  const int2 *seg;         // (problem, first output zone) per wave
  int nseg;
  const DevProb *prob;
  DevClock *clock;
  unsigned long long *counters;
  RunParams rp;
  DevEos eos;              // tabulated EOS (when the batch has one)
  const double *grav;      // monopole gravity at zone centres, packed like the state
  int ncu;                 // compute units of the device
  int persist;             // 1: persistent grid (resident workgroups walk the segment table)
  int gen;                 // the general instances: PROB/EXTRAP ghost layers, hydro_interp_order < 3
  int zpl;                 // two-stage FP32-bulk kernel: zones per lane (1 or 2)
};

struct Sweep2DArgs {
  const DevProb *prob;     // single problem
  DevClock *clock;
  unsigned long long *counters;
  unsigned char *flags;    // stage-1 shock tags, one byte per zone
  int nstrip, nchunk, chunk;
  int src;                 // source terms present (fixed gravity)
  int slab;                // the general instances: PROB/EXTRAP ghost layers, hydro_interp_order < 3
};

struct Sweep3DArgs {
  const DevProb *prob;     // single problem
  DevClock *clock;
  unsigned long long *counters;
  unsigned char *flags;    // stage-1 shock tags, one byte per zone
  int nstrip, ntile, nchunk, chunk;
  int src;                 // source terms present (fixed gravity)
  int slab;                // the general instances: PROB/EXTRAP ghost layers, hydro_interp_order < 3
};

// multipole gravity of a 3D/2D spherical grid (gravity.c:499-1110): the
// host-made grid tables (mds_mp.h) on the device and the stage work arrays
struct MpDev {
  int lmax, mmax, nlm, maxklev;
  int nlist;               // (l, m) pairs evaluated: l = 0..lmax, |m| <= min(l, mmax), in that order
  int npart;               // row parts of the moment pass
  const short2 *lms;       // per pair: (lm index, l)
  const double *S;         // zone integrals of Y sin(th) per level pair and level-local (j, k)
  const long long *Soff;
  const double *Y;         // harmonics at the corner angles
  const double *rc, *rd;   // moment weights per shell and l
  const double *Cf, *Df;   // potential factors per corner radius and l
  const double *rm;        // corner radii
  const double *wx, *wy, *wz;   // centroid weights of the corner differences
  double *part;            // [pair][part][n1][2]: moment parts
  double *C, *D;           // [pair][n1]: moments inside / outside, then their running sums
  double *phi;             // [(n1 + 1) (n2 + 1) (n3 + 1)]: potential at the corners (level-local indices)
  double *src;             // [6][zones]: -rho dPhi/dx and -<rho v> dPhi/dx per direction
};

#define MDS_MP_ROWS 128    // (j, k) rows per wavefront in the multipole moment pass

// buffers of the direction-split 3D spherical stage
struct Sph3Launch {
  int fp32, tab, mono;
  void *F[3];              // face fluxes per direction, [slot][face]: FP64 area weighted,
                           // or FP32 per unit area on the FP32-bulk path
  long long nf[3];         // faces per direction
  void *P, *G, *CS;        // table EOS arrays (state precision)
  unsigned short *tg;      // temperature bracket of each zone's last EOS call
  double *ms, *grav;       // shell masses and monopole field (n1 each)
  double *part;            // shell-mass parts [npart][n1]
  int npart;
  hipStream_t side;        // second stream of a stage (default priority)
  hipEvent_t ev[5];        // fork/join events of a stage (no timing)
  unsigned char *flags;    // shock tags
  int dend;                // dendritic grid
  const int4 *irr[2];      // dendritic: radial and theta faces whose pencils cross a level change
  long long nirr[2];       // (i, j, k, 0) per face; k is the face's own phi index
  int mpl;                 // multipole gravity: lmax (0 = none)
  MpDev mp;
  int upd;                 // update workgroup shape (MDS_TILE_UPD3S)
};

#define MDS_SHELL_ROWS 32  // (j, k) rows per wavefront in the 3D shell-mass pass
#define MDS_SEG1D 60       // output zones per 64-lane wave in the 1D stage sweep
#define MDS_SEG1D2 56      // output zones per 64-lane wave in the 1D two-stage sweep
#define MDS_SEG1D2Z 120    // the same with two zones per lane (FP32-bulk)
