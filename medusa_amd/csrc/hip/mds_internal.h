// mds_internal.h -- device-side descriptors shared by the sweep kernels and
// the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "../../../include/mds_hydro.h"

// One problem as the kernels see it.
struct DevProb {
  long long off;           // first interior zone in the packed state arrays
  int n[3];                // interior zones
  int ndim;
  int bc_lo[3], bc_hi[3];
  int sph;                 // spherical 1D column
  double gam;
  double g0[3];            // fixed gravity (contravariant)
  double startx[3], dx[3];
  double volume, area[3];  // Cartesian uniform geometry
  const double *mom[3];    // per direction (n[d]+2NG)*3: alpha, beta, Gamma per zone
  const float *rc[3];      // per direction (n[d]+2NG)*MDS_RC_COUNT
  const double *zone;      // spherical: (n[0]+2NG)*MDS_ZG_COUNT
};

// Device-resident clock of one problem (the reference globals t, dt, istep,
// min_dt and main.c's dtold).
struct DevClock {
  double t, dt, dtold, tmax;
  unsigned long long mindt;   // order-preserving bits of the running min_dt
  int istep, active;
  int status;                 // 1: dt fell below dtmin (main.c:229-238)
  int pad;
};

struct RunParams {
  double cfl, dt_init_safety, dt_grow_safety, dtmin, dtmax, initial_dt;
};

template <class R>
struct StateArrays {
  R *v[5];                 // rho, u, v1, v2, v3 (packed over problems)
};

struct Sweep1DArgs {
  const int2 *seg;         // (problem, first output zone) per wave
  int nseg;
  const DevProb *prob;
  DevClock *clock;
  unsigned long long *counters;
  RunParams rp;
};

struct Sweep2DArgs {
  const DevProb *prob;     // single problem
  DevClock *clock;
  unsigned long long *counters;
  unsigned char *flags;    // stage-1 shock tags, one byte per zone
  int nstrip, nchunk, chunk;
  int src;                 // source terms present (fixed gravity)
};

struct Sweep3DArgs {
  const DevProb *prob;     // single problem
  DevClock *clock;
  unsigned long long *counters;
  unsigned char *flags;    // stage-1 shock tags, one byte per zone
  int nstrip, ntile, nchunk, chunk;
  int src;                 // source terms present (fixed gravity)
};

#define MDS_SEG1D 60       // output zones per 64-lane wave in the 1D sweep
