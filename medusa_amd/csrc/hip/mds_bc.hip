// mds_bc.hip -- ghost layers of the PROB and EXTRAP boundary sides (gfx950).
//
// Every other boundary code is a map from a ghost index onto an interior zone
// and is applied by the stage kernels at load time.  PROB (the problem's own
// ghost values, bc.c:125-140), EXTRAP (linear extrapolation, bc.c:806-821,
// 937-954) and DISK (the same at the inner radius with an inflow-only v^1,
// bc.c:863-878) produce values that are no interior zone, so they live in FP64
// ghost layers beside the state (DevProb::gs).  Before every stage this
// kernel refills the EXTRAP layers from the stage's input state and the
// pressure and Gamma_1 entries of every layer (the reference's EOS pass also
// covers the ghost zones, eos.c:4-32), which the 3D spherical kernels read
// for a table EOS.  One thread per ghost zone; blockIdx.y is the problem.
#include "mds_internal.h"
#include "mds_eos.h"

namespace {

template <class R>
__global__ __launch_bounds__(256) void slab_fill(const DevProb *probs, StateArrays<R> st, DevEos eos, int tab, int nv)
{
  const DevProb &pb = probs[blockIdx.y];
  if (!pb.has_slab) return;
  long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (int d = 0; d < 3; d++)
    for (int side = 0; side < 2; side++) {
      double *sl = pb.gs[d][side];
      if (!sl) continue;
      const long long T = slab_T(pb, d);
      if (it >= MDS_NG * T) {
        it -= MDS_NG * T;
        continue;
      }
      const int g = (int)(it / T);
      const long long t = it % T;
      double *e = sl + (size_t)g * MDS_SLAB_NV * T + t;       // entry v at e[v * T]
      const int code = side ? pb.bc_hi[d] : pb.bc_lo[d];
      if (code == MDS_BC_EXTRAP || code == MDS_BC_DISK) {
        // direction 0: edge zone + (distance) x (edge - next), at the ghost's
        // own (j, k) as the reference indexes it; DISK (inner radius) then
        // clamps an outward v^1 to 0 (bc.c:863-878)
        const int n = pb.n[0];
        const int c = side ? n + g : g - MDS_NG;
        const int i0 = side ? n - 1 : 0, i1 = side ? n - 2 : 1;
        const double w = side ? (double)(c - n + 1) : (double)c;
        const int j = (int)(t % pb.n[1]), k = (int)(t / pb.n[1]);
        const bool ok = dend_valid(pb, i0, j, k);
        const long long z0 = pb.off + ((long long)k * pb.n[1] + j) * pb.n[0] + i0;
        const long long z1 = z0 + (i1 - i0);
        for (int v = 0; v < 6; v++) {
          const bool held = v < 2 + nv || (v == 5 && tab);
          double x = 0.0;
          if (held && ok) {
            const double q0 = (double)st.v[v][z0], q1 = (double)st.v[v][z1];
            x = side ? q0 + w * (q0 - q1) : q0 + w * (q1 - q0);
          }
          if (code == MDS_BC_DISK && v == 2 && x > 0.0) x = 0.0;
          e[v * T] = x;
        }
      }
      // pressure and Gamma_1 of the ghost zone (eos_gamma.c:45-57 or the table)
      const double rho = e[0], u = e[T];
      double press = 0.0, g1 = 0.0;
      if (!tab) {
        press = (pb.gam - 1.0) * u;
        g1 = pb.gam;
      } else if (rho > 0.0) {
        double cs, lt;
        eos_fill_x64(eos, rho, u, e[5 * T], press, g1, cs, lt);
      }
      e[6 * T] = press;
      e[7 * T] = g1;
      return;
    }
}

}  // namespace

extern "C" hipError_t mds_launch_slab_fill(int fp32, const DevProb *probs, int nprob, long long items,
                                           void *const *state, const DevEos *eos, int tab, int nv, hipStream_t st)
{
  if (items <= 0 || nprob <= 0) return hipSuccess;
  if (nprob > 65535 || (items + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((items + 255) / 256), (unsigned)nprob), blk(256);
  if (fp32) {
    StateArrays<float> s;
    for (int k = 0; k < 6; k++) s.v[k] = (float *)state[k];
    hipLaunchKernelGGL(slab_fill<float>, grid, blk, 0, st, probs, s, *eos, tab, nv);
  } else {
    StateArrays<double> s;
    for (int k = 0; k < 6; k++) s.v[k] = (double *)state[k];
    hipLaunchKernelGGL(slab_fill<double>, grid, blk, 0, st, probs, s, *eos, tab, nv);
  }
  return hipGetLastError();
}
