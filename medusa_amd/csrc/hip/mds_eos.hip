// mds_eos.hip -- tabulated EOS: device table residency and the pointwise
// parity entry point (replaces eos_fill / eos_given_rex_per_zone,
// eos/burrows/eos_stuff.c:768-901, for arrays of states).
#include <hip/hip_runtime.h>
#include <vector>
#include "mds_internal.h"
#include "mds_eos.h"
#include "../../../include/mds_eos.h"

// Describe a host table for the device and copy its records to HBM
// (*alloc receives the device allocation, to be released with hipFree).
hipError_t mds_eos_to_device(const mds_eos_table *t, DevEos *d, void **alloc)
{
  *alloc = nullptr;
  const size_t rec = (size_t)t->ny * t->nr * t->nt;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, MDS_EOS_NFIELD * rec * sizeof(float));
  if (e != hipSuccess) return e;
  e = hipMemcpy(p, t->data, MDS_EOS_NFIELD * rec * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) { hipFree(p); return e; }
  *alloc = p;
  d->nr = t->nr; d->nt = t->nt; d->ny = t->ny;
  d->lr0 = t->lr0; d->lr1 = t->lr1; d->lt0 = t->lt0; d->lt1 = t->lt1; d->ye0 = t->ye0; d->ye1 = t->ye1;
  d->dlr = (t->lr1 - t->lr0) / (t->nr - 1);
  d->dlt = (t->lt1 - t->lt0) / (t->nt - 1);
  d->dye = (t->ye1 - t->ye0) / (t->ny - 1);
  d->escale = t->e_scale > 0.0 ? t->e_scale : 1.0;
  d->eshift = t->e_shift;
  d->pscale = t->p_scale > 0.0 ? t->p_scale : 1.0;
  d->rmin = pow(10.0, t->lr0);
  d->rmax = pow(10.0, t->lr1);
  d->e = (const float *)p + MDS_EOS_E * rec;
  d->lnp = (const float *)p + MDS_EOS_LNP * rec;
  d->g1 = (const float *)p + MDS_EOS_GAM1 * rec;
  return hipSuccess;
}

__global__ void eos_points_kernel(DevEos t, int fp32, int n, const double *rho, const double *u, const double *ye,
                                  double *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p, lt, g, c;
  if (fp32) {
    float g32, c32;
    eos_fill_f32(t, rho[i], u[i], ye[i], p, g32, c32, lt);
    g = g32;
    c = c32;
  } else {
    eos_fill_x64(t, rho[i], u[i], ye[i], p, g, c, lt);
  }
  out[4 * i + 0] = p;
  out[4 * i + 1] = c;
  out[4 * i + 2] = pow(10.0, lt);
  out[4 * i + 3] = g;
}

extern "C" int mds_eos_points(int precision, const mds_eos_table *t, int n, const double *rho, const double *u,
                              const double *ye, double *press, double *cs, double *temp, double *gam1)
{
  if (!t || !t->data || n < 0 || !rho || !u || !ye) return MDS_EINVAL;
  if (t->nr < 2 || t->nt < 3 || t->ny < 2) return MDS_EINVAL;
  if (n == 0) return MDS_OK;
  DevEos d;
  void *tab = nullptr;
  if (mds_eos_to_device(t, &d, &tab) != hipSuccess) return MDS_EDEVICE;
  double *dbuf = nullptr;
  hipError_t e = hipMalloc(&dbuf, sizeof(double) * 7 * (size_t)n);
  if (e == hipSuccess) e = hipMemcpy(dbuf, rho, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbuf + n, u, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbuf + 2 * (size_t)n, ye, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(eos_points_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d,
                       precision == MDS_FP32_BULK, n, dbuf, dbuf + n, dbuf + 2 * (size_t)n, dbuf + 3 * (size_t)n);
    e = hipGetLastError();
  }
  std::vector<double> out(4 * (size_t)n);
  if (e == hipSuccess) e = hipMemcpy(out.data(), dbuf + 3 * (size_t)n, sizeof(double) * 4 * n, hipMemcpyDeviceToHost);
  if (dbuf) hipFree(dbuf);
  hipFree(tab);
  if (e != hipSuccess) return MDS_EDEVICE;
  for (int i = 0; i < n; i++) {
    if (press) press[i] = out[4 * i + 0];
    if (cs) cs[i] = out[4 * i + 1];
    if (temp) temp[i] = out[4 * i + 2];
    if (gam1) gam1[i] = out[4 * i + 3];
  }
  return MDS_OK;
}
