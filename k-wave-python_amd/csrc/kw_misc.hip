// Launchers instantiating the time-loop kernels (see kw_kernels.hpp).
#include "kw_kernels.hpp"

namespace kw {

// ---------------------------------------------------------------------------
// sensors, step counter, small utilities
// ---------------------------------------------------------------------------
__global__ void k_gather(Geom g, Fields fl, const Inst* __restrict__ inst) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int t = fl.tdev[b];
  if (I.sens_kind == 0 || t >= I.nt || t < I.rec_start) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= I.nsens) return;
  const int64_t off = b * g.ntot;
  const int64_t row = (int64_t)(t - I.rec_start) * I.nsens + s;
  auto sample = [&](const float* f) {
    float acc = 0.f;
    for (int c = 0; c < I.ncorner; ++c) {
      const int64_t ii = I.sidx[(int64_t)s * I.ncorner + c];
      const float w = I.sens_kind == 2 ? I.sw[(int64_t)s * I.ncorner + c] : 1.f;
      acc += w * f[off + ii];
    }
    return acc;
  };
  if (I.rec_p || I.agg_p) {
    const float v = sample(fl.p);
    if (I.rec_p) I.rec_p[row] = v;
    if (I.agg_p) {
      float* a = I.agg_p;
      if (t == I.rec_start) { a[s] = v; a[I.nsens + s] = v; a[2 * I.nsens + s] = v * v; }
      else { a[s] = fmaxf(a[s], v); a[I.nsens + s] = fminf(a[I.nsens + s], v); a[2 * I.nsens + s] += v * v; }
    }
  }
  for (int i = 0; i < g.nd; ++i) {
    if (I.rec_u[i]) I.rec_u[i][row] = sample(fl.u[i]);
    if (I.rec_uc[i]) I.rec_uc[i][row] = sample(fl.uc[i]);
  }
}

__global__ void k_tick(Fields fl, const Inst* __restrict__ inst, int nb) {
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    if (fl.tdev[b] < inst[b].nt) fl.tdev[b] += 1;
}

__global__ void k_fill(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_setint(int* __restrict__ p, int n, int v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

bool size_supported(int n) { return size_supported_impl(n); }

hipError_t launch_gather(const Geom& g, int nb, int max_sens, const Fields& fl, const Inst* inst, hipStream_t st) {
  if (max_sens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((max_sens + 255) / 256, nb), dim3(256), 0, st, g, fl, inst);
  return hipGetLastError();
}

hipError_t launch_tick(const Fields& fl, const Inst* inst, int nb, hipStream_t st) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(256), 0, st, fl, inst, nb);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_setint(int* p, int n, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_setint, dim3(1), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace kw
