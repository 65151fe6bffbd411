// Launchers instantiating the time-loop kernels (see kw_kernels.hpp).
#include <algorithm>

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

// ---------------------------------------------------------------------------
// reverse sweep (kw_plan_adjoint_c; see kw_adjoint.hpp)
// ---------------------------------------------------------------------------
// pbar += S^T r_t: the transpose of k_gather for step t.  Binary masks hit
// distinct voxels; interpolated points may share corners and add atomically.
__global__ void k_adj_scatter(Geom g, AdjFields af, const Inst* __restrict__ inst, int t) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int j = t - I.rec_start;
  if (I.sens_kind == 0 || j < 0 || j >= af.nrec) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= I.nsens) return;
  const float r = af.adj[((int64_t)b * af.nrec + j) * af.nsens + s];
  float* pb = af.pb + b * g.ntot;
  if (I.ncorner == 1) {
    pb[I.sidx[s]] += r;
  } else {
    for (int c = 0; c < I.ncorner; ++c) {
      const int64_t k = (int64_t)s * I.ncorner + c;
      atomicAdd(pb + I.sidx[k], (I.sens_kind == 2 ? I.sw[k] : 1.f) * r);
    }
  }
}

// Source-scale term: the injected value scale * sig(t) depends on c at the source
// point.  Pressure (which = -1): scale = 2 dt / (D c d) (kspace_solver.py:619-624), so
// cbar += -(scale / c) sig(t) Z with Z = K R (additive) or R (uncorrected);
// velocity along axis which: scale = 2 c dt / d_i (:636-642), so cbar += (scale / c)
// sig(t) Z with Z = K U_i or U_i.  Zn is the unfiltered array (R or U_i), Zk the
// filtered one.  One thread per grid row walks the row's source points (CSR).
__global__ void k_adj_srcgrad(Geom g, AdjFields af, int which, const float* __restrict__ Zn, const float* __restrict__ Zk,
                              const Inst* __restrict__ inst, int t) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const Src& s = which < 0 ? I.srcp : I.srcu[which];
  if (s.mode == SRC_NONE || s.mode == SRC_DIRICHLET || t >= s.len) return;
  const float* Z = (s.mode == SRC_NOCORR ? Zn : Zk) + b * g.ntot;
  const float sgn = which < 0 ? -1.f : 1.f;
  float* cb = af.cb + b * g.ntot;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < g.rows; r += gridDim.x * blockDim.x) {
    for (int e = s.row_ptr[r]; e < s.row_ptr[r + 1]; ++e) {
      const int64_t idx = (int64_t)r * g.nz + s.col[e];
      const float c = sqrtf(I.c0sq[idx * I.c0sq_s]);
      const float sv = s.sig[(int64_t)s.sigrow[e] * s.len + t];
      cb[idx] += sgn * (s.scale[e] / c) * sv * Z[idx];
    }
  }
}

// Initial-pressure term of dJ/dc0 (one instance): the state after step 0 holds
// rho_i = p0 / (c0^2 p0div_i), so cb += -2 p0 / c0^3 sum_i rb_i / p0div_i with rb_i
// the adjoints of that state (p0div_i infinite for an axis that gets no share).
__global__ void k_adj_p0(int64_t n, int D, const float* __restrict__ p0, const float* __restrict__ c0sq, int c0sq_s,
                         const float* __restrict__ rb0, const float* __restrict__ rb1, const float* __restrict__ rb2, float w0,
                         float w1, float w2, float* __restrict__ cb) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float r = w0 * rb0[i];
    if (D > 1) r += w1 * rb1[i];
    if (D > 2) r += w2 * rb2[i];
    const float c2 = c0sq[i * c0sq_s];
    cb[i] += -2.f * p0[i] / (c2 * sqrtf(c2)) * r;
  }
}

hipError_t launch_adj_p0(int64_t n, int D, const float* p0, const float* c0sq, int c0sq_s, const float* const rb[3], const float w[3],
                         float* cb, hipStream_t st) {
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_adj_p0, dim3((unsigned)blocks), dim3(256), 0, st, n, D, p0, c0sq, c0sq_s, rb[0], D > 1 ? rb[1] : rb[0],
                     D > 2 ? rb[2] : rb[0], w[0], w[1], w[2], cb);
  return hipGetLastError();
}

hipError_t launch_adj_scatter(const Geom& g, int nb, int max_sens, const AdjFields& af, const Inst* inst, int t, hipStream_t st) {
  if (max_sens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_scatter, dim3((max_sens + 255) / 256, nb), dim3(256), 0, st, g, af, inst, t);
  return hipGetLastError();
}

hipError_t launch_adj_srcgrad(const Geom& g, int nb, const AdjFields& af, int which, const float* Zn, const float* Zk, const Inst* inst, int t,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_adj_srcgrad, dim3((g.rows + 255) / 256, nb), dim3(256), 0, st, g, af, which, Zn, Zk, inst, t);
  return hipGetLastError();
}


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
