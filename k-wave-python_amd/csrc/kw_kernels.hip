// Time-loop kernels of the k-space pseudospectral solver (gfx950).
//
// Every kernel works on a *batch* of same-shape instances (blockIdx.y is the
// instance) and reads the instance's step counter from device memory, so that
// one captured sequence serves every step and finished instances drop out.
//
// z-pass kernels own the contiguous axis: a workgroup takes L pairs of grid
// rows, runs the real<->half-complex transform of both rows of a pair as one
// complex FFT (row a in the real part, row b in the imaginary part), and fuses
// the pointwise physics of the step between the inverse and forward transforms.
// Column kernels own the strided axes; the outermost-axis pass applies the
// spectral multipliers (regenerated in registers from 1-D tables) between its
// forward and inverse transforms.
#include "kw_fft.hpp"
#include "kw_launch.hpp"
#include "kw_types.hpp"

namespace kw {

// ---------------------------------------------------------------------------
// z-pass configuration
// ---------------------------------------------------------------------------
template <int N>
struct ZCfg {
  static constexpr int L0 = 2048 / N;
  static constexpr int L = L0 < 1 ? 1 : (L0 > 32 ? 32 : L0);  // row pairs per workgroup
  static constexpr int NT = 256;
  static constexpr int LP = L + 1;
  static constexpr int NH = N / 2 + 1;
  static constexpr int M = (N * L + NT - 1) / NT;  // (pair, n) elements per thread
  static constexpr int LDS = N * LP;
};

__device__ __forceinline__ float ld(const float* p, int64_t i, int s) { return p[i * s]; }

// pml value of axis `ax` at grid row r, contiguous position n
__device__ __forceinline__ float pml_at(const float* const* tab, const Geom& g, int ax, int r, int n) {
  if (ax == g.nd - 1) return tab[ax][n];
  if (ax == 0) return tab[0][g.nd == 2 ? r : r / g.n[1]];
  return tab[1][r % g.n[1]];
}

// Load the half spectra of the L row pairs of this workgroup into LDS as the
// full spectrum of z = a + i b (Hermitian extension of each half spectrum).
template <int N>
__device__ __forceinline__ void load_pair_spec(float2* buf, const float2* __restrict__ spec, int row0, int rows, int tid) {
  using C = ZCfg<N>;
  for (int e = tid; e < C::L * C::NH; e += C::NT) {
    const int p = e / C::NH, k = e % C::NH;
    const int ra = row0 + 2 * p, rb = ra + 1;
    float2 A = ra < rows ? spec[(int64_t)ra * C::NH + k] : make_float2(0.f, 0.f);
    float2 B = rb < rows ? spec[(int64_t)rb * C::NH + k] : make_float2(0.f, 0.f);
    if (k == 0 || (N % 2 == 0 && k == N / 2)) {
      buf[k * C::LP + p] = make_float2(A.x, B.x);
    } else {
      buf[k * C::LP + p] = make_float2(A.x - B.y, A.y + B.x);
      buf[(N - k) * C::LP + p] = make_float2(A.x + B.y, B.x - A.y);
    }
  }
}

// After a forward FFT of z = a + i b in LDS, split into the two half spectra.
template <int N>
__device__ __forceinline__ void store_pair_spec(const float2* buf, float2* __restrict__ spec, int row0, int rows, int tid) {
  using C = ZCfg<N>;
  for (int e = tid; e < C::L * C::NH; e += C::NT) {
    const int p = e / C::NH, k = e % C::NH;
    const int ra = row0 + 2 * p, rb = ra + 1;
    const float2 Z = buf[k * C::LP + p];
    const float2 Zc = buf[((N - k) % N) * C::LP + p];
    if (ra < rows) spec[(int64_t)ra * C::NH + k] = make_float2(0.5f * (Z.x + Zc.x), 0.5f * (Z.y - Zc.y));
    if (rb < rows) spec[(int64_t)rb * C::NH + k] = make_float2(0.5f * (Z.y + Zc.y), -0.5f * (Z.x - Zc.x));
  }
}

template <int N>
__device__ __forceinline__ void inv_to_regs(float2* buf, const float2* __restrict__ spec, const float2* __restrict__ tw,
                                            int row0, int rows, int tid, float2 (&out)[ZCfg<N>::M]) {
  using C = ZCfg<N>;
  load_pair_spec<N>(buf, spec, row0, rows, tid);
  __syncthreads();
  fft_lds<N, C::L, C::NT, +1>(buf, tw, tid);
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if ((N * C::L) % C::NT == 0 || e < N * C::L) out[m] = buf[(e % N) * C::LP + e / N];
  }
  __syncthreads();
}

template <int N>
__device__ __forceinline__ void fwd_from_regs(float2* buf, const float2 (&in)[ZCfg<N>::M], float2* __restrict__ spec,
                                              const float2* __restrict__ tw, int row0, int rows, int tid) {
  using C = ZCfg<N>;
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if ((N * C::L) % C::NT == 0 || e < N * C::L) buf[(e % N) * C::LP + e / N] = in[m];
  }
  __syncthreads();
  fft_lds<N, C::L, C::NT, -1>(buf, tw, tid);
  store_pair_spec<N>(buf, spec, row0, rows, tid);
  __syncthreads();
}

// Source value added/imposed at (row r, position n) for a CSR source term.
// Returns true when (r, n) is a source point; v receives scale * signal(t).
__device__ __forceinline__ bool csr_lookup(const Src& s, int r, int n, int t, float& v) {
  int lo = s.row_ptr[r], hi = s.row_ptr[r + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int c = s.col[mid];
    if (c == n) {
      v = s.scale[mid] * s.sig[(int64_t)s.sigrow[mid] * s.len + t];
      return true;
    }
    if (c < n) lo = mid + 1; else hi = mid;
  }
  return false;
}

// Apply a source term to field value f at flat index idx (row r, position n),
// axis factor fac. chain holds the inverse-transformed filtered source (mode CHAIN).
__device__ __forceinline__ float apply_src(const Src& s, float f, int64_t idx, int r, int n, int t, float fac, float chain) {
  if (s.mode == SRC_NONE || t >= s.len) return f;
  if (s.mode == SRC_PATTERN) return f + fac * s.sig[t] * s.pattern[idx];
  if (s.mode == SRC_CHAIN) return f + fac * chain;
  float v;
  if (!csr_lookup(s, r, n, t, v)) return f;
  return s.mode == SRC_DIRICHLET ? fac * v : f + fac * v;
}

#define KW_ZPASS_PROLOGUE                                              \
  using C = ZCfg<N>;                                                   \
  __shared__ float2 buf[C::LDS];                                       \
  const int tid = threadIdx.x;                                         \
  const int b = blockIdx.y;                                            \
  const int row0 = blockIdx.x * 2 * C::L;                              \
  const int rows = g.rows;

// ---------------------------------------------------------------------------
// generic real <-> half-spectrum transforms along the contiguous axis
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_zfwd(Geom g, const float* __restrict__ f, int64_t fstride, float2* __restrict__ spec,
                                              int64_t sstride, const float2* __restrict__ tw) {
  KW_ZPASS_PROLOGUE
  const float* fb = f + b * fstride;
  float2 v[C::M];
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    const int p = e / N, n = e % N;
    const int ra = row0 + 2 * p, rb = ra + 1;
    v[m] = make_float2(ra < rows && e < N * C::L ? fb[(int64_t)ra * N + n] : 0.f, rb < rows && e < N * C::L ? fb[(int64_t)rb * N + n] : 0.f);
  }
  fwd_from_regs<N>(buf, v, spec + b * sstride, tw, row0, rows, tid);
}

template <int N>
__global__ __launch_bounds__(256) void k_zinv(Geom g, const float2* __restrict__ spec, int64_t sstride, float* __restrict__ f,
                                              int64_t fstride, const float2* __restrict__ tw, float scale) {
  KW_ZPASS_PROLOGUE
  float* fb = f + b * fstride;
  float2 v[C::M];
  inv_to_regs<N>(buf, spec + b * sstride, tw, row0, rows, tid, v);
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if (e >= N * C::L) continue;
    const int p = e / N, n = e % N;
    const int ra = row0 + 2 * p, rb = ra + 1;
    if (ra < rows) fb[(int64_t)ra * N + n] = scale * v[m].x;
    if (rb < rows) fb[(int64_t)rb * N + n] = scale * v[m].y;
  }
}

// ---------------------------------------------------------------------------
// momentum: u_i = pml_sg (pml_sg u_i - dt/rho0_sg * grad_i p) (+ source), then
// forward transform of u_i along z (for the divergence).  blockIdx.z = axis.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_momentum(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw) {
  KW_ZPASS_PROLOGUE
  const int ax = blockIdx.z;
  const Inst& I = inst[b];
  const int t = fl.tdev[b];
  if (t >= I.nt) return;
  float2* sp = fl.spec[ax] + b * g.nspec;
  float* u = fl.u[ax] + b * g.ntot;
  const Src& src = I.srcu[ax];
  float2 gv[C::M];
  inv_to_regs<N>(buf, sp, tw, row0, rows, tid, gv);
  float2 ch[C::M];
  const bool chain = src.mode == SRC_CHAIN && t < src.len;
  if (chain) inv_to_regs<N>(buf, fl.spec[src.chain_buf] + b * g.nspec, tw, row0, rows, tid, ch);
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if (e >= N * C::L) continue;
    const int p = e / N, n = e % N;
    const int ra = row0 + 2 * p;
    float res[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const int64_t idx = (int64_t)r * N + n;
      const float s = pml_at(I.pmlsg, g, ax, r, n);
      const float gval = h ? gv[m].y : gv[m].x;
      float uv = s * (s * u[idx] - ld(I.dtr0[ax], idx, I.dtr0_s) * gval);
      uv = apply_src(src, uv, idx, r, n, t, src.fac[ax], chain ? (h ? ch[m].y : ch[m].x) : 0.f);
      u[idx] = uv;
      res[h] = uv;
    }
    gv[m] = make_float2(res[0], res[1]);
  }
  fwd_from_regs<N>(buf, gv, sp, tw, row0, rows, tid);
}

// ---------------------------------------------------------------------------
// mass conservation (+ equation of state when no fractional Laplacian is needed)
// ---------------------------------------------------------------------------
template <int N, int ND>
__global__ __launch_bounds__(256) void k_mass(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw, int flags) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  const int t = fl.tdev[b];
  if (t >= I.nt) return;
  const int64_t off = b * g.ntot;
  float2 d[ND][C::M];
#pragma unroll
  for (int i = 0; i < ND; ++i) inv_to_regs<N>(buf, fl.spec[i] + b * g.nspec, tw, row0, rows, tid, d[i]);
  const Src& src = I.srcp;
  const bool chain = src.mode == SRC_CHAIN && t < src.len;
  float2 ch[C::M];
  if (chain) inv_to_regs<N>(buf, fl.spec[src.chain_buf] + b * g.nspec, tw, row0, rows, tid, ch);
  const bool powerlaw = flags & (FLAG_POWERLAW_ABS | FLAG_POWERLAW_DISP);
  float2 oa[C::M], ob[C::M];
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if (e >= N * C::L) continue;
    const int p = e / N, n = e % N;
    const int ra = row0 + 2 * p;
    float ra_out[2] = {0.f, 0.f}, rb_out[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const int64_t idx = (int64_t)r * N + n;
      const float r0 = ld(I.rho0, idx, I.rho0_s);
      float rho[ND];
      float rsum = 0.f;
#pragma unroll
      for (int i = 0; i < ND; ++i) { rho[i] = fl.rho[i][off + idx]; rsum += rho[i]; }
      const float nl = (flags & FLAG_NONLINEAR) ? (2.f * rsum + r0) / r0 : 1.f;
      float divu = 0.f, rt = 0.f;
      const float chv = chain ? (h ? ch[m].y : ch[m].x) : 0.f;
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const float dv = h ? d[i][m].y : d[i][m].x;
        const float s = pml_at(I.pml, g, i, r, n);
        float rv = s * (s * rho[i] - I.dt * r0 * dv * nl);
        rv = apply_src(src, rv, idx, r, n, t, src.fac[i], chv);
        fl.rho[i][off + idx] = rv;
        divu += dv;
        rt += rv;
      }
      if (powerlaw) {
        fl.rhot[off + idx] = rt;
        ra_out[h] = r0 * divu;
        rb_out[h] = rt;
      } else {
        float q = rt;
        if (flags & FLAG_STOKES) q += ld(I.tau, idx, I.tau_s) * r0 * divu;
        if (flags & FLAG_NONLINEAR) q += ld(I.bona, idx, I.bona_s) * rt * rt / (2.f * r0);
        const float pv = ld(I.c0sq, idx, I.c0sq_s) * q;
        fl.p[off + idx] = pv;
        ra_out[h] = pv;
      }
    }
    oa[m] = make_float2(ra_out[0], ra_out[1]);
    ob[m] = make_float2(rb_out[0], rb_out[1]);
  }
  if (!powerlaw) {
    fwd_from_regs<N>(buf, oa, fl.spec[0] + b * g.nspec, tw, row0, rows, tid);
  } else {
    if (flags & FLAG_POWERLAW_ABS) fwd_from_regs<N>(buf, oa, fl.spec[0] + b * g.nspec, tw, row0, rows, tid);
    if (flags & FLAG_POWERLAW_DISP) fwd_from_regs<N>(buf, ob, fl.spec[1] + b * g.nspec, tw, row0, rows, tid);
  }
}

// ---------------------------------------------------------------------------
// equation of state with power-law absorption / dispersion
// p = c0^2 (rho + tau F^-1[|k|^{y-2} F[rho0 div u]] - eta F^-1[|k|^{y-1} F[rho]] + B/A rho^2 / (2 rho0))
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_eos(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw, int flags) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  const int t = fl.tdev[b];
  if (t >= I.nt) return;
  const int64_t off = b * g.ntot;
  float2 av[C::M], bv[C::M];
  if (flags & FLAG_POWERLAW_ABS) inv_to_regs<N>(buf, fl.spec[0] + b * g.nspec, tw, row0, rows, tid, av);
  if (flags & FLAG_POWERLAW_DISP) inv_to_regs<N>(buf, fl.spec[1] + b * g.nspec, tw, row0, rows, tid, bv);
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if (e >= N * C::L) continue;
    const int p = e / N, n = e % N;
    const int ra = row0 + 2 * p;
    float out[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const int64_t idx = (int64_t)r * N + n;
      const float rt = fl.rhot[off + idx];
      float q = rt;
      if (flags & FLAG_POWERLAW_ABS) q += ld(I.tau, idx, I.tau_s) * (h ? av[m].y : av[m].x);
      if (flags & FLAG_POWERLAW_DISP) q -= ld(I.eta, idx, I.eta_s) * (h ? bv[m].y : bv[m].x);
      if (flags & FLAG_NONLINEAR) q += ld(I.bona, idx, I.bona_s) * rt * rt / (2.f * ld(I.rho0, idx, I.rho0_s));
      const float pv = ld(I.c0sq, idx, I.c0sq_s) * q;
      fl.p[off + idx] = pv;
      out[h] = pv;
    }
    av[m] = make_float2(out[0], out[1]);
  }
  fwd_from_regs<N>(buf, av, fl.spec[0] + b * g.nspec, tw, row0, rows, tid);
}

// ---------------------------------------------------------------------------
// t = 0 initial-value override: u_i = dt/(2 rho0_sg_i) grad_i p0, rho_j = p0 / (c0^2 D)
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_init_u(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw) {
  KW_ZPASS_PROLOGUE
  const int ax = blockIdx.z;
  const Inst& I = inst[b];
  const int64_t off = b * g.ntot;
  float2 gv[C::M];
  inv_to_regs<N>(buf, fl.spec[ax] + b * g.nspec, tw, row0, rows, tid, gv);
#pragma unroll
  for (int m = 0; m < C::M; ++m) {
    const int e = tid + m * C::NT;
    if (e >= N * C::L) continue;
    const int p = e / N, n = e % N;
    const int ra = row0 + 2 * p;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const int64_t idx = (int64_t)r * N + n;
      fl.u[ax][off + idx] = 0.5f * ld(I.dtr0[ax], idx, I.dtr0_s) * (h ? gv[m].y : gv[m].x);
      if (ax == 0) {
        const float rv = fl.p[off + idx] / (ld(I.c0sq, idx, I.c0sq_s) * (float)g.nd);
        for (int j = 0; j < g.nd; ++j) fl.rho[j][off + idx] = rv;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// build the z-spectrum of a multi-row additive source at step t (mode CHAIN)
// which = -1: pressure source, 0..2: velocity component
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_src_build(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw, int which) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  const int t = fl.tdev[b];
  const Src& s = which < 0 ? I.srcp : I.srcu[which];
  if (t >= I.nt || s.mode != SRC_CHAIN || t >= s.len) return;
  float2* sp = fl.spec[s.chain_buf] + b * g.nspec;
  const int rend = min(row0 + 2 * C::L, rows);
  const int e0 = s.row_ptr[row0], e1 = s.row_ptr[rend];
  if (e0 == e1) {
    for (int e = tid; e < C::L * C::NH; e += C::NT) {
      const int p = e / C::NH, k = e % C::NH;
      const int ra = row0 + 2 * p, rb = ra + 1;
      if (ra < rows) sp[(int64_t)ra * C::NH + k] = make_float2(0.f, 0.f);
      if (rb < rows) sp[(int64_t)rb * C::NH + k] = make_float2(0.f, 0.f);
    }
    return;
  }
  for (int e = tid; e < N * C::L; e += C::NT) buf[(e % N) * C::LP + e / N] = make_float2(0.f, 0.f);
  __syncthreads();
  float* fb = reinterpret_cast<float*>(buf);
  for (int e = e0 + tid; e < e1; e += C::NT) {
    // row of entry e: smallest r with row_ptr[r+1] > e
    int r = row0;
    while (s.row_ptr[r + 1] <= e) ++r;
    const int p = (r - row0) >> 1, h = (r - row0) & 1;
    fb[2 * (s.col[e] * C::LP + p) + h] = s.scale[e] * s.sig[(int64_t)s.sigrow[e] * s.len + t];
  }
  __syncthreads();
  fft_lds<N, C::L, C::NT, -1>(buf, tw, tid);
  store_pair_spec<N>(buf, sp, row0, rows, tid);
}

// ---------------------------------------------------------------------------
// strided column passes
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sincf(float x) { return fabsf(x) < 1e-4f ? 1.f - x * x * (1.f / 6.f) : sinf(x) / x; }

template <int N, int L>
struct ColCfg {
  static constexpr int NT = (N * L >= 8192) ? 512 : 256;
  static constexpr int LP = L + 1;
  static constexpr int M = (N * L + NT - 1) / NT;
  static constexpr int LDS = N * LP;
};

template <int N, int L>
__global__ __launch_bounds__((ColCfg<N, L>::NT)) void k_col(Geom g, Fields fl, Ops ops, const Inst* __restrict__ inst, ColArgs a) {
  using C = ColCfg<N, L>;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  if (fl.tdev && fl.tdev[b] >= inst[b].nt) return;
  const int ntile = (a.ncols + L - 1) / L;
  const int o = blockIdx.x / ntile, tile = blockIdx.x % ntile;
  const int64_t base = (int64_t)b * g.nspec + (int64_t)o * a.ostride + (int64_t)tile * L;
  const int cvalid = min(L, a.ncols - tile * L);
  const float2* tw = ops.tw[a.axis];

  auto load = [&](const float2* src, float2 (&v)[C::M]) {
#pragma unroll
    for (int m = 0; m < C::M; ++m) {
      const int e = tid + m * C::NT;
      const int c = e % L, n = e / L;
      v[m] = (e < N * L && c < cvalid) ? src[base + (int64_t)n * a.estride + c] : make_float2(0.f, 0.f);
    }
  };
  auto store = [&](float2* dst, const float2 (&v)[C::M]) {
#pragma unroll
    for (int m = 0; m < C::M; ++m) {
      const int e = tid + m * C::NT;
      const int c = e % L, n = e / L;
      if (e < N * L && c < cvalid) dst[base + (int64_t)n * a.estride + c] = v[m];
    }
  };
  auto to_lds = [&](const float2 (&v)[C::M]) {
#pragma unroll
    for (int m = 0; m < C::M; ++m) {
      const int e = tid + m * C::NT;
      if (e < N * L) buf[(e / L) * C::LP + e % L] = v[m];
    }
  };
  auto from_lds = [&](float2 (&v)[C::M]) {
#pragma unroll
    for (int m = 0; m < C::M; ++m) {
      const int e = tid + m * C::NT;
      if (e < N * L) v[m] = buf[(e / L) * C::LP + e % L];
    }
  };

  float2 v[C::M];
  if (a.mode < 2) {
    for (int q = 0; q < a.nbuf; ++q) {
      float2* bp = fl.spec[a.bufs[q]];
      load(bp, v);
      to_lds(v);
      __syncthreads();
      if (a.mode == 0) fft_lds<N, L, C::NT, -1>(buf, tw, tid);
      else fft_lds<N, L, C::NT, +1>(buf, tw, tid);
      from_lds(v);
      __syncthreads();
      store(bp, v);
    }
    return;
  }
  // forward, multiply, inverse (outermost axis = axis 0)
  const float kap = inst[b].kap;
  float2 st[C::M];
  int cur_in = -1;
  for (int j = 0; j < a.njob; ++j) {
    const XJob jb = a.job[j];
    if (jb.in != cur_in) {
      cur_in = jb.in;
      load(fl.spec[jb.in], v);
      to_lds(v);
      __syncthreads();
      fft_lds<N, L, C::NT, -1>(buf, tw, tid);
      from_lds(st);
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < C::M; ++m) {
      const int e = tid + m * C::NT;
      if (e >= N * L) { v[m] = make_float2(0.f, 0.f); continue; }
      const int c = e % L, n = e / L;
      const int col = tile * L + c;
      int ix[3];
      ix[0] = n;
      if (g.nd == 2) { ix[1] = col < g.n[1] ? col : 0; ix[2] = 0; }
      else { ix[1] = o; ix[2] = col < g.nh ? col : 0; }
      float k2 = 0.f;
      for (int i = 0; i < g.nd; ++i) { const float kk = ops.k[i][ix[i]]; k2 += kk * kk; }
      const float km = sqrtf(k2);
      float2 op;
      switch (jb.kind) {
        case XK_GRAD: op = cscale(ops.grad[jb.axis][ix[jb.axis]], kap > 0.f ? sincf(kap * km) : 1.f); break;
        case XK_DIV: op = cscale(ops.div[jb.axis][ix[jb.axis]], kap > 0.f ? sincf(kap * km) : 1.f); break;
        case XK_FRAC1: op = make_float2(km == 0.f ? 0.f : powf(km, ops.frac1), 0.f); break;
        case XK_FRAC2: op = make_float2(km == 0.f ? 0.f : powf(km, ops.frac2), 0.f); break;
        case XK_SRCKAPPA: op = make_float2(kap > 0.f ? cosf(kap * km) : 1.f, 0.f); break;
        case XK_UNSTAGGER: op = ops.unst[jb.axis][ix[jb.axis]]; break;
        default: op = make_float2(1.f, 0.f); break;
      }
      v[m] = cscale(cmul(st[m], op), ops.inv_ntot);
    }
    to_lds(v);
    __syncthreads();
    fft_lds<N, L, C::NT, +1>(buf, tw, tid);
    from_lds(v);
    __syncthreads();
    store(fl.spec[jb.out], v);
  }
}

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
  for (int i = 0; i < g.nd; ++i)
    if (I.rec_u[i]) I.rec_u[i][row] = sample(fl.u[i]);
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
// launchers
// ---------------------------------------------------------------------------
#define KW_FOR_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(24) X(48) X(96) X(192) X(384) X(768) X(40) X(80) X(160) X(320) X(640) X(56) X(112) X(168) X(224) X(336)

bool size_supported(int n) {
  switch (n) {
#define KW_CASE(S) case S: return true;
    KW_FOR_SIZES(KW_CASE)
#undef KW_CASE
    default: return false;
  }
}

static inline dim3 zgrid(const Geom& g, int lpairs, int nb, int nz) {
  const int npairs = (g.rows + 1) / 2;
  return dim3((npairs + lpairs - 1) / lpairs, nb, nz);
}

#define KW_ZDISPATCH(NAME, ...)                                            \
  switch (g.nz) {                                                          \
    KW_FOR_SIZES(NAME)                                                     \
    default: return hipErrorInvalidValue;                                  \
  }

hipError_t launch_zfwd(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw, hipStream_t st) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_zfwd<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(256), 0, st, g, f, fs, spec, ss, tw); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t launch_zinv(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale, hipStream_t st) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_zinv<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(256), 0, st, g, spec, ss, f, fs, tw, scale); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t launch_momentum(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_momentum<S>, zgrid(g, ZCfg<S>::L, nb, g.nd), dim3(256), 0, st, g, fl, inst, tw); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t launch_mass(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st) {
  if (g.nd == 2) {
#define KW_C(S) case S: hipLaunchKernelGGL((k_mass<S, 2>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(256), 0, st, g, fl, inst, tw, flags); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  } else {
#define KW_C(S) case S: hipLaunchKernelGGL((k_mass<S, 3>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(256), 0, st, g, fl, inst, tw, flags); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  }
  return hipGetLastError();
}

hipError_t launch_eos(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_eos<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(256), 0, st, g, fl, inst, tw, flags); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t launch_init_u(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_init_u<S>, zgrid(g, ZCfg<S>::L, nb, g.nd), dim3(256), 0, st, g, fl, inst, tw); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t launch_src_build(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int which, hipStream_t st) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_src_build<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(256), 0, st, g, fl, inst, tw, which); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t launch_col(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st) {
  const int64_t work = (int64_t)a.nouter * nb * ((a.ncols + 15) / 16);
  const bool wide = work >= 1024;
#define KW_C(S)                                                                                                     \
  case S:                                                                                                           \
    if (wide) {                                                                                                     \
      const int nt = (a.ncols + 15) / 16;                                                                           \
      hipLaunchKernelGGL((k_col<S, 16>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, 16>::NT), 0, st, g, fl, ops, inst, a); \
    } else {                                                                                                        \
      const int nt = (a.ncols + 3) / 4;                                                                             \
      hipLaunchKernelGGL((k_col<S, 4>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, 4>::NT), 0, st, g, fl, ops, inst, a);   \
    }                                                                                                               \
    break;
  switch (nline) {
    KW_FOR_SIZES(KW_C)
    default: return hipErrorInvalidValue;
  }
#undef KW_C
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
