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
#pragma once
#include "kw_fft.hpp"
#include "kw_launch.hpp"
#include "kw_types.hpp"

namespace kw {

// ---------------------------------------------------------------------------
// z-pass configuration
// ---------------------------------------------------------------------------
constexpr int pow2floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

#ifndef KW_ZPV
#define KW_ZPV 0
#endif
template <int N>
struct ZCfg {
  // radix plan of the row passes: KW_ZPV=2 selects the radix-4 plan (4 elements per
  // thread) for power-of-two lengths
  static constexpr int PV = (KW_ZPV == 2 && (N & (N - 1)) == 0) ? 2 : 0;
  using F = FftShape<N, 1, PV>;
  static constexpr int E = F::E;
  static constexpr int TPL = F::TPL;
  static constexpr int L0 = clampi(pow2floor(2048 / N), 1, 64);
  static constexpr int L = (L0 * TPL < 64) ? clampi(64 / TPL, L0, 64) : L0;  // row pairs per workgroup
  static constexpr int NT = TPL * L;
  static constexpr int NH = N / 2 + 1;
  using Idx = RowIdx<N>;
  static constexpr int LDS = Idx::size(L);
};

// Global-memory views of pointers.  Pointers read from parameter blocks are
// generic to the compiler, and generic (flat) loads also count against the LDS
// wait counter, so every LDS wait inside a transform would drain them too; the
// global view keeps them on the vector-memory counter alone.
#define KW_AS1 __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const KW_AS1 T* gp(const T* p) { return (const KW_AS1 T*)p; }
template <class T>
__device__ __forceinline__ KW_AS1 T* gw(T* p) { return (KW_AS1 T*)p; }
// float2 through a global view, by components (the vector type's members take
// generic references)
__device__ __forceinline__ float2 ld2(const KW_AS1 float2* p, int64_t i) {
  const KW_AS1 float2& v = p[i];
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st2(KW_AS1 float2* p, int64_t i, float2 v) {
  KW_AS1 float2& d = p[i];
  d.x = v.x;
  d.y = v.y;
}

__device__ __forceinline__ float ld(const float* p, int64_t i, int s) { return gp(p)[i * s]; }

// absorption / dispersion coefficient at idx given c0^2 there
__device__ __forceinline__ float tau_at(const Inst& I, int64_t idx, float c2) {
  return I.tau ? ld(I.tau, idx, I.tau_s) : I.tau_c * exp2f(I.ym1h * log2f(c2));
}
__device__ __forceinline__ float eta_at(const Inst& I, int64_t idx, float c2) {
  return I.eta ? ld(I.eta, idx, I.eta_s) : I.eta_c * exp2f(I.yh * log2f(c2));
}

// pml value of axis `ax` at grid row r, contiguous position n
__device__ __forceinline__ float pml_at(const float* const* tab, const Geom& g, int ax, int r, int n) {
  if (ax == g.nd - 1) return gp(tab[ax])[n];
  if (ax == 0) return gp(tab[0])[g.nd == 2 ? r : r / g.n[1]];
  return gp(tab[1])[r % g.n[1]];
}

// Load the half spectra of the L row pairs of this workgroup into LDS as the
// full spectrum of z = a + i b (Hermitian extension of each half spectrum).
template <int N>
__device__ __forceinline__ void load_pair_spec(float2* lds, const float2* __restrict__ spec_, int row0, int rows, int pitch, int tid) {
  using C = ZCfg<N>;
  const auto spec = gp(spec_);
  const typename C::Idx idx{};
  // all of this thread's loads are issued before the first LDS write, so their
  // latencies overlap instead of adding up
  constexpr int IT = (C::L * C::NH + C::NT - 1) / C::NT;
  float2 As[IT], Bs[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * C::NT;
    const int p = e / C::NH, k = e % C::NH;
    const int ra = row0 + 2 * p, rb = ra + 1;
    const bool in = e < C::L * C::NH;
    As[it] = in && ra < rows ? ld2(spec, (int64_t)ra * pitch + k) : make_float2(0.f, 0.f);
    Bs[it] = in && rb < rows ? ld2(spec, (int64_t)rb * pitch + k) : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = tid + it * C::NT;
    if (e >= C::L * C::NH) break;
    const int p = e / C::NH, k = e % C::NH;
    const float2 A = As[it], B = Bs[it];
    if (k == 0 || (N % 2 == 0 && k == N / 2)) {
      lds[idx(k, p)] = make_float2(A.x, B.x);
    } else {
      lds[idx(k, p)] = make_float2(A.x - B.y, A.y + B.x);
      lds[idx(N - k, p)] = make_float2(A.x + B.y, B.x - A.y);
    }
  }
}

// After a forward FFT of z = a + i b (in LDS), split into the two half spectra.
template <int N>
__device__ __forceinline__ void store_pair_spec(const float2* lds, float2* __restrict__ spec_, int row0, int rows, int pitch, int tid) {
  using C = ZCfg<N>;
  const auto spec = gw(spec_);
  const typename C::Idx idx{};
  for (int e = tid; e < C::L * C::NH; e += C::NT) {
    const int p = e / C::NH, k = e % C::NH;
    const int ra = row0 + 2 * p, rb = ra + 1;
    const float2 Z = lds[idx(k, p)];
    const float2 Zc = lds[idx((N - k) % N, p)];
    if (ra < rows) st2(spec, (int64_t)ra * pitch + k, make_float2(0.5f * (Z.x + Zc.x), 0.5f * (Z.y - Zc.y)));
    if (rb < rows) st2(spec, (int64_t)rb * pitch + k, make_float2(0.5f * (Z.y + Zc.y), -0.5f * (Z.x - Zc.x)));
  }
}

// half spectra of a row pair -> real values of both rows at the owned positions
template <int N>
__device__ __forceinline__ void inv_to_regs(float2* lds, const float2* __restrict__ spec, const float2* __restrict__ tw,
                                            int row0, int rows, int pitch, int tid, float2 (&out)[ZCfg<N>::E]) {
  using C = ZCfg<N>;
  const typename C::Idx idx{};
  const int j0 = tid % C::TPL, c = tid / C::TPL;
  __syncthreads();
  load_pair_spec<N>(lds, spec, row0, rows, pitch, tid);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < C::E; ++m) out[m] = lds[idx(j0 + m * C::TPL, c)];
  __syncthreads();
  fft_reg<N, +1, typename C::Idx, 1, C::PV>(out, lds, tw, c, j0, idx);
}

// real values of a row pair at the owned positions -> half spectra of both rows
template <int N>
__device__ __forceinline__ void fwd_from_regs(float2* lds, float2 (&v)[ZCfg<N>::E], float2* __restrict__ spec,
                                              const float2* __restrict__ tw, int row0, int rows, int pitch, int tid) {
  using C = ZCfg<N>;
  const typename C::Idx idx{};
  const int j0 = tid % C::TPL, c = tid / C::TPL;
  fft_reg<N, -1, typename C::Idx, 1, C::PV>(v, lds, tw, c, j0, idx);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < C::E; ++m) lds[idx(j0 + m * C::TPL, c)] = v[m];
  __syncthreads();
  store_pair_spec<N>(lds, spec, row0, rows, pitch, tid);
  // leave the buffer free: a following transform writes it in its first stage
  // while slower waves could still be reading the pair spectra here
  __syncthreads();
}

// Source value added/imposed at (row r, position n) for a CSR source term.
// Returns true when (r, n) is a source point; v receives scale * signal(t).
__device__ __forceinline__ bool csr_lookup(const Src& s, int r, int n, int t, float& v) {
  int lo = gp(s.row_ptr)[r], hi = gp(s.row_ptr)[r + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int c = gp(s.col)[mid];
    if (c == n) {
      v = gp(s.scale)[mid] * gp(s.sig)[(int64_t)gp(s.sigrow)[mid] * s.len + t];
      return true;
    }
    if (c < n) lo = mid + 1; else hi = mid;
  }
  return false;
}

// Source terms applied inside the unrolled element loops: the dense filtered
// pattern (PATTERN, value sig(t) * pattern) and the inverse-transformed
// multi-row source held in registers (CHAIN, value chv).  mode is SRC_NONE once
// the signal has ended.  The point-source modes (DIRICHLET, NOCORR) are applied
// afterwards by a rolled pass over the owned source points (csr_apply), which
// keeps their lookup out of the unrolled code.
__device__ __forceinline__ int src_mode_at(const Src& s, int t) { return t < s.len ? s.mode : SRC_NONE; }
__device__ __forceinline__ float src_inline(const Src& s, int mode, float f, int64_t idx, float fac, float sig_t, float chv) {
  if (mode == SRC_PATTERN) return f + fac * (sig_t * gp(s.pattern)[idx]);
  if (mode == SRC_CHAIN) return f + fac * chv;
  return f;
}
__device__ __forceinline__ bool src_is_point(int mode) { return mode == SRC_DIRICHLET || mode == SRC_NOCORR; }
__device__ __forceinline__ float src_point(int mode, float f, float fac, float v) { return mode == SRC_DIRICHLET ? fac * v : f + fac * v; }

// Register-array element (m, h) with a run-time m (selects, no indexed registers).
template <int E>
__device__ __forceinline__ float get_lane(const float2 (&a)[E], int m, int h) {
  float v = 0.f;
#pragma unroll
  for (int k = 0; k < E; ++k)
    if (k == m) v = h ? a[k].y : a[k].x;
  return v;
}
template <int E>
__device__ __forceinline__ void set_lane(float2 (&a)[E], int m, int h, float v) {
#pragma unroll
  for (int k = 0; k < E; ++k)
    if (k == m) { if (h) a[k].y = v; else a[k].x = v; }
}

// Minimum waves per SIMD the register allocator must leave room for (kernels
// with short register arrays only; 1 = unconstrained).  Measured at 512^3: every
// floor that forces spills loses.  With the batched spectrum loads the mass kernel
// runs 2409 us unconstrained (2 waves) against 2595 us forced to 3 (29 spilled
// registers); the momentum kernel ran 1904 us at its natural 3 waves and 2808 us
// forced to 4.
#ifndef KW_WPE_MOM
#define KW_WPE_MOM 1
#endif
#ifndef KW_WPE_MASS
#define KW_WPE_MASS 1
#endif
#ifndef KW_WPE_EOS
#define KW_WPE_EOS 1
#endif
#ifndef KW_WPE_COL
#define KW_WPE_COL 1
#endif
#ifndef KW_WPE_COLOPS
#define KW_WPE_COLOPS 4  // the prefetch of the next job input would take it to 130 VGPRs, 3 waves
#endif
#define KW_WPE(N, W) __attribute__((amdgpu_waves_per_eu((FftShape<N>::E <= 8 ? (W) : 1))))

// Stage-ordered twiddles (kw_fft.hpp stage_tw_*) of an N-point plan from the
// natural table src[m] = exp(-2 pi i m / N), one stage at a time so that every
// index is a compile-time expression of the entry.
template <int N, int PV, int S = 1>
__device__ __forceinline__ void fill_stage_tw(float2* dst, const float2* src, int tid, int nt) {
  if constexpr (S < plan_nstages(N, PV)) {
    constexpr int R = plan_stage_radix(N, S, PV), NS = plan_ns_before(N, S, PV);
    constexpr int B = stage_tw_base<N, PV>(S), TWS = N / (NS * R);
    for (int e = tid; e < NS * (R - 1); e += nt) dst[B + e] = ld2(gp(src), (e % (R - 1) + 1) * (e / (R - 1)) * TWS);
    fill_stage_tw<N, PV, S + 1>(dst, src, tid, nt);
  }
}

#define KW_ZPASS_PROLOGUE                                              \
  using C = ZCfg<N>;                                                   \
  __shared__ float2 lds[C::LDS];                                       \
  __shared__ float2 twl[stage_tw_size<N, C::PV>()];                    \
  const int tid = threadIdx.x;                                         \
  fill_stage_tw<N, C::PV>(twl, tw_g, tid, C::NT);                      \
  __syncthreads();                                                     \
  const float2* tw = twl;                                              \
  const int b = blockIdx.y;                                            \
  const int row0 = blockIdx.x * 2 * C::L;                              \
  const int rows = g.rows;                                             \
  const int j0 = tid % C::TPL, pc = tid / C::TPL;                      \
  const int ra = row0 + 2 * pc;                                        \
  (void)j0; (void)ra;

// ---------------------------------------------------------------------------
// generic real <-> half-spectrum transforms along the contiguous axis
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_zfwd(Geom g, const float* __restrict__ f, int64_t fstride, float2* __restrict__ spec,
                                                      int64_t sstride, const float2* __restrict__ tw_g) {
  KW_ZPASS_PROLOGUE
  const auto fb = gp(f + b * fstride);
  float2 v[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int n = j0 + m * C::TPL;
    v[m] = make_float2(ra < rows ? fb[(int64_t)ra * N + n] : 0.f, ra + 1 < rows ? fb[(int64_t)(ra + 1) * N + n] : 0.f);
  }
  fwd_from_regs<N>(lds, v, spec + b * sstride, tw, row0, rows, g.nhp, tid);
}

template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_zinv(Geom g, const float2* __restrict__ spec, int64_t sstride, float* __restrict__ f,
                                                      int64_t fstride, const float2* __restrict__ tw_g, float scale) {
  KW_ZPASS_PROLOGUE
  const auto fb = gw(f + b * fstride);
  float2 v[C::E];
  inv_to_regs<N>(lds, spec + b * sstride, tw, row0, rows, g.nhp, tid, v);
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int n = j0 + m * C::TPL;
    if (ra < rows) fb[(int64_t)ra * N + n] = scale * v[m].x;
    if (ra + 1 < rows) fb[(int64_t)(ra + 1) * N + n] = scale * v[m].y;
  }
}

// Element addressing shared by the step kernels.  A thread owns positions
// n = j0 + m * TPL of the two rows ra, ra + 1 of its pair.  Row counts are even,
// so both rows of a pair are inside the grid or both outside (tail workgroups of
// small grids); loads use rows clamped into the grid so that the unrolled
// element loops carry no branches and their loads issue together, and only the
// stores are guarded.  Values computed for rows outside the grid are never stored.
#define KW_ROWS_CLAMPED                                                          \
  const bool inside = ra < rows;                                                 \
  const int rA = min(ra, rows - 2);                                              \
  const int64_t iA = (int64_t)rA * N + j0; /* element (m, h): iA + h N + m TPL */

// pml multipliers of the rows of a thread along axis ax (1 for the contiguous axis)
__device__ __forceinline__ void pml_rows(const float* const* tab, const Geom& g, int ax, int rA, float (&ph)[2]) {
  if (ax == g.nd - 1) { ph[0] = ph[1] = 1.f; return; }
  ph[0] = pml_at(tab, g, ax, rA, 0);
  ph[1] = pml_at(tab, g, ax, rA + 1, 0);
}

// ---------------------------------------------------------------------------
// momentum: u_i = pml_sg (pml_sg u_i - dt/rho0_sg * grad_i p) (+ source), then
// forward transform of u_i along z (for the divergence).  blockIdx.z = axis.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) KW_WPE(N, KW_WPE_MOM) void k_momentum(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw_g) {
  KW_ZPASS_PROLOGUE
  const int ax = blockIdx.z;
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  if (t >= I.nt) return;
  // pml along the contiguous axis, staged next to the twiddles (read after the
  // first barrier of the transform below)
  __shared__ float pmlz[N];
  const bool zax = ax == g.nd - 1;
  if (zax)
    for (int i_ = tid; i_ < N; i_ += C::NT) pmlz[i_] = gp(I.pmlsg[ax])[i_];
  KW_ROWS_CLAMPED
  float2* sp = fl.spec[ax] + b * g.nspec;
  const auto u = gw(fl.u[ax] + b * g.ntot);
  const Src& src = I.srcu[ax];
  float ph[2];
  pml_rows(I.pmlsg, g, ax, rA, ph);
  // operands of the update, fetched before the transform so their latency overlaps it
  float2 uo[C::E], dr[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int64_t i0 = iA + m * C::TPL;
    uo[m] = make_float2(u[i0], u[i0 + N]);
    dr[m] = make_float2(ld(I.dtr0[ax], i0, I.dtr0_s), ld(I.dtr0[ax], i0 + N, I.dtr0_s));
  }
  float2 gv[C::E];
  inv_to_regs<N>(lds, sp, tw, row0, rows, g.nhp, tid, gv);
  float2 ch[C::E];
  const int smode = src_mode_at(src, t);
  const bool chain = smode == SRC_CHAIN;
  if (chain) inv_to_regs<N>(lds, fl.spec[src.chain_buf] + b * g.nspec, tw, row0, rows, g.nhp, tid, ch);
  const float sig_t = smode == SRC_PATTERN ? gp(src.sig)[t] : 0.f;
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const float sz = zax ? pmlz[j0 + m * C::TPL] : 1.f;
    float res[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t idx = iA + h * N + m * C::TPL;
      const float s = sz * ph[h];
      const float gval = h ? gv[m].y : gv[m].x;
      float uv = s * (s * (h ? uo[m].y : uo[m].x) - (h ? dr[m].y : dr[m].x) * gval);
      uv = src_inline(src, smode, uv, idx, src.fac[ax], sig_t, chain ? (h ? ch[m].y : ch[m].x) : 0.f);
      res[h] = uv;
    }
    if (inside) {
      u[iA + m * C::TPL] = res[0];
      u[iA + N + m * C::TPL] = res[1];
    }
    gv[m] = make_float2(res[0], res[1]);
  }
  if (src_is_point(smode) && inside) {
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (gp(src.row_ptr)[r] == gp(src.row_ptr)[r + 1]) continue;
      for (int m = 0; m < C::E; ++m) {
        const int n = j0 + m * C::TPL;
        float v;
        if (!csr_lookup(src, r, n, t, v)) continue;
        const float uv = src_point(smode, get_lane(gv, m, h), src.fac[ax], v);
        u[(int64_t)r * N + n] = uv;
        set_lane(gv, m, h, uv);
      }
    }
  }
  fwd_from_regs<N>(lds, gv, sp, tw, row0, rows, g.nhp, tid);
}

// ---------------------------------------------------------------------------
// mass conservation (+ equation of state when no fractional Laplacian is needed)
// SREG: the pressure source term is staged in registers once per element
// (CHAIN: the inverse transform of its filtered spectrum; PATTERN: sig(t) times
// the dense filtered pattern) and added to every split density from there.
// The other instantiation serves plans without such a source and the steps
// after the signal has ended.
// CAP: the adjoint replay of a nonlinear plan also keeps each axis's Dop_i u_i
// (Fields::adjD).  A template switch rather than a runtime test, so the
// stepping kernels' code, and with it their FP32 contractions, stays unchanged.
// ---------------------------------------------------------------------------
template <int N, int ND, bool SREG, bool CAP>
__global__ __launch_bounds__(ZCfg<N>::NT) KW_WPE(N, KW_WPE_MASS) void k_mass(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw_g, int flags) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  if (t >= I.nt) return;
  __shared__ float pmlz[N];  // pml along the contiguous axis (read after the first transform barrier)
  for (int i_ = tid; i_ < N; i_ += C::NT) pmlz[i_] = gp(I.pml[ND - 1])[i_];
  KW_ROWS_CLAMPED
  const int64_t off = b * g.ntot;
  const Src& src = I.srcp;
  const int smode = src_mode_at(src, t);
  const bool chain = SREG && smode == SRC_CHAIN;
  const bool patt = SREG && smode == SRC_PATTERN;
  const float sig_t = smode == SRC_PATTERN ? gp(src.sig)[t] : 0.f;
  const bool powerlaw = flags & (FLAG_POWERLAW_ABS | FLAG_POWERLAW_DISP);
  float ph[ND][2];
#pragma unroll
  for (int i = 0; i < ND; ++i) pml_rows(I.pml, g, i, rA, ph[i]);
  // per owned element and row: the coefficient dt (2 rho_old + rho0) of the
  // density update (dt rho0 when linear; rho_old = sum of the split densities
  // of the previous step, kept in rhot), and the running sums of div u and of
  // the new densities
  // rho0 stays in registers for the outputs at the end where that costs no occupancy (3-D,
  // power-of-two lengths: two waves per SIMD either way).  A re-read there misses L2 at
  // large grids, about 4 bytes per voxel of extra HBM traffic; the 2-D kernels would drop
  // from three waves per SIMD to two, the mixed-radix ones from two to one.
  constexpr bool R0R = ND == 3 && C::E <= 16 && (N & (N - 1)) == 0;
  float2 qv[C::E], divu[C::E], rt[C::E], r0v[R0R ? C::E : 1];
  float2 ch[SREG ? C::E : 1];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int64_t i0 = iA + m * C::TPL;
    float qa = ld(I.rho0, i0, I.rho0_s), qb = ld(I.rho0, i0 + N, I.rho0_s);
    if constexpr (R0R) r0v[m] = make_float2(qa, qb);
    if (flags & FLAG_NONLINEAR) {
      qa += 2.f * gw(fl.rhot)[off + i0];
      qb += 2.f * gw(fl.rhot)[off + i0 + N];
    }
    qv[m] = make_float2(I.dt * qa, I.dt * qb);
    divu[m] = make_float2(0.f, 0.f);
    rt[m] = make_float2(0.f, 0.f);
  }
  if constexpr (SREG) {
    if (chain) {
      inv_to_regs<N>(lds, fl.spec[src.chain_buf] + b * g.nspec, tw, row0, rows, g.nhp, tid, ch);
    } else if (patt) {
      const auto pat = gp(src.pattern);
#pragma unroll
      for (int m = 0; m < C::E; ++m) ch[m] = make_float2(sig_t * pat[iA + m * C::TPL], sig_t * pat[iA + N + m * C::TPL]);
    }
  }
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const auto rho = gw(fl.rho[i] + off);
    // old split density, fetched before the transform so its latency overlaps it
    // (for short register arrays; the long mixed-radix ones load it in place)
    constexpr bool PF = C::E <= 16;
    float2 ro[PF ? C::E : 1];
    if constexpr (PF) {
#pragma unroll
      for (int m = 0; m < C::E; ++m) ro[m] = make_float2(rho[iA + m * C::TPL], rho[iA + N + m * C::TPL]);
    }
    float2 d[C::E];
    inv_to_regs<N>(lds, fl.spec[i] + b * g.nspec, tw, row0, rows, g.nhp, tid, d);
    if constexpr (CAP) {
      if (inside) {
        const auto dd = gw(fl.adjD + (int64_t)i * gridDim.y * g.ntot + off);
#pragma unroll
        for (int m = 0; m < C::E; ++m) {
          dd[iA + m * C::TPL] = d[m].x;
          dd[iA + N + m * C::TPL] = d[m].y;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      const float sz = (i == ND - 1) ? pmlz[j0 + m * C::TPL] : 1.f;
      float res[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t idx = iA + h * N + m * C::TPL;
        const float dv = h ? d[m].y : d[m].x;
        const float s = sz * ph[i][h];
        const float q = h ? qv[m].y : qv[m].x;
        float rold;
        if constexpr (PF) rold = h ? ro[m].y : ro[m].x;
        else rold = rho[idx];
        float rv = s * (s * rold - q * dv);
        if constexpr (SREG) {
          if (chain || patt) rv = rv + src.fac[i] * (h ? ch[m].y : ch[m].x);
        } else {
          rv = src_inline(src, smode, rv, idx, src.fac[i], sig_t, 0.f);
        }
        res[h] = rv;
      }
      if (inside) {
        rho[iA + m * C::TPL] = res[0];
        rho[iA + N + m * C::TPL] = res[1];
      }
      divu[m].x += d[m].x;
      divu[m].y += d[m].y;
      rt[m].x += res[0];
      rt[m].y += res[1];
    }
  }
  if (src_is_point(smode) && inside) {
    // point sources: update the owned source points of every split density and
    // rebuild their density sums (same order of additions as the loop above)
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (gp(src.row_ptr)[r] == gp(src.row_ptr)[r + 1]) continue;
      for (int m = 0; m < C::E; ++m) {
        const int n = j0 + m * C::TPL;
        float v;
        if (!csr_lookup(src, r, n, t, v)) continue;
        const int64_t idx = off + (int64_t)r * N + n;
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < ND; ++i) {
          const auto rho = gw(fl.rho[i]);
          const float x = src_point(smode, rho[idx], src.fac[i], v);
          rho[idx] = x;
          sum += x;
        }
        set_lane(rt, m, h, sum);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    float a_out[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t idx = iA + h * N + m * C::TPL;
      float r0;
      if constexpr (R0R) r0 = h ? r0v[m].y : r0v[m].x;
      else r0 = ld(I.rho0, idx, I.rho0_s);  // the lines this block read at its start, from L2 if still there
      const float du = h ? divu[m].y : divu[m].x;
      const float rtv = h ? rt[m].y : rt[m].x;
      if (powerlaw) {
        a_out[h] = r0 * du;
      } else {
        float q = rtv;
        const float c2 = ld(I.c0sq, idx, I.c0sq_s);
        if (flags & FLAG_STOKES) {
          q += tau_at(I, idx, c2) * r0 * du;
          if (fl.adjA && inside) gw(fl.adjA)[off + idx] = r0 * du;  // adjoint replay: the Stokes operand
        }
        if (flags & FLAG_NONLINEAR) q += ld(I.bona, idx, I.bona_s) * rtv * rtv / (2.f * r0);
        a_out[h] = c2 * q;
      }
    }
    if (inside) {
      const int64_t i0 = off + iA + m * C::TPL;
      if (fl.rhot) {
        gw(fl.rhot)[i0] = rt[m].x;
        gw(fl.rhot)[i0 + N] = rt[m].y;
      }
      if (!powerlaw) {
        gw(fl.p)[i0] = a_out[0];
        gw(fl.p)[i0 + N] = a_out[1];
      }
    }
    divu[m] = make_float2(a_out[0], a_out[1]);
  }
  if (!powerlaw) {
    fwd_from_regs<N>(lds, divu, fl.spec[0] + b * g.nspec, tw, row0, rows, g.nhp, tid);
  } else {
    if (flags & FLAG_POWERLAW_ABS) fwd_from_regs<N>(lds, divu, fl.spec[0] + b * g.nspec, tw, row0, rows, g.nhp, tid);
    if (flags & FLAG_POWERLAW_DISP) fwd_from_regs<N>(lds, rt, fl.spec[1] + b * g.nspec, tw, row0, rows, g.nhp, tid);
  }
}

// ---------------------------------------------------------------------------
// equation of state with power-law absorption / dispersion
// p = c0^2 (rho + tau F^-1[|k|^{y-2} F[rho0 div u]] - eta F^-1[|k|^{y-1} F[rho]] + B/A rho^2 / (2 rho0))
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) KW_WPE(N, KW_WPE_EOS) void k_eos(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw_g, int flags) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  if (t >= I.nt) return;
  KW_ROWS_CLAMPED
  const int64_t off = b * g.ntot;
  const auto rhot = gw(fl.rhot + off);
  float2 rtv[C::E], c2v[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int64_t i0 = iA + m * C::TPL;
    rtv[m] = make_float2(rhot[i0], rhot[i0 + N]);
    c2v[m] = make_float2(ld(I.c0sq, i0, I.c0sq_s), ld(I.c0sq, i0 + N, I.c0sq_s));
  }
  float2 av[C::E], bv[C::E];
  if (flags & FLAG_POWERLAW_ABS) inv_to_regs<N>(lds, fl.spec[0] + b * g.nspec, tw, row0, rows, g.nhp, tid, av);
  if (flags & FLAG_POWERLAW_DISP) inv_to_regs<N>(lds, fl.spec[1] + b * g.nspec, tw, row0, rows, g.nhp, tid, bv);
  if (inside && fl.adjA && (flags & FLAG_POWERLAW_ABS)) {  // adjoint replay: keep the operands
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      gw(fl.adjA)[off + iA + m * C::TPL] = av[m].x;
      gw(fl.adjA)[off + iA + N + m * C::TPL] = av[m].y;
    }
  }
  if (inside && fl.adjB && (flags & FLAG_POWERLAW_DISP)) {
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      gw(fl.adjB)[off + iA + m * C::TPL] = bv[m].x;
      gw(fl.adjB)[off + iA + N + m * C::TPL] = bv[m].y;
    }
  }
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    float out[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t idx = iA + h * N + m * C::TPL;
      const float rt = h ? rtv[m].y : rtv[m].x;
      const float c2 = h ? c2v[m].y : c2v[m].x;
      float q = rt;
      if (flags & FLAG_POWERLAW_ABS) q += tau_at(I, idx, c2) * (h ? av[m].y : av[m].x);
      if (flags & FLAG_POWERLAW_DISP) q -= eta_at(I, idx, c2) * (h ? bv[m].y : bv[m].x);
      if (flags & FLAG_NONLINEAR) q += ld(I.bona, idx, I.bona_s) * rt * rt / (2.f * ld(I.rho0, idx, I.rho0_s));
      out[h] = c2 * q;
    }
    if (inside) {
      gw(fl.p)[off + iA + m * C::TPL] = out[0];
      gw(fl.p)[off + iA + N + m * C::TPL] = out[1];
    }
    av[m] = make_float2(out[0], out[1]);
  }
  fwd_from_regs<N>(lds, av, fl.spec[0] + b * g.nspec, tw, row0, rows, g.nhp, tid);
}

// ---------------------------------------------------------------------------
// t = 0 initial-value override: u_i = dt/(2 rho0_sg_i) grad_i p0, rho_j = p0 / (c0^2 D)
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_init_u(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw_g) {
  KW_ZPASS_PROLOGUE
  const int ax = blockIdx.z;
  const Inst& I = inst[b];
  const int64_t off = b * g.ntot;
  float2 gv[C::E];
  inv_to_regs<N>(lds, fl.spec[ax] + b * g.nspec, tw, row0, rows, g.nhp, tid, gv);
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int n = j0 + m * C::TPL;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const int64_t idx = (int64_t)r * N + n;
      gw(fl.u[ax])[off + idx] = 0.5f * ld(I.dtr0[ax], idx, I.dtr0_s) * (h ? gv[m].y : gv[m].x);
      if (ax == 0) {
        const float pv = gw(fl.p)[off + idx], c2 = ld(I.c0sq, idx, I.c0sq_s);
        float rs = 0.f;
        for (int j = 0; j < g.nd; ++j) {
          const float rv = pv / (c2 * I.p0div[j]);
          gw(fl.rho[j])[off + idx] = rv;
          rs += rv;
        }
        if (fl.rhot) gw(fl.rhot)[off + idx] = rs;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// build the z-spectrum of a multi-row additive source at step t (mode CHAIN)
// which = -1: pressure source, 0..2: velocity component
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_src_build(Geom g, Fields fl, const Inst* __restrict__ inst, const float2* __restrict__ tw_g, int which) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  const Src& s = which < 0 ? I.srcp : I.srcu[which];
  if (t >= I.nt || s.mode != SRC_CHAIN || t >= s.len) return;
  const auto sp = gw(fl.spec[s.chain_buf] + b * g.nspec);
  const int rend = min(row0 + 2 * C::L, rows);
  const int e0 = gp(s.row_ptr)[row0], e1 = gp(s.row_ptr)[rend];
  if (e0 == e1) {
    for (int e = tid; e < C::L * C::NH; e += C::NT) {
      const int p = e / C::NH, k = e % C::NH;
      const int r0 = row0 + 2 * p;
      if (r0 < rows) st2(sp, (int64_t)r0 * g.nhp + k, make_float2(0.f, 0.f));
      if (r0 + 1 < rows) st2(sp, (int64_t)(r0 + 1) * g.nhp + k, make_float2(0.f, 0.f));
    }
    return;
  }
  const typename C::Idx idx{};
  for (int e = tid; e < N * C::L; e += C::NT) lds[idx(e % N, e / N)] = make_float2(0.f, 0.f);
  __syncthreads();
  float* fb = reinterpret_cast<float*>(lds);
  for (int e = e0 + tid; e < e1; e += C::NT) {
    int r = row0;
    while (gp(s.row_ptr)[r + 1] <= e) ++r;
    const int p = (r - row0) >> 1, h = (r - row0) & 1;
    fb[2 * idx(gp(s.col)[e], p) + h] = gp(s.scale)[e] * gp(s.sig)[(int64_t)gp(s.sigrow)[e] * s.len + t];
  }
  __syncthreads();
  float2 v[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) v[m] = lds[idx(j0 + m * C::TPL, pc)];
  __syncthreads();
  fwd_from_regs<N>(lds, v, fl.spec[s.chain_buf] + b * g.nspec, tw, row0, rows, g.nhp, tid);
}

// ---------------------------------------------------------------------------
// strided column passes
// ---------------------------------------------------------------------------
// sin(x)/x and cos(x) by their Taylor series through x^14 (truncation error
// below 2e-10 for |x| <= 2, which covers c_ref dt |k| / 2 at any stable CFL);
// larger arguments fall back to the library functions.
// The k-space operators of the outermost-axis pass take |k|^2 = kx^2 + ky^2 + kz^2
// (no square root on the common path).  kappa = sinc(x) and cos(x) for
// x = c_ref dt |k| / 2 from x^2: Taylor polynomials in x^2 up to x^2 = 4 (FP32
// rounding), sin(x)/x and cos(x) beyond.
__device__ __forceinline__ float sinc_k2(float x2) {
  if (x2 > 4.f) {
    const float x = sqrtf(x2);
    return sinf(x) / x;
  }
  float p = -1.f / 1307674368000.f;                 // -1/15!
  p = fmaf(p, x2, 1.f / 6227020800.f);              //  1/13!
  p = fmaf(p, x2, -1.f / 39916800.f);               // -1/11!
  p = fmaf(p, x2, 1.f / 362880.f);                  //  1/9!
  p = fmaf(p, x2, -1.f / 5040.f);                   // -1/7!
  p = fmaf(p, x2, 1.f / 120.f);                     //  1/5!
  p = fmaf(p, x2, -1.f / 6.f);                      // -1/3!
  return fmaf(p, x2, 1.f);
}
__device__ __forceinline__ float cos_k2(float x2) {
  if (x2 > 4.f) return cosf(sqrtf(x2));
  float p = -1.f / 87178291200.f;                   // -1/14!
  p = fmaf(p, x2, 1.f / 479001600.f);               //  1/12!
  p = fmaf(p, x2, -1.f / 3628800.f);                // -1/10!
  p = fmaf(p, x2, 1.f / 40320.f);                   //  1/8!
  p = fmaf(p, x2, -1.f / 720.f);                    // -1/6!
  p = fmaf(p, x2, 1.f / 24.f);                      //  1/4!
  p = fmaf(p, x2, -0.5f);                           // -1/2!
  return fmaf(p, x2, 1.f);
}
// |k|^a from |k|^2 for |k| > 0 (the fractional Laplacian is 0 at |k| = 0)
__device__ __forceinline__ float kpow2(float km2, float a) { return km2 == 0.f ? 0.f : exp2f(0.5f * a * log2f(km2)); }

// Consecutive workgroup ids go to different XCDs (round robin over 8); give each
// XCD a contiguous range of tiles so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int nfull = nblk & ~7;
  if (bid >= nfull) return bid;
  return (bid & 7) * (nfull >> 3) + (bid >> 3);
}

template <int N, int L, int EM = 1, int PV = 0>
struct ColCfg {
  using F = FftShape<N, EM, PV>;
  static constexpr int E = F::E;
  static constexpr int TPL = F::TPL;
  static constexpr int NT = TPL * L;
  using Idx = ColIdx<L>;
  static constexpr int LDS = Idx::size(N);
};

template <int N>
constexpr int col_lbig() { return clampi(pow2floor(8192 / N), 1, 64); }
template <int N>
constexpr int col_lsmall() { return col_lbig<N>() / 4 < 1 ? 1 : col_lbig<N>() / 4; }
// Tiles of the outermost-axis pass with spectral jobs: half as wide, since that
// kernel holds more registers (at 512: 942 vs 1015 us for the gradient pass,
// 1197 vs 1335 us for the divergence pass).
template <int N>
constexpr int col_lops() { return col_lbig<N>() / 2 < 1 ? 1 : col_lbig<N>() / 2; }

template <int N, int L, bool OPS, int EM = 1, int PV = 0>
__global__ __launch_bounds__((ColCfg<N, L, EM, PV>::NT)) KW_WPE(N, (OPS ? KW_WPE_COLOPS : KW_WPE_COL)) void k_col(Geom g, Fields fl, Ops ops, const Inst* __restrict__ inst, ColArgs a) {
  using C = ColCfg<N, L, EM, PV>;
  __shared__ float2 lds[C::LDS];
  __shared__ float2 twl[stage_tw_size<N, PV>()];
  const typename C::Idx idx{};
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  if (fl.tdev && gp(fl.tdev)[b] >= inst[b].nt) return;
  fill_stage_tw<N, PV>(twl, ops.tw[a.axis], tid, C::NT);
  __syncthreads();
  const int ntile = (a.ncols + L - 1) / L;
  const int lid = xcd_remap(blockIdx.x, gridDim.x);
  const int o = lid / ntile, tile = lid % ntile;
  const int c = tid % L, j0 = tid / L;
  const int col = tile * L + c;
  const bool valid = col < a.ncols;
  const int64_t base = (int64_t)b * g.nspec + (int64_t)o * a.ostride + col;
  const float2* tw = twl;

  float2 v[C::E];
  if constexpr (!OPS) {
    // software pipeline over the buffers and tiles: the loads of the next buffer
    // (after a tile's last buffer, the first buffer of this workgroup's next tile)
    // are in flight while the current one is transformed and stored
    const int ntiles = a.nouter * ntile;
    float2 nx[C::E];
    if (valid) {
      const auto bp = gp(fl.spec[a.bufs[0]]);
#pragma unroll
      for (int m = 0; m < C::E; ++m) nx[m] = ld2(bp, base + (int64_t)(j0 + m * C::TPL) * a.estride);
    }
    for (int lt = lid; lt < ntiles; lt += (int)gridDim.x) {
      const int ot = lt / ntile, colt = (lt % ntile) * L + c;
      const bool vt = colt < a.ncols;
      const int64_t bt = (int64_t)b * g.nspec + (int64_t)ot * a.ostride + colt;
      const int ltn = lt + (int)gridDim.x;
      const int coln = (ltn % ntile) * L + c;
      const bool vn = ltn < ntiles && coln < a.ncols;
      const int64_t bn = (int64_t)b * g.nspec + (int64_t)(ltn / ntile) * a.ostride + coln;
      for (int q = 0; q < a.nbuf; ++q) {
        const auto bp = gw(fl.spec[a.bufs[q]]);
#pragma unroll
        for (int m = 0; m < C::E; ++m) v[m] = vt ? nx[m] : make_float2(0.f, 0.f);
        if (q + 1 < a.nbuf) {
          if (vt) {
            const auto bq = gp(fl.spec[a.bufs[q + 1]]);
#pragma unroll
            for (int m = 0; m < C::E; ++m) nx[m] = ld2(bq, bt + (int64_t)(j0 + m * C::TPL) * a.estride);
          }
        } else if (vn) {
          const auto bq = gp(fl.spec[a.bufs[0]]);
#pragma unroll
          for (int m = 0; m < C::E; ++m) nx[m] = ld2(bq, bn + (int64_t)(j0 + m * C::TPL) * a.estride);
        }
        if (a.mode == 0) fft_reg<N, -1, typename C::Idx, EM, PV>(v, lds, tw, c, j0, idx);
        else fft_reg<N, +1, typename C::Idx, EM, PV>(v, lds, tw, c, j0, idx);
        if (vt) {
#pragma unroll
          for (int m = 0; m < C::E; ++m) st2(bp, bt + (int64_t)(j0 + m * C::TPL) * a.estride, v[m]);
        }
      }
    }
  } else {
  // forward, multiply, inverse (outermost axis = axis 0), over the tiles lid,
  // lid + G, ... of this instance (G = workgroups per instance: one tile each when
  // the grid covers every tile, else a resident grid that loops; with the
  // XCD-aware lid, the tiles an XCD runs at one time are neighbours).
  const float kap = inst[b].kap;
  // Operator values read between the transforms come from LDS, so no global
  // latency sits between a forward and an inverse transform: the x wavenumbers,
  // the x-axis operator vectors of up to two tabled jobs (loaded once), and per
  // column of the tile the value of every y/z-axis job.  (The twiddle barrier
  // above has passed; the first transform's barriers publish these before they
  // are read, and the previous tile's last reads of cop sit before its last
  // transform's barriers.)
  __shared__ float kxs[N];
  __shared__ float2 xop[2][N];
  __shared__ float2 cop[kMaxJobs][L];
  for (int i_ = tid; i_ < N; i_ += C::NT) kxs[i_] = gp(ops.k[0])[i_];
  for (int j = 0, slot = 0; j < a.njob; ++j) {
    const XJob jb = a.job[j];
    const float2* t = jb.kind == XK_GRAD ? ops.grad[jb.axis] : jb.kind == XK_DIV ? ops.div[jb.axis]
                    : jb.kind == XK_UNSTAGGER ? ops.unst[jb.axis] : nullptr;
    if (!t || jb.axis != 0) continue;
    if (slot < 2)
      for (int i_ = tid; i_ < N; i_ += C::NT) xop[slot][i_] = ld2(gp(t), i_);
    ++slot;
  }
  const int ntiles = a.nouter * ntile;
  // The next distinct input is loaded while the current one is transformed and
  // its jobs stored: issued after those stores, its loads would also wait for
  // them (the vector-memory counter drains in order).  Within a tile, not when
  // one of the jobs before it writes that buffer; after a tile's last input, the
  // next tile's first input (other columns, so never written by this tile).
  // Only for short register arrays (the mixed-radix ones would lose occupancy).
  constexpr bool PRE = C::E <= 8;
  float2 nx[PRE ? C::E : 1];
  bool next_loaded = false;  // nx holds this tile's first input
  for (int lt = lid; lt < ntiles; lt += (int)gridDim.x) {
    const int o = lt / ntile, tile = lt % ntile;
    const int col = tile * L + c;
    const bool valid = col < a.ncols;
    const int64_t base = (int64_t)b * g.nspec + (int64_t)o * a.ostride + col;
    int iy, iz;
    if (g.nd == 2) { iy = valid ? col : 0; iz = 0; }
    else { iy = o; iz = valid ? col : 0; }
    for (int j = 0; j < a.njob; ++j) {
      const XJob jb = a.job[j];
      const float2* t = jb.kind == XK_GRAD ? ops.grad[jb.axis] : jb.kind == XK_DIV ? ops.div[jb.axis]
                      : jb.kind == XK_UNSTAGGER ? ops.unst[jb.axis] : nullptr;
      if (!t || jb.axis == 0) continue;
      for (int cc = tid; cc < L; cc += C::NT) {
        const int colc = min(tile * L + cc, a.ncols - 1);
        const int ia = (jb.axis == 1 && g.nd == 3) ? o : colc;
        cop[j][cc] = ld2(gp(t), ia);
      }
    }
    float kyz2 = 0.f;
    {
      const float ky = gp(ops.k[1])[iy];
      kyz2 = ky * ky;
      if (g.nd == 3) { const float kz = gp(ops.k[2])[iz]; kyz2 += kz * kz; }
    }
    float2 st[C::E];
    int cur_in = -1, pre_in = next_loaded ? a.job[0].in : -1;
    bool kgroup = false;  // the current input's spectrum carries kappa / ntot
    next_loaded = false;
    for (int j = 0; j < a.njob; ++j) {
      const XJob jb = a.job[j];
      if (jb.in != cur_in) {
        cur_in = jb.in;
        if (PRE && pre_in == cur_in) {
#pragma unroll
          for (int m = 0; m < C::E; ++m) st[m] = nx[PRE ? m : 0];
        } else {
          const auto src = gp(fl.spec[jb.in]);
#pragma unroll
          for (int m = 0; m < C::E; ++m) st[m] = valid ? ld2(src, base + (int64_t)(j0 + m * C::TPL) * a.estride) : make_float2(0.f, 0.f);
        }
        pre_in = -1;
        int k = j + 1;
        while (k < a.njob && a.job[k].in == cur_in) ++k;
        if (PRE && k < a.njob) {
          bool clobbered = false;
          for (int i = j; i < k; ++i) clobbered |= a.job[i].out == a.job[k].in;
          if (!clobbered) {
            pre_in = a.job[k].in;
            const auto nsrc = gp(fl.spec[pre_in]);
#pragma unroll
            for (int m = 0; m < C::E; ++m) nx[PRE ? m : 0] = valid ? ld2(nsrc, base + (int64_t)(j0 + m * C::TPL) * a.estride) : make_float2(0.f, 0.f);
          }
        } else if (PRE && lt + (int)gridDim.x < ntiles) {
          const int ltn = lt + (int)gridDim.x;
          const int on = ltn / ntile, coln = (ltn % ntile) * L + c;
          const bool validn = coln < a.ncols;
          const int64_t basen = (int64_t)b * g.nspec + (int64_t)on * a.ostride + coln;
          const auto nsrc = gp(fl.spec[a.job[0].in]);
#pragma unroll
          for (int m = 0; m < C::E; ++m) nx[PRE ? m : 0] = validn ? ld2(nsrc, basen + (int64_t)(j0 + m * C::TPL) * a.estride) : make_float2(0.f, 0.f);
          next_loaded = true;
        }
        fft_reg<N, -1, typename C::Idx, EM, PV>(st, lds, tw, c, j0, idx);
        // When every job on this input is a gradient or divergence, the common
        // factor kappa(|k|) / ntot is applied to the spectrum once (the x
        // wavenumbers in LDS are published by the transform's barriers).
        kgroup = true;
        for (int i = j; i < k; ++i) kgroup &= a.job[i].kind == XK_GRAD || a.job[i].kind == XK_DIV;
        if (kgroup) {
#pragma unroll
          for (int m = 0; m < C::E; ++m) {
            const float kx = kxs[j0 + m * C::TPL];
            const float f = kap > 0.f ? sinc_k2(kap * kap * fmaf(kx, kx, kyz2)) : 1.f;
            st[m] = cscale(st[m], f * ops.inv_ntot);
          }
        }
      }
      // where this job's operator values live: an x-table slot, the per-column
      // values, or (a third x-axis table) global memory
      const bool tabled = jb.kind == XK_GRAD || jb.kind == XK_DIV || jb.kind == XK_UNSTAGGER;
      int slot = 0;
      for (int i = 0; i < j; ++i) {
        const XJob q = a.job[i];
        slot += q.axis == 0 && (q.kind == XK_GRAD || q.kind == XK_DIV || q.kind == XK_UNSTAGGER);
      }
      const float2* gtab = jb.kind == XK_GRAD ? ops.grad[0] : jb.kind == XK_DIV ? ops.div[0] : ops.unst[0];
      const float2 colv = (tabled && jb.axis != 0) ? cop[j][c] : make_float2(1.f, 0.f);
#pragma unroll
      for (int m = 0; m < C::E; ++m) {
        const int n = j0 + m * C::TPL;
        float2 op = make_float2(1.f, 0.f);
        if (tabled) op = jb.axis != 0 ? colv : (slot < 2 ? xop[slot][n] : ld2(gp(gtab), n));
        if (kgroup) {
          v[m] = cmul(st[m], op);
          continue;
        }
        const float kx = kxs[n];
        const float km2 = fmaf(kx, kx, kyz2);
        switch (jb.kind) {
          case XK_GRAD:
          case XK_DIV: op = cscale(op, kap > 0.f ? sinc_k2(kap * kap * km2) : 1.f); break;  // kappa(|k|)
          case XK_FRAC1: op = make_float2(kpow2(km2, ops.frac1), 0.f); break;
          case XK_FRAC2: op = make_float2(kpow2(km2, ops.frac2), 0.f); break;
          case XK_SRCKAPPA: op = make_float2(kap > 0.f ? cos_k2(kap * kap * km2) : 1.f, 0.f); break;
          default: break;
        }
        v[m] = cscale(cmul(st[m], op), ops.inv_ntot);
      }
      fft_reg<N, +1, typename C::Idx, EM, PV>(v, lds, tw, c, j0, idx);
      if (valid) {
        const auto dst = gw(fl.spec[jb.out]);
#pragma unroll
        for (int m = 0; m < C::E; ++m) st2(dst, base + (int64_t)(j0 + m * C::TPL) * a.estride, v[m]);
      }
    }
  }
  }
}

// Supported line lengths, in four radix families.  Launcher files are compiled
// once per family (-DKW_GROUP=0..3) so the many template instantiations build
// in parallel; KW_GROUP_SIZES lists the family of the current translation unit
// and KW_GNAME(f) names its launcher f_g<family>.
#define KW_SIZES_G0(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#define KW_SIZES_G1(X) X(24) X(48) X(96) X(192) X(384) X(768)
#define KW_SIZES_G2(X) X(40) X(80) X(160) X(320) X(640)
#define KW_SIZES_G3(X) X(56) X(112) X(224) X(448)
#define KW_FOR_SIZES(X) KW_SIZES_G0(X) KW_SIZES_G1(X) KW_SIZES_G2(X) KW_SIZES_G3(X)
#define KW_CAT_(a, b) a##b
#define KW_CAT(a, b) KW_CAT_(a, b)
#ifdef KW_GROUP
#define KW_GROUP_SIZES KW_CAT(KW_SIZES_G, KW_GROUP)
#define KW_GNAME(f) KW_CAT(f##_g, KW_GROUP)
#endif

inline bool size_supported_impl(int n) {
  switch (n) {
#define KW_CASE(S) case S: return true;
    KW_FOR_SIZES(KW_CASE)
#undef KW_CASE
    default: return false;
  }
}

inline dim3 zgrid(const Geom& g, int lpairs, int nb, int nz) {
  const int npairs = (g.rows + 1) / 2;
  return dim3((npairs + lpairs - 1) / lpairs, nb, nz);
}

// switch over the current family's sizes; a size outside it leaves *handled false
#define KW_ZDISPATCH(NAME)                                                 \
  *handled = true;                                                         \
  switch (g.nz) {                                                          \
    KW_GROUP_SIZES(NAME)                                                   \
    default: *handled = false; return hipSuccess;                          \
  }

}  // namespace kw
