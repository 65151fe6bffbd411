// Generic-length family of the time-loop launchers.
//
// The fused kernels of kw_kernels.hpp are compiled for a fixed set of axis
// lengths (register-resident transforms with compile-time radix plans) and,
// along the contiguous axis, for even row counts.  Every other grid runs here:
// the same launcher entry points with the same argument blocks and results, so
// the step sequence, graphs, batching, sources and sensors of kw_abi.hip are
// unchanged.  Differences:
//
// * transforms are Stockham passes over lines held in LDS with a radix plan
//   chosen at run time: radices 8, 4, 2, 3, 5 and 7 as butterflies, any other
//   prime factor p as a direct DFT stage (each work item one output, O(p) per
//   output), so every length from 2 to kGenMaxN transforms;
// * the contiguous-axis work is unfused: inverse transforms into real
//   temporaries (Fields::tmp), one pointwise kernel for the update, forward
//   transforms of the results; the column passes with spectral jobs keep the
//   forward transform of the current input in a stash spectrum (Fields::gstash)
//   between its jobs.
//
// Arithmetic follows the fused kernels operation for operation (same operator
// tables, same order of the update terms), so the two families agree to FP32
// rounding of the transforms.
#include <algorithm>

#include "kw_kernels.hpp"

namespace kw {

constexpr int kGenMaxN = 4096;
constexpr int kGenThreads = 256;
constexpr int kGenMaxStages = 24;

struct GPlan {
  int n;
  int ns;
  int r[kGenMaxStages];
};

// Factor n: radix 8, 4 and 2 for the power of two, then 3, 5, 7, then every
// other prime factor (direct DFT stage).
static GPlan gplan(int n) {
  GPlan p{};
  p.n = n;
  int m = n;
  while (m % 8 == 0) { p.r[p.ns++] = 8; m /= 8; }
  while (m % 4 == 0) { p.r[p.ns++] = 4; m /= 4; }
  while (m % 2 == 0) { p.r[p.ns++] = 2; m /= 2; }
  for (int f : {3, 5, 7}) while (m % f == 0) { p.r[p.ns++] = f; m /= f; }
  for (int f = 11; m > 1; f += 2)
    while (m % f == 0) { p.r[p.ns++] = f; m /= f; }
  return p;
}

bool compiled_size(int n) { return size_supported_impl(n); }

bool size_supported(int n) { return compiled_size(n) || (n >= 2 && n <= kGenMaxN); }

// The fused contiguous-axis kernels need a compiled length and an even row count
// (they clamp a row pair into the grid).
bool generic_z(const Geom& g) { return !compiled_size(g.nz) || (g.rows & 1); }

// ---------------------------------------------------------------------------
// LDS transform of `nl` lines of length pl.n; element (n, line c) at
// buf[n * sn + c * sc].  Ping-pongs between A and B and returns the buffer that
// holds the result.  Every thread of the workgroup must call it.
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ float2 twid(const float2* __restrict__ tw, int m) {
  float2 w = ld2(gp(tw), m);
  if (DIR > 0) w.y = -w.y;
  return w;
}

// One Stockham stage of radix R (2, 3, 4, 5, 7 or 8) over every line: butterfly
// j of a line reads elements j + r N/R, twiddles them by tw[r k N/(NS R)]
// (k = j mod NS), and writes its R-point DFT to positions o + r NS,
// o = (j / NS) NS R + k -- the stage of the fused transforms (reg_stage).
template <int R, int DIR>
__device__ __forceinline__ void gstage(const float2* src, float2* dst, const float2* __restrict__ tw, int N, int NS, int nl, int sn,
                                       int sc) {
  const int nbf = N / R;
  const int TWS = N / (NS * R);
  for (int e = threadIdx.x; e < nbf * nl; e += blockDim.x) {
    const int c = e % nl, j = e / nl;
    const int k = j % NS;
    float2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = src[(j + r * nbf) * sn + c * sc];
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) x[r] = cmul(x[r], twid<DIR>(tw, r * k * TWS));
    }
    dft<R, DIR>(x);
    const int o = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[(o + r * NS) * sn + c * sc] = x[r];
  }
}

// A stage of any other prime radix R as a direct DFT: one work item per output q
// of one butterfly, O(R) operations each.
template <int DIR>
__device__ __forceinline__ void gstage_prime(const float2* src, float2* dst, const float2* __restrict__ tw, int N, int R, int NS, int nl,
                                             int sn, int sc) {
  const int nbf = N / R;
  const int TWS = N / (NS * R);
  const int root = N / R;  // tw[m * root] = exp(-2 pi i m / R)
  for (int e = threadIdx.x; e < N * nl; e += blockDim.x) {
    const int c = e % nl, jq = e / nl;
    const int j = jq % nbf, q = jq / nbf;
    const int k = j % NS;
    float2 acc = make_float2(0.f, 0.f);
    int m = 0;  // (r q) mod R
    for (int r = 0; r < R; ++r) {
      float2 x = src[(j + r * nbf) * sn + c * sc];
      if (NS > 1 && r > 0) x = cmul(x, twid<DIR>(tw, r * k * TWS));
      const float2 w = twid<DIR>(tw, m * root);
      acc.x += x.x * w.x - x.y * w.y;
      acc.y += x.x * w.y + x.y * w.x;
      m += q;
      if (m >= R) m -= R;
    }
    const int o = (j / NS) * NS * R + k;
    dst[(o + q * NS) * sn + c * sc] = acc;
  }
}

template <int DIR>
__device__ float2* gfft(float2* A, float2* B, const GPlan& pl, const float2* __restrict__ tw, int nl, int sn, int sc) {
  const int N = pl.n;
  float2* src = A;
  float2* dst = B;
  int NS = 1;
  for (int s = 0; s < pl.ns; ++s) {
    const int R = pl.r[s];
    switch (R) {
      case 2: gstage<2, DIR>(src, dst, tw, N, NS, nl, sn, sc); break;
      case 3: gstage<3, DIR>(src, dst, tw, N, NS, nl, sn, sc); break;
      case 4: gstage<4, DIR>(src, dst, tw, N, NS, nl, sn, sc); break;
      case 5: gstage<5, DIR>(src, dst, tw, N, NS, nl, sn, sc); break;
      case 7: gstage<7, DIR>(src, dst, tw, N, NS, nl, sn, sc); break;
      case 8: gstage<8, DIR>(src, dst, tw, N, NS, nl, sn, sc); break;
      default: gstage_prime<DIR>(src, dst, tw, N, R, NS, nl, sn, sc); break;
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    NS *= R;
  }
  return src;
}

// ---------------------------------------------------------------------------
// contiguous axis: real rows <-> half spectra, two rows per complex line
// ---------------------------------------------------------------------------
// Instance filter of a transform launch: skip instances that have finished
// (tdev), and, for source spectra, instances whose source `which` (-1: p,
// 0..2: u_i) is not an active spectral (CHAIN) source at this step.
enum { GF_ALL = -3, GF_ACTIVE = -2 };
__device__ __forceinline__ bool gen_skip(const int* tdev, const Inst* inst, int b, int which) {
  if (which == GF_ALL || !tdev) return false;
  const int t = gp(tdev)[b];
  const Inst& I = inst[b];
  if (t >= I.nt) return true;
  if (which == GF_ACTIVE) return false;
  const Src& s = which < 0 ? I.srcp : I.srcu[which];
  return !(s.mode == SRC_CHAIN && t < s.len);
}

__global__ __launch_bounds__(kGenThreads) void k_gzfwd(Geom g, const float* __restrict__ f, int64_t fs, float2* __restrict__ spec,
                                                       int64_t ss, const float2* __restrict__ tw, GPlan pl, int P, const int* tdev,
                                                       const Inst* __restrict__ inst, int which) {
  extern __shared__ float2 gsm[];
  const int b = blockIdx.y;
  if (gen_skip(tdev, inst, b, which)) return;
  const int N = pl.n, nh = g.nh;
  float2* A = gsm;
  float2* B = gsm + P * N;
  const auto fb = gp(f + b * fs);
  const auto sp = gw(spec + b * ss);
  const int pair0 = blockIdx.x * P;
  for (int e = threadIdx.x; e < P * N; e += blockDim.x) {
    const int lp = e / N, n = e % N;
    const int ra = 2 * (pair0 + lp), rb = ra + 1;
    A[e] = make_float2(ra < g.rows ? fb[(int64_t)ra * N + n] : 0.f, rb < g.rows ? fb[(int64_t)rb * N + n] : 0.f);
  }
  __syncthreads();
  const float2* Z = gfft<-1>(A, B, pl, tw, P, 1, N);
  for (int e = threadIdx.x; e < P * nh; e += blockDim.x) {
    const int lp = e / nh, k = e % nh;
    const int ra = 2 * (pair0 + lp), rb = ra + 1;
    const float2 z = Z[lp * N + k], zc = Z[lp * N + (N - k) % N];
    if (ra < g.rows) st2(sp, (int64_t)ra * g.nhp + k, make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y)));
    if (rb < g.rows) st2(sp, (int64_t)rb * g.nhp + k, make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x)));
  }
}

__global__ __launch_bounds__(kGenThreads) void k_gzinv(Geom g, const float2* __restrict__ spec, int64_t ss, float* __restrict__ f,
                                                       int64_t fs, const float2* __restrict__ tw, GPlan pl, int P, float scale,
                                                       const int* tdev, const Inst* __restrict__ inst, int which) {
  extern __shared__ float2 gsm[];
  const int b = blockIdx.y;
  if (gen_skip(tdev, inst, b, which)) return;
  const int N = pl.n, nh = g.nh;
  float2* A = gsm;
  float2* B = gsm + P * N;
  const auto sp = gp(spec + b * ss);
  const auto fb = gw(f + b * fs);
  const int pair0 = blockIdx.x * P;
  // Hermitian extension of both half spectra into the spectrum of a + i b
  for (int e = threadIdx.x; e < P * nh; e += blockDim.x) {
    const int lp = e / nh, k = e % nh;
    const int ra = 2 * (pair0 + lp), rb = ra + 1;
    const float2 a = ra < g.rows ? ld2(sp, (int64_t)ra * g.nhp + k) : make_float2(0.f, 0.f);
    const float2 c = rb < g.rows ? ld2(sp, (int64_t)rb * g.nhp + k) : make_float2(0.f, 0.f);
    float2* L = A + lp * N;
    if (k == 0 || (N % 2 == 0 && k == N / 2)) {
      L[k] = make_float2(a.x, c.x);
    } else {
      L[k] = make_float2(a.x - c.y, a.y + c.x);
      L[N - k] = make_float2(a.x + c.y, c.x - a.y);
    }
  }
  __syncthreads();
  const float2* Z = gfft<+1>(A, B, pl, tw, P, 1, N);
  for (int e = threadIdx.x; e < P * N; e += blockDim.x) {
    const int lp = e / N, n = e % N;
    const int ra = 2 * (pair0 + lp), rb = ra + 1;
    const float2 z = Z[e];
    if (ra < g.rows) fb[(int64_t)ra * N + n] = scale * z.x;
    if (rb < g.rows) fb[(int64_t)rb * N + n] = scale * z.y;
  }
}

// row pairs per workgroup: up to 16, with two LDS buffers of P N complex values <= 32 KiB
// (64 KiB at the longest lines)
static int gen_pairs(int N) {
  int P = 1;
  while (P < 16 && 2 * P * N <= 2048) P *= 2;
  return P;
}

static hipError_t gzfwd(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw,
                        const int* tdev, const Inst* inst, int which, hipStream_t st) {
  const GPlan pl = gplan(g.nz);
  const int P = gen_pairs(g.nz);
  const int npairs = (g.rows + 1) / 2;
  hipLaunchKernelGGL(k_gzfwd, dim3((npairs + P - 1) / P, nb), dim3(kGenThreads), sizeof(float2) * 2 * P * g.nz, st, g, f, fs, spec,
                     ss, tw, pl, P, tdev, inst, which);
  return hipGetLastError();
}

static hipError_t gzinv(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale,
                        const int* tdev, const Inst* inst, int which, hipStream_t st) {
  const GPlan pl = gplan(g.nz);
  const int P = gen_pairs(g.nz);
  const int npairs = (g.rows + 1) / 2;
  hipLaunchKernelGGL(k_gzinv, dim3((npairs + P - 1) / P, nb), dim3(kGenThreads), sizeof(float2) * 2 * P * g.nz, st, g, spec, ss, f,
                     fs, tw, pl, P, scale, tdev, inst, which);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// strided axes: column passes (k_col semantics)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kGenThreads) void k_gcol(Geom g, Fields fl, Ops ops, const Inst* __restrict__ inst, ColArgs a, GPlan pl,
                                                      int L) {
  extern __shared__ float2 gsm[];
  const int b = blockIdx.y;
  if (fl.tdev && gp(fl.tdev)[b] >= inst[b].nt) return;
  const int N = pl.n;
  float2* A = gsm;
  float2* B = gsm + N * L;
  const float2* tw = ops.tw[a.axis];
  const int ntile = (a.ncols + L - 1) / L;
  const int o = blockIdx.x / ntile, tile = blockIdx.x % ntile;
  const int64_t base0 = (int64_t)b * g.nspec + (int64_t)o * a.ostride + (int64_t)tile * L;
  const int nvalid = min(L, a.ncols - tile * L);
  auto load = [&](const float2* buf, float2* dstl) {
    const auto p = gp(buf);
    for (int e = threadIdx.x; e < N * L; e += blockDim.x) {
      const int c = e % L, n = e / L;
      dstl[e] = c < nvalid ? ld2(p, base0 + c + (int64_t)n * a.estride) : make_float2(0.f, 0.f);
    }
  };
  auto store = [&](const float2* srcl, float2* buf) {
    const auto p = gw(buf);
    for (int e = threadIdx.x; e < N * L; e += blockDim.x) {
      const int c = e % L, n = e / L;
      if (c < nvalid) st2(p, base0 + c + (int64_t)n * a.estride, srcl[e]);
    }
  };
  if (a.mode != 2) {
    for (int q = 0; q < a.nbuf; ++q) {
      float2* buf = fl.spec[a.bufs[q]];
      load(buf, A);
      __syncthreads();
      const float2* R = a.mode == 0 ? gfft<-1>(A, B, pl, tw, L, L, 1) : gfft<+1>(A, B, pl, tw, L, L, 1);
      store(R, buf);
      __syncthreads();
    }
    return;
  }
  // forward, multiply, inverse along axis 0; the forward transform of the current
  // input waits in the stash spectrum while its jobs run
  const float kap = inst[b].kap;
  int cur_in = -1;
  for (int j = 0; j < a.njob; ++j) {
    const XJob jb = a.job[j];
    if (jb.in != cur_in) {
      cur_in = jb.in;
      load(fl.spec[jb.in], A);
      __syncthreads();
      const float2* R = gfft<-1>(A, B, pl, tw, L, L, 1);
      store(R, fl.gstash);
      __syncthreads();
    }
    // multiply the stashed spectrum by the job's operator (regenerated per element)
    const auto sp = gp(fl.gstash);
    for (int e = threadIdx.x; e < N * L; e += blockDim.x) {
      const int c = e % L, n = e / L;
      const int col = min(tile * L + c, a.ncols - 1);
      int iy, iz;
      if (g.nd == 2) { iy = col; iz = 0; }
      else { iy = o; iz = col; }
      const float kx = gp(ops.k[0])[n];
      float kyz2;
      {
        const float ky = gp(ops.k[1])[iy];
        kyz2 = ky * ky;
        if (g.nd == 3) { const float kz = gp(ops.k[2])[iz]; kyz2 += kz * kz; }
      }
      const float km2 = fmaf(kx, kx, kyz2);
      float2 op = make_float2(1.f, 0.f);
      if (jb.kind == XK_GRAD || jb.kind == XK_DIV || jb.kind == XK_UNSTAGGER) {
        const float2* t = jb.kind == XK_GRAD ? ops.grad[jb.axis] : jb.kind == XK_DIV ? ops.div[jb.axis] : ops.unst[jb.axis];
        const int ia = jb.axis == 0 ? n : ((jb.axis == 1 && g.nd == 3) ? o : col);
        op = ld2(gp(t), ia);
      }
      switch (jb.kind) {
        case XK_GRAD:
        case XK_DIV: op = cscale(op, kap > 0.f ? sinc_k2(kap * kap * km2) : 1.f); break;
        case XK_FRAC1: op = make_float2(kpow2(km2, ops.frac1), 0.f); break;
        case XK_FRAC2: op = make_float2(kpow2(km2, ops.frac2), 0.f); break;
        case XK_SRCKAPPA: op = make_float2(kap > 0.f ? cos_k2(kap * kap * km2) : 1.f, 0.f); break;
        default: break;
      }
      const float2 v = c < nvalid ? ld2(sp, base0 + c + (int64_t)n * a.estride) : make_float2(0.f, 0.f);
      A[e] = cscale(cmul(v, op), ops.inv_ntot);
    }
    __syncthreads();
    const float2* R = gfft<+1>(A, B, pl, tw, L, L, 1);
    store(R, fl.spec[jb.out]);
    __syncthreads();
  }
}

// columns per workgroup: up to 16, two LDS buffers of N L complex values <= 32 KiB
// (64 KiB at the longest lines)
static int gen_cols(int N) {
  int L = 1;
  while (L < 16 && 2 * L * N <= 2048) L *= 2;
  return L;
}

hipError_t launch_col_gen(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline,
                          hipStream_t st) {
  if (a.mode == 2 && !fl.gstash) return hipErrorInvalidValue;
  const GPlan pl = gplan(nline);
  const int L = gen_cols(nline);
  const int ntile = (a.ncols + L - 1) / L;
  hipLaunchKernelGGL(k_gcol, dim3(a.nouter * ntile, nb), dim3(kGenThreads), sizeof(float2) * 2 * L * nline, st, g, fl, ops, inst, a, pl,
                     L);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pointwise updates (the fused z-pass kernels' arithmetic, one element per thread)
// ---------------------------------------------------------------------------
// pml multiplier of axis ax at (row r, position n); 1 on the contiguous axis
// unless that is ax
__device__ __forceinline__ float gpml(const float* const* tab, const Geom& g, int ax, int r, int n) { return pml_at(tab, g, ax, r, n); }

#define KW_GEN_ELEMENT_LOOP                                                       \
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < g.ntot; \
       idx += (int64_t)gridDim.x * blockDim.x)

__global__ __launch_bounds__(kGenThreads) void k_gmomentum(Geom g, Fields fl, const Inst* __restrict__ inst, int ax,
                                                           const float* __restrict__ gtmp, const float* __restrict__ ctmp) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  if (t >= I.nt) return;
  const Src& src = I.srcu[ax];
  const int smode = src_mode_at(src, t);
  const float sig_t = smode == SRC_PATTERN ? gp(src.sig)[t] : 0.f;
  const int64_t off = (int64_t)b * g.ntot;
  const auto u = gw(fl.u[ax] + off);
  KW_GEN_ELEMENT_LOOP {
    const int r = (int)(idx / g.nz), n = (int)(idx % g.nz);
    const float s = gpml(I.pmlsg, g, ax, r, n);
    float uv = s * (s * u[idx] - ld(I.dtr0[ax], idx, I.dtr0_s) * gp(gtmp)[off + idx]);
    uv = src_inline(src, smode, uv, idx, src.fac[ax], sig_t, smode == SRC_CHAIN ? gp(ctmp)[off + idx] : 0.f);
    if (src_is_point(smode)) {
      float v;
      if (csr_lookup(src, r, n, t, v)) uv = src_point(smode, uv, src.fac[ax], v);
    }
    u[idx] = uv;
  }
}

// mass conservation; with power law the outputs for the fractional Laplacians
// are rho0 div u (into aout) and the density sum (rhot), otherwise p.
__global__ __launch_bounds__(kGenThreads) void k_gmass(Geom g, Fields fl, const Inst* __restrict__ inst, int flags,
                                                       const float* d0, const float* d1, const float* d2,
                                                       const float* __restrict__ ctmp, float* aout) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  if (t >= I.nt) return;
  const int ND = g.nd;
  const Src& src = I.srcp;
  const int smode = src_mode_at(src, t);
  const float sig_t = smode == SRC_PATTERN ? gp(src.sig)[t] : 0.f;
  const bool powerlaw = flags & (FLAG_POWERLAW_ABS | FLAG_POWERLAW_DISP);
  const int64_t off = (int64_t)b * g.ntot;
  const float* dd[3] = {d0, d1, d2};
  KW_GEN_ELEMENT_LOOP {
    const int r = (int)(idx / g.nz), n = (int)(idx % g.nz);
    float qa = ld(I.rho0, idx, I.rho0_s);
    if (flags & FLAG_NONLINEAR) qa += 2.f * gw(fl.rhot)[off + idx];
    const float q = I.dt * qa;
    float ch = 0.f;
    if (smode == SRC_CHAIN) ch = gp(ctmp)[off + idx];
    else if (smode == SRC_PATTERN) ch = sig_t * gp(src.pattern)[idx];
    float du = 0.f, rt = 0.f;
    for (int i = 0; i < ND; ++i) {
      const auto rho = gw(fl.rho[i] + off);
      const float dv = gp(dd[i])[off + idx];
      if (fl.adjD) gw(fl.adjD)[(int64_t)i * gridDim.y * g.ntot + off + idx] = dv;  // adjoint replay: Dop_i u_i
      const float s = gpml(I.pml, g, i, r, n);
      float rv = s * (s * rho[idx] - q * dv);
      if (smode == SRC_CHAIN || smode == SRC_PATTERN) rv = rv + src.fac[i] * ch;
      rho[idx] = rv;
      du += dv;
      rt += rv;
    }
    if (src_is_point(smode)) {
      float v;
      if (csr_lookup(src, r, n, t, v)) {
        float sum = 0.f;
        for (int i = 0; i < ND; ++i) {
          const auto rho = gw(fl.rho[i] + off);
          const float x = src_point(smode, rho[idx], src.fac[i], v);
          rho[idx] = x;
          sum += x;
        }
        rt = sum;
      }
    }
    if (fl.rhot) gw(fl.rhot)[off + idx] = rt;
    const float r0 = ld(I.rho0, idx, I.rho0_s);
    if (powerlaw) {
      gw(aout)[off + idx] = r0 * du;
    } else {
      float qq = rt;
      const float c2 = ld(I.c0sq, idx, I.c0sq_s);
      if (flags & FLAG_STOKES) {
        qq += tau_at(I, idx, c2) * r0 * du;
        if (fl.adjA) gw(fl.adjA)[off + idx] = r0 * du;  // adjoint replay: the Stokes operand
      }
      if (flags & FLAG_NONLINEAR) qq += ld(I.bona, idx, I.bona_s) * rt * rt / (2.f * r0);
      gw(fl.p)[off + idx] = c2 * qq;
    }
  }
}

__global__ __launch_bounds__(kGenThreads) void k_geos(Geom g, Fields fl, const Inst* __restrict__ inst, int flags,
                                                      const float* __restrict__ av, const float* __restrict__ bv) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  if (t >= I.nt) return;
  const int64_t off = (int64_t)b * g.ntot;
  KW_GEN_ELEMENT_LOOP {
    const float rt = gw(fl.rhot)[off + idx];
    const float c2 = ld(I.c0sq, idx, I.c0sq_s);
    float q = rt;
    if (flags & FLAG_POWERLAW_ABS) q += tau_at(I, idx, c2) * gp(av)[off + idx];
    if (flags & FLAG_POWERLAW_DISP) q -= eta_at(I, idx, c2) * gp(bv)[off + idx];
    // adjoint replay: keep the absorption and dispersion operands
    if (fl.adjA && (flags & FLAG_POWERLAW_ABS)) gw(fl.adjA)[off + idx] = gp(av)[off + idx];
    if (fl.adjB && (flags & FLAG_POWERLAW_DISP)) gw(fl.adjB)[off + idx] = gp(bv)[off + idx];
    if (flags & FLAG_NONLINEAR) q += ld(I.bona, idx, I.bona_s) * rt * rt / (2.f * ld(I.rho0, idx, I.rho0_s));
    gw(fl.p)[off + idx] = c2 * q;
  }
}

// t = 0 initial-value override: u_i = dt/(2 rho0_sg_i) grad_i p0, rho_j = p0 / (c0^2 D)
__global__ __launch_bounds__(kGenThreads) void k_ginit_u(Geom g, Fields fl, const Inst* __restrict__ inst, const float* __restrict__ g0,
                                                         const float* __restrict__ g1, const float* __restrict__ g2) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int64_t off = (int64_t)b * g.ntot;
  const float* gg[3] = {g0, g1, g2};
  KW_GEN_ELEMENT_LOOP {
    for (int ax = 0; ax < g.nd; ++ax) gw(fl.u[ax])[off + idx] = 0.5f * ld(I.dtr0[ax], idx, I.dtr0_s) * gp(gg[ax])[off + idx];
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

// dense field of a spectral (CHAIN) source at step t: scale * sig at the source
// points, 0 elsewhere (instances without an active CHAIN source are skipped)
__global__ __launch_bounds__(kGenThreads) void k_gsrc_scatter(Geom g, Fields fl, const Inst* __restrict__ inst, int which,
                                                              float* __restrict__ out) {
  const int b = blockIdx.y;
  if (gen_skip(fl.tdev, inst, b, which)) return;
  const Inst& I = inst[b];
  const int t = gp(fl.tdev)[b];
  const Src& s = which < 0 ? I.srcp : I.srcu[which];
  const int64_t off = (int64_t)b * g.ntot;
  KW_GEN_ELEMENT_LOOP {
    const int r = (int)(idx / g.nz), n = (int)(idx % g.nz);
    float v = 0.f;
    if (gp(s.row_ptr)[r] != gp(s.row_ptr)[r + 1]) {
      float x;
      if (csr_lookup(s, r, n, t, x)) v = x;
    }
    gw(out)[off + idx] = v;
  }
}

// ---------------------------------------------------------------------------
// reverse sweep: the arithmetic of kw_adjoint.hpp's kernels, one element per
// thread; their transforms run as the launchers below (every instance, no step
// filter: the reverse sweep runs all of them)
// ---------------------------------------------------------------------------
// power law: cbar += 2 p pbar / c + w ((y-1) tau A - y eta B) / c; tau w, eta w out
__global__ __launch_bounds__(kGenThreads) void k_gadj_abs1(Geom g, AdjFields af, const Inst* __restrict__ inst,
                                                           float* __restrict__ tout, float* __restrict__ eout) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int64_t off = (int64_t)b * g.ntot;
  KW_GEN_ELEMENT_LOOP {
    const float c2 = ld(I.c0sq, idx, I.c0sq_s), c = sqrtf(c2);
    const float pbar = gp(af.pb)[off + idx];
    const float w = c2 * pbar;
    const float tau = tau_at(I, idx, c2);
    const float eta = af.disp ? eta_at(I, idx, c2) : 0.f;
    float gc = 2.f * gp(af.pt)[off + idx] * pbar / c + w * af.ym1 * tau / c * gp(af.At)[off + idx];
    if (af.disp) gc -= w * af.y * eta / c * gp(af.Bt)[off + idx];
    gw(af.cb)[off + idx] += gc;
    gw(tout)[off + idx] = tau * w;
    if (af.disp) gw(eout)[off + idx] = eta * w;
  }
}

// equation of state, sources and mass update (k_adj_eos); l1, l2: L1[tau w], L2[eta w]
// on the grid (power law), w_i out for the gradient jobs
__global__ __launch_bounds__(kGenThreads) void k_gadj_eos(Geom g, AdjFields af, const Inst* __restrict__ inst, int t,
                                                          const float* __restrict__ l1, const float* __restrict__ l2, float* w0,
                                                          float* w1, float* w2) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int ND = g.nd;
  const int64_t off = (int64_t)b * g.ntot;
  const int64_t nf = (int64_t)gridDim.y * g.ntot;
  const Src& src = I.srcp;
  const bool srcon = src.mode != SRC_NONE && t < src.len;
  const bool dir = srcon && src.mode == SRC_DIRICHLET;
  float* const wo[3] = {w0, w1, w2};
  KW_GEN_ELEMENT_LOOP {
    const int r = (int)(idx / g.nz), n = (int)(idx % g.nz);
    const float c2 = ld(I.c0sq, idx, I.c0sq_s);
    const float pbar = gp(af.pb)[off + idx];
    const float wc = c2 * pbar;
    const float r0 = ld(I.rho0, idx, I.rho0_s);
    float gc = 0.f;
    if (af.absorb != 2) {  // the power law's speed terms are k_gadj_abs1's
      gc = 2.f * gp(af.pt)[off + idx] * pbar / sqrtf(c2);
      if (af.absorb == 1) gc += wc * tau_at(I, idx, c2) / sqrtf(c2) * gp(af.At)[off + idx];
    }
    float rs = 0.f, sb = 0.f, rn[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < ND; ++i) {
      const float s = gpml(I.pml, g, i, r, n);
      float rbt = gp(af.rb[i])[off + idx] + wc;
      float nl = 1.f;
      if (af.nonlin) {
        rbt += wc * ld(I.bona, idx, I.bona_s) * gp(af.St)[off + idx] / r0;
        nl = 1.f + 2.f * gp(af.Sp)[off + idx] / r0;
      }
      float dbar = 0.f;
      if (af.absorb == 1) dbar = tau_at(I, idx, c2) * wc;
      if (af.absorb == 2) {
        if (af.disp) rbt -= gp(l2)[off + idx];
        dbar = gp(l1)[off + idx];
      }
      if (dir) {
        float v;
        if (csr_lookup(src, r, n, t, v)) {
          gc += -2.f * src.fac[i] * v * rbt / sqrtf(c2);
          rbt = 0.f;
        }
      }
      rs += src.fac[i] * rbt;
      if (af.nonlin) sb += -2.f * I.dt * s * gp(af.Dt)[(int64_t)i * nf + off + idx] * rbt;
      gw(wo[i])[off + idx] = s * I.dt * r0 * nl * rbt - r0 * dbar;
      rn[i] = s * s * rbt;
    }
    for (int i = 0; i < ND; ++i) gw(af.rb[i])[off + idx] = rn[i] + sb;
    gw(af.cb)[off + idx] += gc;
    if (srcon) gw(af.R)[off + idx] = rs;
  }
}

// momentum (k_adj_mom) of axis ax: gin = Gop_ax w_ax on the grid, v out
__global__ __launch_bounds__(kGenThreads) void k_gadj_mom(Geom g, AdjFields af, const Inst* __restrict__ inst, int t, int ax,
                                                          const float* __restrict__ gin, float* __restrict__ vout) {
  const int b = blockIdx.y;
  const Inst& I = inst[b];
  const int64_t off = (int64_t)b * g.ntot;
  const Src& us = I.srcu[ax];
  const bool uon = us.mode != SRC_NONE && t < us.len;
  const bool udir = uon && us.mode == SRC_DIRICHLET;
  float* const Ub = uon && !udir && af.U[ax] ? af.U[ax] + off : nullptr;
  KW_GEN_ELEMENT_LOOP {
    const int r = (int)(idx / g.nz), n = (int)(idx % g.nz);
    const float s = gpml(I.pmlsg, g, ax, r, n);
    float u = gp(af.ub[ax])[off + idx] + gp(gin)[off + idx];
    if (udir) {
      float v;
      if (csr_lookup(us, r, n, t, v)) u = 0.f;
    }
    if (Ub) gw(Ub)[idx] = u;
    gw(vout)[off + idx] = s * ld(I.dtr0[ax], idx, I.dtr0_s) * u;
    gw(af.ub[ax])[off + idx] = s * s * u;
  }
}

// pbar = sum_i Dop_i v_i (k_adj_sum), the terms on the grid
__global__ __launch_bounds__(kGenThreads) void k_gadj_sum(Geom g, AdjFields af, const float* __restrict__ d0,
                                                          const float* __restrict__ d1, const float* __restrict__ d2) {
  const int b = blockIdx.y;
  const int64_t off = (int64_t)b * g.ntot;
  KW_GEN_ELEMENT_LOOP {
    float a = gp(d0)[off + idx];
    if (g.nd > 1) a += gp(d1)[off + idx];
    if (g.nd > 2) a += gp(d2)[off + idx];
    gw(af.pb)[off + idx] = a;
  }
}

static dim3 gen_elem_grid(const Geom& g, int nb) {
  const int64_t blocks = std::min<int64_t>((g.ntot + kGenThreads - 1) / kGenThreads, 4096);
  return dim3((unsigned)blocks, nb);
}

#define KW_GEN_CHECK(x)                    \
  do {                                     \
    const hipError_t e_ = (x);             \
    if (e_ != hipSuccess) return e_;       \
  } while (0)

// ---------------------------------------------------------------------------
// launchers (same contracts as the fused family in kw_launch.hpp)
// ---------------------------------------------------------------------------
hipError_t launch_zfwd_gen(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw, hipStream_t st) {
  return gzfwd(g, nb, f, fs, spec, ss, tw, nullptr, nullptr, GF_ALL, st);
}

hipError_t launch_zinv_gen(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale,
                           hipStream_t st) {
  return gzinv(g, nb, spec, ss, f, fs, tw, scale, nullptr, nullptr, GF_ALL, st);
}

hipError_t launch_momentum_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[3]) return hipErrorInvalidValue;
  for (int ax = 0; ax < g.nd; ++ax) {
    KW_GEN_CHECK(gzinv(g, nb, fl.spec[ax], g.nspec, fl.tmp[0], g.ntot, tw, 1.f, fl.tdev, inst, GF_ACTIVE, st));
    if (fl.spec[g.nd + 1 + ax])
      KW_GEN_CHECK(gzinv(g, nb, fl.spec[g.nd + 1 + ax], g.nspec, fl.tmp[3], g.ntot, tw, 1.f, fl.tdev, inst, ax, st));
    hipLaunchKernelGGL(k_gmomentum, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, fl, inst, ax, fl.tmp[0], fl.tmp[3]);
    KW_GEN_CHECK(hipGetLastError());
    KW_GEN_CHECK(gzfwd(g, nb, fl.u[ax], g.ntot, fl.spec[ax], g.nspec, tw, fl.tdev, inst, GF_ACTIVE, st));
  }
  return hipSuccess;
}

hipError_t launch_mass_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[3]) return hipErrorInvalidValue;
  for (int i = 0; i < g.nd; ++i) KW_GEN_CHECK(gzinv(g, nb, fl.spec[i], g.nspec, fl.tmp[i], g.ntot, tw, 1.f, fl.tdev, inst, GF_ACTIVE, st));
  if (fl.spec[g.nd]) KW_GEN_CHECK(gzinv(g, nb, fl.spec[g.nd], g.nspec, fl.tmp[3], g.ntot, tw, 1.f, fl.tdev, inst, -1, st));
  hipLaunchKernelGGL(k_gmass, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, fl, inst, flags, fl.tmp[0], fl.tmp[1], fl.tmp[2],
                     fl.tmp[3], fl.tmp[0]);
  KW_GEN_CHECK(hipGetLastError());
  if (flags & (FLAG_POWERLAW_ABS | FLAG_POWERLAW_DISP)) {
    if (flags & FLAG_POWERLAW_ABS) KW_GEN_CHECK(gzfwd(g, nb, fl.tmp[0], g.ntot, fl.spec[0], g.nspec, tw, fl.tdev, inst, GF_ACTIVE, st));
    if (flags & FLAG_POWERLAW_DISP) KW_GEN_CHECK(gzfwd(g, nb, fl.rhot, g.ntot, fl.spec[1], g.nspec, tw, fl.tdev, inst, GF_ACTIVE, st));
    return hipSuccess;
  }
  return gzfwd(g, nb, fl.p, g.ntot, fl.spec[0], g.nspec, tw, fl.tdev, inst, GF_ACTIVE, st);
}

hipError_t launch_eos_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[1]) return hipErrorInvalidValue;
  if (flags & FLAG_POWERLAW_ABS) KW_GEN_CHECK(gzinv(g, nb, fl.spec[0], g.nspec, fl.tmp[0], g.ntot, tw, 1.f, fl.tdev, inst, GF_ACTIVE, st));
  if (flags & FLAG_POWERLAW_DISP) KW_GEN_CHECK(gzinv(g, nb, fl.spec[1], g.nspec, fl.tmp[1], g.ntot, tw, 1.f, fl.tdev, inst, GF_ACTIVE, st));
  hipLaunchKernelGGL(k_geos, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, fl, inst, flags, fl.tmp[0], fl.tmp[1]);
  KW_GEN_CHECK(hipGetLastError());
  return gzfwd(g, nb, fl.p, g.ntot, fl.spec[0], g.nspec, tw, fl.tdev, inst, GF_ACTIVE, st);
}

hipError_t launch_init_u_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[2]) return hipErrorInvalidValue;
  for (int ax = 0; ax < g.nd; ++ax) KW_GEN_CHECK(gzinv(g, nb, fl.spec[ax], g.nspec, fl.tmp[ax], g.ntot, tw, 1.f, nullptr, inst, GF_ALL, st));
  hipLaunchKernelGGL(k_ginit_u, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, fl, inst, fl.tmp[0], fl.tmp[1], fl.tmp[2]);
  return hipGetLastError();
}

hipError_t launch_src_build_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int which, hipStream_t st) {
  if (!fl.tmp[3]) return hipErrorInvalidValue;
  const int buf = which < 0 ? g.nd : g.nd + 1 + which;
  hipLaunchKernelGGL(k_gsrc_scatter, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, fl, inst, which, fl.tmp[3]);
  KW_GEN_CHECK(hipGetLastError());
  return gzfwd(g, nb, fl.tmp[3], g.ntot, fl.spec[buf], g.nspec, tw, fl.tdev, inst, which, st);
}

hipError_t launch_adj_abs1_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw,
                               hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[1]) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gadj_abs1, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, af, inst, fl.tmp[0], fl.tmp[1]);
  KW_GEN_CHECK(hipGetLastError());
  KW_GEN_CHECK(gzfwd(g, nb, fl.tmp[0], g.ntot, fl.spec[0], g.nspec, tw, nullptr, nullptr, GF_ALL, st));
  if (af.disp) KW_GEN_CHECK(gzfwd(g, nb, fl.tmp[1], g.ntot, fl.spec[1], g.nspec, tw, nullptr, nullptr, GF_ALL, st));
  return hipSuccess;
}

hipError_t launch_adj_eos_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t,
                              hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[2]) return hipErrorInvalidValue;
  if (af.absorb == 2) {  // L1[tau w], L2[eta w] from spectra 0 and 1 (the fractional column passes' outputs)
    if (!af.gl[0] || !af.gl[1]) return hipErrorInvalidValue;
    KW_GEN_CHECK(gzinv(g, nb, fl.spec[0], g.nspec, af.gl[0], g.ntot, tw, 1.f, nullptr, inst, GF_ALL, st));
    if (af.disp) KW_GEN_CHECK(gzinv(g, nb, fl.spec[1], g.nspec, af.gl[1], g.ntot, tw, 1.f, nullptr, inst, GF_ALL, st));
  }
  hipLaunchKernelGGL(k_gadj_eos, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, af, inst, t, af.gl[0], af.gl[1], fl.tmp[0],
                     fl.tmp[1], fl.tmp[2]);
  KW_GEN_CHECK(hipGetLastError());
  for (int i = 0; i < g.nd; ++i)
    KW_GEN_CHECK(gzfwd(g, nb, fl.tmp[i], g.ntot, fl.spec[i], g.nspec, tw, nullptr, nullptr, GF_ALL, st));
  return hipSuccess;
}

hipError_t launch_adj_mom_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t,
                              hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[1]) return hipErrorInvalidValue;
  for (int ax = 0; ax < g.nd; ++ax) {
    KW_GEN_CHECK(gzinv(g, nb, fl.spec[ax], g.nspec, fl.tmp[0], g.ntot, tw, 1.f, nullptr, inst, GF_ALL, st));
    hipLaunchKernelGGL(k_gadj_mom, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, af, inst, t, ax, fl.tmp[0], fl.tmp[1]);
    KW_GEN_CHECK(hipGetLastError());
    KW_GEN_CHECK(gzfwd(g, nb, fl.tmp[1], g.ntot, fl.spec[ax], g.nspec, tw, nullptr, nullptr, GF_ALL, st));
  }
  return hipSuccess;
}

hipError_t launch_adj_sum_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw,
                              hipStream_t st) {
  if (!fl.tmp[0] || !fl.tmp[2]) return hipErrorInvalidValue;
  for (int i = 0; i < g.nd; ++i)
    KW_GEN_CHECK(gzinv(g, nb, fl.spec[i], g.nspec, fl.tmp[i], g.ntot, tw, 1.f, nullptr, inst, GF_ALL, st));
  hipLaunchKernelGGL(k_gadj_sum, gen_elem_grid(g, nb), dim3(kGenThreads), 0, st, g, af, fl.tmp[0], fl.tmp[1], fl.tmp[2]);
  return hipGetLastError();
}

}  // namespace kw
