// Register-resident mixed-radix Stockham FFT for gfx950.
//
// A workgroup transforms L independent lines of length N.  Each line is owned by
// TPL = N / E threads; thread j0 of a line holds the E elements
//     n = j0 + m * TPL,   m = 0 .. E-1
// in registers (E = lcm of the radices in the plan).  With that ownership the
// inputs of every Stockham butterfly j = j0 + q*TPL (q < E/R) of a radix-R stage
// are already in the thread's registers (they sit at n = j + r*N/R), so a stage
// is: twiddle + R-point DFT in registers, then one LDS exchange to move the
// outputs to their Stockham positions before the next stage.  The last stage's
// outputs land exactly on the owned positions, so an S-stage transform needs
// S-1 exchanges and global loads/stores go straight from/to registers.
//
// Radices 8, 4, 2, 3, 5, 7.  Twiddles come from a table tw[m] = exp(-2 pi i m/N)
// built in FP64 on the host and stored as FP32, rearranged stage by stage
// (stage_tw_*) into the array the transform reads.
#pragma once
#include <hip/hip_runtime.h>

namespace kw {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (DIR=-1, forward) or +i (DIR=+1, inverse)
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// compile-time radix plan
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 12;
struct RadixPlan {
  int n;
  int ns;
  int r[kMaxStages];
};

// Powers of two use radix 8 (E = 8).  With an odd factor f the elements per
// thread are lcm(power-of-two radix, f), so the power-of-two part uses radix 4
// next to 3 (E = 12) and radix 2 next to 5 or 7 (E = 10, 14): a few more LDS
// exchanges, but register arrays short enough for several waves per SIMD (radix
// 8 gave E = 24 / 40 / 56 and single-wave kernels that spilled).
//
// Plan variant pv = 1 (lengths with a factor 5 or 7 only): radix 4 next to 5 or 7
// (E = 20, 28) - fewer stages, and so fewer LDS exchanges, for twice the
// registers; the column passes time both variants per plan and keep the faster.
// Plan variant pv = 2 (powers of two only): radix 4 then 2 (E = 4) - twice the
// exchanges of the radix-8 plan for half the registers per thread.
constexpr RadixPlan make_plan(int n, int pv = 0) {
  RadixPlan p{n, 0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int odd = n;
  while (odd > 1 && odd % 2 == 0) odd /= 2;
  const bool f57 = odd % 5 == 0 || odd % 7 == 0;
  const int p2max = odd == 1 ? (pv == 2 ? 4 : 8) : f57 ? (pv == 1 ? 4 : 2) : 4;
  int m = n;
  while (p2max >= 8 && m % 8 == 0 && p.ns < kMaxStages) { p.r[p.ns++] = 8; m /= 8; }
  while (p2max >= 4 && m % 4 == 0 && p.ns < kMaxStages) { p.r[p.ns++] = 4; m /= 4; }
  while (m % 2 == 0 && p.ns < kMaxStages) { p.r[p.ns++] = 2; m /= 2; }
  while (m % 3 == 0) { p.r[p.ns++] = 3; m /= 3; }
  while (m % 5 == 0) { p.r[p.ns++] = 5; m /= 5; }
  while (m % 7 == 0) { p.r[p.ns++] = 7; m /= 7; }
  if (m != 1) p.ns = -1;
  return p;
}
// lengths whose plan variant 1 differs from variant 0
constexpr bool has_plan_variant(int n) {
  int odd = n;
  while (odd > 1 && odd % 2 == 0) odd /= 2;
  return (odd % 5 == 0 || odd % 7 == 0) && n % 4 == 0;
}
constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
constexpr int plan_stage_radix(int n, int s, int pv = 0) { return make_plan(n, pv).r[s]; }
constexpr int plan_nstages(int n, int pv = 0) { return make_plan(n, pv).ns; }
constexpr int plan_ns_before(int n, int s, int pv = 0) {
  int v = 1;
  for (int i = 0; i < s; ++i) v *= make_plan(n, pv).r[i];
  return v;
}
constexpr int plan_elems(int n, int pv = 0) {  // lcm of the radices
  int e = 1;
  for (int i = 0; i < make_plan(n, pv).ns; ++i) {
    const int r = make_plan(n, pv).r[i];
    e = e / gcd_c(e, r) * r;
  }
  return e;
}

// Stage-ordered twiddle table.  Stage s >= 1 (NS = product of the radices before
// it, R its radix) multiplies input r of butterfly class k = j mod NS by
// w^(r k N / (NS R)), w = exp(-2 pi i / N); its R - 1 twiddles of class k sit next
// to each other at stage_tw_base(s) + k (R - 1) + (r - 1), so a thread addresses a
// butterfly's twiddles as one base plus immediate offsets.  The stages take
// N - R0 entries in all (R0 = the first radix).
template <int N, int PV = 0>
constexpr int stage_tw_base(int s) { return plan_ns_before(N, s, PV) - plan_stage_radix(N, 0, PV); }
template <int N, int PV = 0>
constexpr int stage_tw_size() { return plan_nstages(N, PV) > 1 ? N - plan_stage_radix(N, 0, PV) : 1; }
// power of w held by entry e of the stage-ordered table
template <int N, int PV = 0>
__host__ __device__ constexpr int stage_tw_power(int e) {
  for (int s = 1; s < plan_nstages(N, PV); ++s) {
    const int R = plan_stage_radix(N, s, PV), NS = plan_ns_before(N, s, PV), b = stage_tw_base<N, PV>(s);
    if (e < b + NS * (R - 1)) {
      const int i = e - b;
      return (i % (R - 1) + 1) * (i / (R - 1)) * (N / (NS * R));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// small DFTs (in registers)
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* v) {
  float2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
  float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
  float2 s13 = cadd(a1, a3), d13 = mul_dir_i<DIR>(csub(a1, a3));
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  constexpr float h = 0.70710678118654752440f;
  float2 a[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a[j] = cadd(v[j], v[j + 4]);
    a[j + 4] = csub(v[j], v[j + 4]);
  }
  {
    float2 t = a[5];
    a[5] = DIR < 0 ? make_float2(h * (t.x + t.y), h * (t.y - t.x)) : make_float2(h * (t.x - t.y), h * (t.y + t.x));
    a[6] = mul_dir_i<DIR>(a[6]);
    t = a[7];
    a[7] = DIR < 0 ? make_float2(h * (-t.x + t.y), h * (-t.y - t.x)) : make_float2(h * (-t.x - t.y), h * (-t.y + t.x));
  }
  float2 e[4] = {a[0], a[1], a[2], a[3]};
  float2 o[4] = {a[4], a[5], a[6], a[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
  v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
  constexpr float c = -0.5f;
  constexpr float s = 0.86602540378443864676f;
  float2 a0 = v[0], a1 = v[1], a2 = v[2];
  float2 t1 = cadd(a1, a2);
  float2 t2 = make_float2(a0.x + c * t1.x, a0.y + c * t1.y);
  float2 d = csub(a1, a2);
  float2 isd = DIR < 0 ? make_float2(s * d.y, -s * d.x) : make_float2(-s * d.y, s * d.x);
  v[0] = cadd(a0, t1);
  v[1] = cadd(t2, isd);
  v[2] = csub(t2, isd);
}

template <int R, int DIR>
__device__ __forceinline__ void dft_odd(float2* v) {
  constexpr float c5[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f, -0.80901699437494742f, 0.30901699437494742f};
  constexpr float s5[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f, -0.58778525229247313f, -0.95105651629515357f};
  constexpr float c7[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f,
                           -0.90096886790241913f, -0.22252093395631440f, 0.62348980185873353f};
  constexpr float s7[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f,
                           -0.43388373911755812f, -0.97492791218182361f, -0.78183148246802981f};
  float2 in[R];
#pragma unroll
  for (int j = 0; j < R; ++j) in[j] = v[j];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float2 acc = in[0];
#pragma unroll
    for (int j = 1; j < R; ++j) {
      const int m = (j * k) % R;
      const float cc = R == 5 ? c5[m] : c7[m];
      const float ss = (R == 5 ? s5[m] : s7[m]) * (DIR < 0 ? -1.0f : 1.0f);
      acc.x += in[j].x * cc - in[j].y * ss;
      acc.y += in[j].x * ss + in[j].y * cc;
    }
    v[k] = acc;
  }
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else dft_odd<R, DIR>(v);
}

// ---------------------------------------------------------------------------
// LDS index maps used for the exchanges.  Each is idx(n, c) for position n of
// line c, plus lin(k) with idx(n + k, c) = idx(n, c) + lin(k) whenever k is a
// multiple of A, or n is a multiple of A and k < A; the transform uses it to
// address a thread's accesses as one base register plus immediate offsets
// (A = 0: not used, every access computes its own address).
// ---------------------------------------------------------------------------
// LDS bank model behind these maps (it matches the SQ_LDS_BANK_CONFLICT counts
// measured for the column passes at 512^3): a b64 access of a wave is served 16
// lanes at a time over 32 four-byte banks, i.e. 16 float2 slots, so a 16-lane
// group is conflict-free when its 16 slot indices are distinct mod 16.  In the
// row passes, base-plus-offset addressing let the compiler merge accesses into
// two-address forms and raised their conflict cycles by a quarter (97e6 to
// 122e6 per momentum dispatch at 512^3), so the row map keeps A = 0.
//
// lines-fastest (column passes: consecutive threads = consecutive columns).
// A 16-lane group touches 16 / L rows of L lines: consecutive rows (the reads,
// and the writes of stages after the first) or rows R apart (the first stage's
// writes, R = its radix).  Tiles of 16 lines or more keep a pitch of L + 1 (one
// row per group: conflict-free).  Narrower tiles store rows in groups of
// G = 32 / L with 4 slots after each group: consecutive rows stay contiguous and
// rows 8 apart are 4 slots apart mod 16, where a pitch of L + 1 or L + 2 put them
// on the same banks (at 512 with 8-wide tiles: 3.84 conflict cycles per LDS
// instruction with pitch 10, 0.29 grouped).
#ifndef KW_COLGRP
#define KW_COLGRP 1
#endif
#ifndef KW_COLPAD8
#define KW_COLPAD8 2
#endif
#ifndef KW_COLPAD4
#define KW_COLPAD4 2
#endif
template <int L>
struct ColIdx {
#if KW_COLGRP
  static constexpr int G = L >= 16 ? 0 : 32 / L;  // rows per group (a power of two)
  static constexpr int GS = G >= 32 ? 5 : G >= 16 ? 4 : G >= 8 ? 3 : G >= 4 ? 2 : G >= 2 ? 1 : 0;  // log2 G
  static_assert(G == 0 || (1 << GS) == G, "rows per group must be a power of two");
  static constexpr int PAD = G ? 4 : 0;
  static constexpr int P = G ? L : L + 1;  // row pitch inside a group
  static constexpr int A = G ? G : 1;
  __device__ __forceinline__ int operator()(int n, int c) const { return n * P + c + PAD * (n >> GS); }
  static constexpr int lin(int k) { return k * P + PAD * (k >> GS); }
  static constexpr int size(int N) { return N * P + (G ? PAD * ((N + G - 1) / G) : 0); }
#else
  static constexpr int P = L + (L == 8 ? KW_COLPAD8 : L == 4 ? KW_COLPAD4 : 1);  // row pitch
  static constexpr int A = 1;
  __device__ __forceinline__ int operator()(int n, int c) const { return n * P + c; }
  static constexpr int lin(int k) { return k * P; }
  static constexpr int size(int N) { return N * P; }
#endif
};
// line-major (row passes: consecutive threads = consecutive positions of one
// line; a 16-lane group is 16 consecutive positions j0 + m TPL, or in the first
// stage's writes positions 8 apart, or in the later stages' writes two runs of 8
// that are 64 apart), with a pad slot every 8 positions: about 2 conflict cycles
// per LDS instruction at 512, mostly the stride-8 writes and the k / N - k pair
// exchange of the real transforms.  KW_ROWSWZ=1 selects, for N a multiple of 16,
// slot n ^ h(n) with h taking bits 4-6 of n into bits 0-2 and bit 6 into bit 3,
// which separates all three transform patterns; it left the 512^3 step unchanged
// (the row passes are not LDS-bound) and made the batched 2-D sweep 8 % slower
// (its per-access address arithmetic), so it is off.
#ifndef KW_ROWSWZ
#define KW_ROWSWZ 0
#endif
template <int N>
struct RowIdx {
  static constexpr bool SW = KW_ROWSWZ && N % 16 == 0;
  static constexpr int LS = SW ? N + 8 : N + N / 8 + 1;
  static constexpr int A = 0;
  __device__ __forceinline__ int operator()(int n, int c) const {
    if constexpr (SW) return c * LS + (n ^ (((n >> 4) & 7) | (((n >> 6) & 1) << 3)));
    else return c * LS + n + (n >> 3);
  }
  static constexpr int lin(int k) { return k; }
  static constexpr int size(int L) { return L * LS; }
};

// EM multiplies the per-thread element count (more loads in flight per thread,
// fewer threads per line); N must stay divisible by E.  PV selects the plan variant.
template <int N, int EM = 1, int PV = 0>
struct FftShape {
  static constexpr int E = plan_elems(N, PV) * EM;
  static constexpr int TPL = N / E;
  static constexpr int S = plan_nstages(N, PV);
  static_assert(N % E == 0, "elements per thread must divide the line length");
};

// One stage on the owned elements v (see header); Idx maps (n, line) -> LDS slot.
template <int N, int EM, int S, int DIR, class Idx, int PV = 0>
__device__ __forceinline__ void reg_stage(float2 (&v)[(FftShape<N, EM, PV>::E)], float2* lds, const float2* __restrict__ tw, int c,
                                          int j0, const Idx& idx) {
  constexpr int E = FftShape<N, EM, PV>::E;
  constexpr int TPL = FftShape<N, EM, PV>::TPL;
  constexpr int R = plan_stage_radix(N, S, PV);
  constexpr int NS = plan_ns_before(N, S, PV);
  constexpr int Q = E / R;
  constexpr bool LAST = (S == plan_nstages(N, PV) - 1);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = j0 + q * TPL;
    float2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
    if constexpr (NS > 1) {
      constexpr int TB = stage_tw_base<N, PV>(S);
      const int k = j % NS;
      const float2* twk = tw + TB + k * (R - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = twk[r - 1];
        if (DIR > 0) w.y = -w.y;
        x[r] = cmul(x[r], w);
      }
    }
    dft<R, DIR>(x);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = x[r];
    } else {
      const int o = (j / NS) * NS * R + (j % NS);
      // o + r NS splits into idx(o) + lin(r NS) when the steps are multiples of
      // A, or when NS divides A and A divides NS R (then o mod A = j mod NS and
      // r NS mod A <= A - NS never carry past A)
      constexpr int A = Idx::A;
      if constexpr (A > 0 && (NS % A == 0 || ((NS * R) % A == 0 && A % NS == 0))) {
        const int b = idx(o, c);
#pragma unroll
        for (int r = 0; r < R; ++r) lds[b + Idx::lin(r * NS)] = x[r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) lds[idx(o + r * NS, c)] = x[r];
      }
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
    if constexpr (Idx::A > 0 && TPL % Idx::A == 0) {
      const int b = idx(j0, c);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = lds[b + Idx::lin(m * TPL)];
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = lds[idx(j0 + m * TPL, c)];
    }
    __syncthreads();
  }
}

template <int N, int EM, int S, int DIR, class Idx, int PV = 0>
__device__ __forceinline__ void reg_stages(float2 (&v)[(FftShape<N, EM, PV>::E)], float2* lds, const float2* __restrict__ tw, int c,
                                           int j0, const Idx& idx) {
  if constexpr (S < plan_nstages(N, PV)) {
    reg_stage<N, EM, S, DIR, Idx, PV>(v, lds, tw, c, j0, idx);
    reg_stages<N, EM, S + 1, DIR, Idx, PV>(v, lds, tw, c, j0, idx);
  }
}

// Transform the owned elements of line c in place (natural order in, natural
// order out at the same owned positions).  Every thread of the workgroup must
// call it (it contains barriers); the LDS buffer must not be in use.
template <int N, int DIR, class Idx, int EM = 1, int PV = 0>
__device__ __forceinline__ void fft_reg(float2 (&v)[(FftShape<N, EM, PV>::E)], float2* lds, const float2* __restrict__ tw, int c,
                                        int j0, const Idx& idx) {
  static_assert(plan_nstages(N, PV) > 0, "unsupported FFT length");
  reg_stages<N, EM, 0, DIR, Idx, PV>(v, lds, tw, c, j0, idx);
}

}  // namespace kw
