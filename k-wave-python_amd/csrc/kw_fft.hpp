// LDS-resident mixed-radix Stockham FFT building blocks for gfx950.
//
// A workgroup transforms L independent lines of length N that sit in LDS in a
// "lines-fastest" layout:  element n of line c lives at  buf[n * LP + c]  with
// LP = L + 1 (one float2 of padding per n-row keeps the transposing global<->LDS
// copies free of bank conflicts).  Each stage reads R values per butterfly,
// applies the Stockham twiddle, does an R-point DFT in registers and writes the
// results back in place after a barrier.  The output is in natural order.
//
// Radices: 8, 4, 2, 3, 5, 7 (constexpr plan per N).  Twiddles come from a table
// tw[m] = exp(-2 pi i m / N) built in FP64 on the host and stored as FP32.
#pragma once
#include <hip/hip_runtime.h>

namespace kw {

struct cf { float x, y; };

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
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

constexpr RadixPlan make_plan(int n) {
  RadixPlan p{n, 0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n;
  // powers of two: as many 8s as possible, then one 4 or 2
  while (m % 8 == 0 && p.ns < kMaxStages) { p.r[p.ns++] = 8; m /= 8; }
  if (m % 4 == 0) { p.r[p.ns++] = 4; m /= 4; }
  if (m % 2 == 0) { p.r[p.ns++] = 2; m /= 2; }
  while (m % 3 == 0) { p.r[p.ns++] = 3; m /= 3; }
  while (m % 5 == 0) { p.r[p.ns++] = 5; m /= 5; }
  while (m % 7 == 0) { p.r[p.ns++] = 7; m /= 7; }
  if (m != 1) p.ns = -1;  // unsupported prime factor
  return p;
}

constexpr int plan_stage_radix(int n, int s) { return make_plan(n).r[s]; }
constexpr int plan_nstages(int n) { return make_plan(n).ns; }
constexpr int plan_ns_before(int n, int s) {
  int v = 1;
  for (int i = 0; i < s; ++i) v *= make_plan(n).r[i];
  return v;
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
  // first level: pairs (j, j+4)
  float2 a[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a[j] = cadd(v[j], v[j + 4]);
    a[j + 4] = csub(v[j], v[j + 4]);
  }
  // twiddles W8^j on the difference half (j = 1, 2, 3)
  {
    float2 t = a[5];
    // W8^1 = (1 -/+ i)/sqrt2 for DIR -/+
    a[5] = DIR < 0 ? make_float2(h * (t.x + t.y), h * (t.y - t.x)) : make_float2(h * (t.x - t.y), h * (t.y + t.x));
    a[6] = mul_dir_i<DIR>(a[6]);
    t = a[7];
    // W8^3 = (-1 -/+ i)/sqrt2
    a[7] = DIR < 0 ? make_float2(h * (-t.x + t.y), h * (-t.y - t.x)) : make_float2(h * (-t.x - t.y), h * (-t.y + t.x));
  }
  float2 e[4] = {a[0], a[1], a[2], a[3]};
  float2 o[4] = {a[4], a[5], a[6], a[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // outputs: even indices from e, odd from o
  v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
  v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
  constexpr float c = -0.5f;
  constexpr float s = 0.86602540378443864676f;  // sin(2 pi/3)
  float2 a0 = v[0], a1 = v[1], a2 = v[2];
  float2 t1 = cadd(a1, a2);
  float2 t2 = make_float2(a0.x + c * t1.x, a0.y + c * t1.y);
  float2 d = csub(a1, a2);
  // forward: y1 = t2 - i s d ; y2 = t2 + i s d
  float2 isd = DIR < 0 ? make_float2(s * d.y, -s * d.x) : make_float2(-s * d.y, s * d.x);
  v[0] = cadd(a0, t1);
  v[1] = cadd(t2, isd);
  v[2] = csub(t2, isd);
}

// generic odd-radix DFT from constant cos/sin tables (radix 5 and 7)
template <int R, int DIR>
__device__ __forceinline__ void dft_odd(float2* v) {
  // cos(2 pi m / R), sin(2 pi m / R) for m = 0..R-1
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
// one Stockham stage over L lines of length N in LDS (in place)
// ---------------------------------------------------------------------------
template <int N, int L, int NT, int R, int NS, int DIR>
__device__ __forceinline__ void stockham_stage(float2* buf, const float2* __restrict__ tw, int tid) {
  constexpr int LP = L + 1;
  constexpr int NB = N / R;            // butterflies per line
  constexpr int TOT = NB * L;          // butterflies per workgroup
  constexpr int PER = (TOT + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);    // twiddle stride in the length-N table
  float2 v[PER][R];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int b = tid + i * NT;
    if (TOT % NT == 0 || b < TOT) {
      const int c = b % L;
      const int j = b / L;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * NB) * LP + c];
      if (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[(r * k * TWS) % N];
          if (DIR > 0) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      dft<R, DIR>(v[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int b = tid + i * NT;
    if (TOT % NT == 0 || b < TOT) {
      const int c = b % L;
      const int j = b / L;
      const int k = j % NS;
      const int o = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(o + r * NS) * LP + c] = v[i][r];
    }
  }
  __syncthreads();
}

template <int N, int L, int NT, int DIR, int S>
__device__ __forceinline__ void fft_stages(float2* buf, const float2* __restrict__ tw, int tid) {
  if constexpr (S < plan_nstages(N)) {
    constexpr int R = plan_stage_radix(N, S);
    constexpr int NS = plan_ns_before(N, S);
    stockham_stage<N, L, NT, R, NS, DIR>(buf, tw, tid);
    fft_stages<N, L, NT, DIR, S + 1>(buf, tw, tid);
  }
}

// Full in-LDS FFT of L lines (caller must __syncthreads() before, data in
// natural order after; ends with a barrier).  DIR = -1 forward, +1 inverse
// (unnormalised).
template <int N, int L, int NT, int DIR>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw, int tid) {
  static_assert(plan_nstages(N) > 0, "unsupported FFT length");
  fft_stages<N, L, NT, DIR, 0>(buf, tw, tid);
}

}  // namespace kw
