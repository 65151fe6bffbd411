// Spectral-mode contraction of the Fourier neural operator surrogate (BF16 MFMA).
//
// An FNO layer multiplies the retained Fourier modes of its input channels by a
// complex weight matrix per mode: Y[m] = X[m] W[m] (batch x Cin) (Cin x Cout).
// With real stacking, X2 = [Re X, Im X] and W2 = [[Re W, Im W], [-Im W, Re W]],
// forward, input gradient and weight gradient are all real mode-batched GEMMs
//   C[m] = A[m] Bt[m]^T      A: [M][R][K], Bt: [M][N][K] (bf16), C: [M][R][N] (fp32)
// (forward A = X2, Bt = W2^T; input gradient A = dY2, Bt = W2; weight gradient
// A = X2^T, Bt = dY2^T).  One wavefront owns a 32 x 32 tile of C and consumes K
// eight at a time with v_mfma_f32_32x32x8_bf16 (FP32 accumulation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "kw_launch.hpp"

namespace kw {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Operand layout of v_mfma_f32_32x32x8_bf16 for lane l of the wavefront: the A
// operand holds row l % 32, k = 4 (l / 32) + 0..3; the B operand column l % 32,
// same k; result register r holds C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
__global__ __launch_bounds__(64) void k_mode_gemm(int R, int N, int K, const uint16_t* __restrict__ A,
                                                  const uint16_t* __restrict__ Bt, float* __restrict__ C) {
  const int m = blockIdx.z;
  const int r0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int l = threadIdx.x;
  const int i = l & 31, kh = (l >> 5) * 4;
  const uint16_t* a = A + ((int64_t)m * R + r0 + i) * K + kh;
  const uint16_t* b = Bt + ((int64_t)m * N + n0 + i) * K + kh;
  f32x16 acc = {};
  for (int k = 0; k < K; k += 8) {
    const bf16x4 av = *reinterpret_cast<const bf16x4*>(a + k);
    const bf16x4 bv = *reinterpret_cast<const bf16x4*>(b + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  float* c = C + ((int64_t)m * R + r0) * N + n0 + i;
#pragma unroll
  for (int r = 0; r < 16; ++r) c[(int64_t)(8 * (r >> 2) + 4 * (l >> 5) + (r & 3)) * N] = acc[r];
}

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// The same contraction with the mode gather and scatter fused in: the A operand
// is read straight from a complex spectrum X[b][c][kx][ky] at the retained modes (kx rows 0..md-1 and
// NX-md..NX-1, ky < md; mode m = kx_index * md + ky), rounded to bf16 in
// registers, and the result is written straight into the complex spectrum Y at
// the same modes.  Strides (in complex elements) are free on both sides.
// Real and imaginary parts are interleaved per channel
// (k = 2 channel + part), so K = N = 2 C.  Batch rows >= B read zero and store
// nothing, so the batch needs no padding.
__global__ __launch_bounds__(64) void k_mode_contract(int B, int K, int md, int NX, const float2* __restrict__ X, int64_t xsb,
                                                      int64_t xsc, int64_t xsx, int64_t xsy, const uint16_t* __restrict__ Bt,
                                                      float* __restrict__ Y, int64_t ysb, int64_t ysc, int64_t ysx, int64_t ysy) {
  const int m = blockIdx.z;
  const int r0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int l = threadIdx.x;
  const int i = l & 31, kh = (l >> 5) * 4;
  const int kxi = m / md, ky = m % md;
  const int row = kxi < md ? kxi : NX - 2 * md + kxi;
  const int r = r0 + i;
  const bool rv = r < B;
  const float2* xa = X + (int64_t)(rv ? r : 0) * xsb + (int64_t)row * xsx + (int64_t)ky * xsy;
  const uint16_t* b = Bt + ((int64_t)m * K + n0 + i) * K + kh;
  f32x16 acc = {};
  for (int k = 0; k < K; k += 8) {
    const int ch = (k + kh) >> 1;  // channels ch, ch + 1
    bf16x4 av = {0, 0, 0, 0};
    if (rv) {
      const float2 x0 = xa[(int64_t)ch * xsc], x1 = xa[(int64_t)(ch + 1) * xsc];
      av = bf16x4{bf16_rne(x0.x), bf16_rne(x0.y), bf16_rne(x1.x), bf16_rne(x1.y)};
    }
    const bf16x4 bv = *reinterpret_cast<const bf16x4*>(b + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  const int col = n0 + i, o = col >> 1, part = col & 1;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int rr = r0 + 8 * (q >> 2) + 4 * (l >> 5) + (q & 3);
    if (rr < B) Y[2 * ((int64_t)rr * ysb + (int64_t)o * ysc + (int64_t)row * ysx + (int64_t)ky * ysy) + part] = acc[q];
  }
}

// Tiled form for 16 retained ky modes stored next to each other (ky stride 1 on
// both sides: one (sample, channel, kx) run of modes is 128 contiguous bytes).
// A workgroup takes one kx row, its 16 modes and 32 batch rows with all channels:
//   1. the 32 x C x 16 complex inputs are read run by run (16 lanes per 128-byte
//      run), rounded to bf16 and stored mode-major in LDS (As[ky][b][k]);
//   2. wavefront w multiplies modes 4w .. 4w+3, the A operand from LDS and the
//      weights Bt[m][n][k] of each mode from global memory (L2-resident: 8 KB per
//      mode), into FP32 accumulators;
//   3. the results go back through LDS (32 output columns at a time, reusing the
//      input space) and out run by run again.
// The per-mode kernel above reads and writes one strided 8-byte value per lane.
template <int K>
__global__ __launch_bounds__(256) void k_mode_contract16(int B, int NX, const float2* __restrict__ X, int64_t xsb, int64_t xsc,
                                                         int64_t xsx, const uint16_t* __restrict__ Bt, float2* __restrict__ Y,
                                                         int64_t ysb, int64_t ysc, int64_t ysx) {
  constexpr int MD = 16, RB = 32, C = K / 2;
  constexpr int RP = K + 4;        // bf16 pitch of a batch row in As (8 pad bytes: the MFMA reads spread over the banks)
  constexpr int KS = RB * RP + 8;  // bf16 pitch of a mode block in As
  constexpr int NPASS = K / 32;    // 32-column passes of the output
  constexpr int YP = MD + 1;       // float pitch of one (row, column) run of modes in Ys
  constexpr int A_BYTES = MD * KS * 2, Y_BYTES = RB * 32 * YP * 4;
  static_assert(K % 32 == 0 && K <= 64, "2 C must be 32 or 64");
  __shared__ __align__(16) unsigned char smem[A_BYTES > Y_BYTES ? A_BYTES : Y_BYTES];
  uint16_t* As = reinterpret_cast<uint16_t*>(smem);
  float* Ys = reinterpret_cast<float*>(smem);
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int kxi = blockIdx.y;
  const int row = kxi < MD ? kxi : NX - 2 * MD + kxi;
  const int b0 = blockIdx.x * RB;
  for (int e = tid; e < RB * C * MD; e += 256) {
    const int ky = e % MD, c = (e / MD) % C, bl = e / (MD * C);
    const int b = b0 + bl;
    float2 x = make_float2(0.f, 0.f);
    if (b < B) x = X[(int64_t)b * xsb + (int64_t)c * xsc + (int64_t)row * xsx + ky];
    const uint32_t packed = (uint32_t)(uint16_t)bf16_rne(x.x) | ((uint32_t)(uint16_t)bf16_rne(x.y) << 16);
    *reinterpret_cast<uint32_t*>(&As[ky * KS + bl * RP + 2 * c]) = packed;
  }
  __syncthreads();
  f32x16 acc[4][NPASS];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ky = 4 * w + q, m = kxi * MD + ky;
    const uint16_t* aa = As + ky * KS + (l & 31) * RP + (l >> 5) * 4;
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const uint16_t* bb = Bt + ((int64_t)m * K + p * 32 + (l & 31)) * K + (l >> 5) * 4;
      acc[q][p] = f32x16{};
#pragma unroll
      for (int k = 0; k < K; k += 8) {
        const bf16x4 av = *reinterpret_cast<const bf16x4*>(aa + k);
        const bf16x4 bv = *reinterpret_cast<const bf16x4*>(bb + k);
        acc[q][p] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc[q][p], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NPASS; ++p) {
    __syncthreads();  // As (first pass) or the previous pass's Ys is no longer read
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ky = 4 * w + q;
#pragma unroll
      for (int r = 0; r < 16; ++r) Ys[((8 * (r >> 2) + 4 * (l >> 5) + (r & 3)) * 32 + (l & 31)) * YP + ky] = acc[q][p][r];
    }
    __syncthreads();
    for (int e = tid; e < RB * 16 * MD; e += 256) {
      const int ky = e % MD, o = (e / MD) % 16, bl = e / (MD * 16);
      const int b = b0 + bl;
      if (b < B)
        Y[(int64_t)b * ysb + (int64_t)(p * 16 + o) * ysc + (int64_t)row * ysx + ky] =
            make_float2(Ys[(bl * 32 + 2 * o) * YP + ky], Ys[(bl * 32 + 2 * o + 1) * YP + ky]);
    }
  }
}

hipError_t launch_mode_contract(int B, int K, int md, int NX, const float2* X, const int64_t xs[4], const uint16_t* Bt, float2* Y,
                                const int64_t ys[4], hipStream_t st) {
  if (md == 16 && xs[3] == 1 && ys[3] == 1 && (K == 32 || K == 64)) {
    const dim3 grid((B + 31) / 32, 2 * md);
    if (K == 64)
      hipLaunchKernelGGL(k_mode_contract16<64>, grid, dim3(256), 0, st, B, NX, X, xs[0], xs[1], xs[2], Bt, Y, ys[0], ys[1], ys[2]);
    else
      hipLaunchKernelGGL(k_mode_contract16<32>, grid, dim3(256), 0, st, B, NX, X, xs[0], xs[1], xs[2], Bt, Y, ys[0], ys[1], ys[2]);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_mode_contract, dim3(K / 32, (B + 31) / 32, 2 * md * md), dim3(64), 0, st, B, K, md, NX, X, xs[0], xs[1],
                     xs[2], xs[3], Bt, reinterpret_cast<float*>(Y), ys[0], ys[1], ys[2], ys[3]);
  return hipGetLastError();
}

// Pointwise part of a layer, fused for inference:
//   out[b][o][p] = act(s[b][o][p] + pb[o] + sum_c pw[o][c] v[b][c][p])
// v, s, out bf16 [B][C][P]; pw [C][C], pb [C] fp32; act = exact GELU
// 0.5 x (1 + erf(x / sqrt 2)) or identity.  Each thread owns four consecutive
// positions of one sample (8-byte loads per channel), accumulates in FP32 and
// rounds the result to bf16 once.  The weights are uniform across the
// workgroup and are read through scalar loads.
template <int C>
__global__ __launch_bounds__(256) void k_fno_pointwise(int64_t P4, const uint16_t* __restrict__ v, const uint16_t* __restrict__ s,
                                                       const float* __restrict__ pw, const float* __restrict__ pb,
                                                       uint16_t* __restrict__ out, int gelu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= P4) return;
  const int64_t P = 4 * P4;
  const int64_t base = (int64_t)b * C * P + 4 * t;
  float x[C][4];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const ushort4 q = *reinterpret_cast<const ushort4*>(v + base + (int64_t)c * P);
    x[c][0] = __uint_as_float((unsigned)q.x << 16);
    x[c][1] = __uint_as_float((unsigned)q.y << 16);
    x[c][2] = __uint_as_float((unsigned)q.z << 16);
    x[c][3] = __uint_as_float((unsigned)q.w << 16);
  }
#pragma unroll 4
  for (int o = 0; o < C; ++o) {
    const ushort4 q = *reinterpret_cast<const ushort4*>(s + base + (int64_t)o * P);
    float y[4] = {__uint_as_float((unsigned)q.x << 16), __uint_as_float((unsigned)q.y << 16),
                  __uint_as_float((unsigned)q.z << 16), __uint_as_float((unsigned)q.w << 16)};
    float w[4] = {pb[o], pb[o], pb[o], pb[o]};
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const float a = pw[o * C + c];
#pragma unroll
      for (int j = 0; j < 4; ++j) w[j] = fmaf(a, x[c][j], w[j]);
    }
    short r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float z = y[j] + w[j];
      if (gelu) z = 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
      r[j] = bf16_rne(z);
    }
    *reinterpret_cast<ushort4*>(out + base + (int64_t)o * P) =
        make_ushort4((unsigned short)r[0], (unsigned short)r[1], (unsigned short)r[2], (unsigned short)r[3]);
  }
}

hipError_t launch_fno_pointwise(int B, int C, int64_t P, const uint16_t* v, const uint16_t* s, const float* pw, const float* pb,
                                uint16_t* out, int gelu, hipStream_t st) {
  const int64_t P4 = P / 4;
  const dim3 grid((unsigned)((P4 + 255) / 256), (unsigned)B);
  if (C == 16) hipLaunchKernelGGL(k_fno_pointwise<16>, grid, dim3(256), 0, st, P4, v, s, pw, pb, out, gelu);
  else if (C == 32) hipLaunchKernelGGL(k_fno_pointwise<32>, grid, dim3(256), 0, st, P4, v, s, pw, pb, out, gelu);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// Truncated 2-D DFT of the surrogate's 64 x 64 channel slabs for inference, both
// axes in one kernel (one workgroup of four waves per slab, bf16 operands, FP32
// accumulation on v_mfma_f32_32x32x8_bf16).  With the real DFT matrices of
// 32 retained kx (rows interleaved re/im) and 16 retained ky:
//   forward  T = Mx V (64 x 64, rounded to bf16 in LDS), X = T' My with
//            T' = T viewed 32 x 128 (row j = [T_2j | T_2j+1]); X 32 x 32 fp32,
//            columns interleaved re/im per ky (four 16 x 16 tiles, one per wave);
//   inverse  C = Y Ny (32 x 128, rounded to bf16 in LDS), out = Nx C' with
//            C' = C viewed 64 x 64; out bf16.
// My and Ny are passed transposed (MyT 32 x 128, NyT 128 x 32) so every MFMA
// operand is four consecutive bf16 values.
namespace {
constexpr int DN = 64, DK = 32;  // slab edge, retained kx rows (= interleaved ky columns)
// LDS copies of B operands are stored transposed (column n contiguous in k) with
// this row pitch: four k values are one 8-byte read, and the pad spreads the
// transposing writes over the banks.
constexpr int LDT = DN + 4;
}  // namespace

// Both kernels run on a resident grid whose workgroups loop over slabs: the
// fixed DFT-matrix operands of every MFMA stay in registers for the whole loop,
// and the next slab's input is loaded while the current one is transformed.
__global__ __launch_bounds__(256) void k_dft_fwd(int64_t S, const uint16_t* __restrict__ v, const uint16_t* __restrict__ mx,
                                                 const uint16_t* __restrict__ myt, float* __restrict__ x) {
  __shared__ __attribute__((aligned(16))) uint16_t vt[DN * LDT];  // the slab, transposed: vt[y][x]
  __shared__ __attribute__((aligned(16))) uint16_t ts[DN * DN];   // T = Mx V
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int i = l & 31, kh = (l >> 5) * 4;
  const int mt = w >> 1, nt = w & 1;
  // T' row j = T rows 2j, 2j+1: element (j, q) sits at ts[j * 2 DN + q].  X = T' My as four
  // 16 x 16 output tiles, one per wave (v_mfma_f32_16x16x16_bf16: lane l holds row / column
  // l % 16 and the four k values 4 (l / 16) .. + 3 of each 16-wide k block)
  const int r16 = l & 15, kq = (l >> 4) * 4;
  bf16x4 am[DN / 8], bm[2 * DN / 16];
#pragma unroll
  for (int k = 0; k < DN / 8; ++k) am[k] = *reinterpret_cast<const bf16x4*>(mx + (mt * 32 + i) * DN + 8 * k + kh);
#pragma unroll
  for (int k = 0; k < 2 * DN / 16; ++k) bm[k] = *reinterpret_cast<const bf16x4*>(myt + (nt * 16 + r16) * 2 * DN + 16 * k + kq);
  constexpr int QL = DN * DN / 8 / 256;  // 16-byte loads per thread and slab
  uint4 q[QL];
  int64_t slab = blockIdx.x;
  if (slab < S) {
    const uint4* src = reinterpret_cast<const uint4*>(v + slab * DN * DN);
#pragma unroll
    for (int t = 0; t < QL; ++t) q[t] = src[tid + 256 * t];
  }
  for (; slab < S; slab += gridDim.x) {
    __syncthreads();  // the previous slab's reads of vt and ts are done
#pragma unroll
    for (int t = 0; t < QL; ++t) {
      const int e = tid + 256 * t;
      const int xr = e >> 3, y0 = (e & 7) * 8;
      const unsigned u[4] = {q[t].x, q[t].y, q[t].z, q[t].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        vt[(y0 + 2 * j) * LDT + xr] = (uint16_t)(u[j] & 0xffffu);
        vt[(y0 + 2 * j + 1) * LDT + xr] = (uint16_t)(u[j] >> 16);
      }
    }
    if (slab + gridDim.x < S) {
      const uint4* src = reinterpret_cast<const uint4*>(v + (slab + gridDim.x) * DN * DN);
#pragma unroll
      for (int t = 0; t < QL; ++t) q[t] = src[tid + 256 * t];
    }
    __syncthreads();
    {
      f32x16 acc = {};
#pragma unroll
      for (int k = 0; k < DN / 8; ++k) {
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(vt + (nt * 32 + i) * LDT + 8 * k + kh);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(am[k], b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) ts[(mt * 32 + 8 * (r >> 2) + kh + (r & 3)) * DN + nt * 32 + i] = (uint16_t)bf16_rne(acc[r]);
    }
    __syncthreads();
    {
      f32x4 acc = {};
#pragma unroll
      for (int k = 0; k < 2 * DN / 16; ++k) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(ts + (mt * 16 + r16) * 2 * DN + 16 * k + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bm[k], acc, 0, 0, 0);
      }
      float* xo = x + slab * DK * DK;
#pragma unroll
      for (int r = 0; r < 4; ++r) xo[(mt * 16 + kq + r) * DK + nt * 16 + r16] = acc[r];
    }
  }
}

__global__ __launch_bounds__(256) void k_dft_inv(int64_t S, const float* __restrict__ y, const uint16_t* __restrict__ nyt,
                                                 const uint16_t* __restrict__ nx, uint16_t* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint16_t ct[DN * LDT];  // C' = (Y Ny viewed 64 x 64), transposed: ct[n][k]
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int i = l & 31, kh = (l >> 5) * 4;
  const int mt = w >> 1, nt = w & 1;
  bf16x4 bn[DK / 8], an[DN / 8];
#pragma unroll
  for (int k = 0; k < DK / 8; ++k) bn[k] = *reinterpret_cast<const bf16x4*>(nyt + (w * 32 + i) * DK + 8 * k + kh);
#pragma unroll
  for (int k = 0; k < DN / 8; ++k) an[k] = *reinterpret_cast<const bf16x4*>(nx + (mt * 32 + i) * DN + 8 * k + kh);
  float4 yq[DK / 8];
  int64_t slab = blockIdx.x;
  if (slab < S) {
    const float* ys = y + slab * DK * DK;
#pragma unroll
    for (int k = 0; k < DK / 8; ++k) yq[k] = *reinterpret_cast<const float4*>(ys + i * DK + 8 * k + kh);
  }
  for (; slab < S; slab += gridDim.x) {
    // wave w: columns [32 w, 32 w + 32) of C, K = 32
    f32x16 acc = {};
#pragma unroll
    for (int k = 0; k < DK / 8; ++k) {
      const bf16x4 a = bf16x4{bf16_rne(yq[k].x), bf16_rne(yq[k].y), bf16_rne(yq[k].z), bf16_rne(yq[k].w)};
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bn[k], acc, 0, 0, 0);
    }
    if (slab + gridDim.x < S) {
      const float* ys = y + (slab + gridDim.x) * DK * DK;
#pragma unroll
      for (int k = 0; k < DK / 8; ++k) yq[k] = *reinterpret_cast<const float4*>(ys + i * DK + 8 * k + kh);
    }
    __syncthreads();  // the previous slab's reads of ct are done
    {
      // C[j][q] is C'[2 j + q / 64][q % 64], with q = 32 w + i
      const int n = (w & 1) * 32 + i, kk = w >> 1;
#pragma unroll
      for (int r = 0; r < 16; ++r) ct[n * LDT + 2 * (8 * (r >> 2) + kh + (r & 3)) + kk] = (uint16_t)bf16_rne(acc[r]);
    }
    __syncthreads();
    {
      f32x16 acc2 = {};
#pragma unroll
      for (int k = 0; k < DN / 8; ++k) {
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(ct + (nt * 32 + i) * LDT + 8 * k + kh);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(an[k], b, acc2, 0, 0, 0);
      }
      uint16_t* o = out + slab * DN * DN;
#pragma unroll
      for (int r = 0; r < 16; ++r) o[(mt * 32 + 8 * (r >> 2) + kh + (r & 3)) * DN + nt * 32 + i] = (uint16_t)bf16_rne(acc2[r]);
    }
  }
}

// Input lift of the surrogate for inference, one pass from the fine map to the
// first layer's bf16 channels: with x the 2 x 2 average of the normalised fine
// map (cmap - cmid) / chalf,
//   v[b][c][i][j] = bf16(x[b][i][j] lc[c] + g[c][i][j])
// (g the lift of the coordinate grid, bf16).  A thread owns four consecutive j of
// one row i of one sample: two 32-byte reads of the fine map, then per channel
// one 8-byte read of g (cached) and one 8-byte write.
template <int C>
__global__ __launch_bounds__(256) void k_fno_lift(int64_t B, int N, const float* __restrict__ cmap, float cmid, float chalf,
                                                  const float* __restrict__ lc, const uint16_t* __restrict__ g,
                                                  uint16_t* __restrict__ v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = N / 4;
  if (t >= B * N * q) return;
  const int64_t b = t / ((int64_t)N * q);
  const int i = (int)((t / q) % N), j0 = (int)(t % q) * 4;
  const float* r0 = cmap + (b * 2 * N + 2 * i) * 2 * N + 2 * j0;
  const float* r1 = r0 + 2 * N;
  const float4 a0 = *reinterpret_cast<const float4*>(r0), a1 = *reinterpret_cast<const float4*>(r0 + 4);
  const float4 c0 = *reinterpret_cast<const float4*>(r1), c1 = *reinterpret_cast<const float4*>(r1 + 4);
  const float top[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  const float bot[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
  float x[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float s0 = (top[2 * k] - cmid) / chalf + (top[2 * k + 1] - cmid) / chalf;
    const float s1 = (bot[2 * k] - cmid) / chalf + (bot[2 * k + 1] - cmid) / chalf;
    x[k] = 0.25f * (s0 + s1);
  }
  const int64_t pos = (int64_t)i * N + j0;
#pragma unroll 8
  for (int c = 0; c < C; ++c) {
    const ushort4 gq = *reinterpret_cast<const ushort4*>(g + (int64_t)c * N * N + pos);
    const float l = lc[c];
    const unsigned short o0 = (unsigned short)bf16_rne(fmaf(x[0], l, __uint_as_float((unsigned)gq.x << 16)));
    const unsigned short o1 = (unsigned short)bf16_rne(fmaf(x[1], l, __uint_as_float((unsigned)gq.y << 16)));
    const unsigned short o2 = (unsigned short)bf16_rne(fmaf(x[2], l, __uint_as_float((unsigned)gq.z << 16)));
    const unsigned short o3 = (unsigned short)bf16_rne(fmaf(x[3], l, __uint_as_float((unsigned)gq.w << 16)));
    *reinterpret_cast<ushort4*>(v + (b * C + c) * N * N + pos) = make_ushort4(o0, o1, o2, o3);
  }
}

hipError_t launch_fno_lift(int64_t B, int C, int N, const float* cmap, float cmid, float chalf, const float* lc, const uint16_t* g,
                           uint16_t* v, hipStream_t st) {
  const int64_t threads = B * N * (N / 4);
  const dim3 grid((unsigned)((threads + 255) / 256));
  if (C == 32)
    hipLaunchKernelGGL(k_fno_lift<32>, grid, dim3(256), 0, st, B, N, cmap, cmid, chalf, lc, g, v);
  else
    hipLaunchKernelGGL(k_fno_lift<16>, grid, dim3(256), 0, st, B, N, cmap, cmid, chalf, lc, g, v);
  return hipGetLastError();
}

// Sensor-head gather of the surrogate: sel[b][k][c][x] = v[b][c][x][cols[k]] for
// k < K (bf16, v [B][C][64][64]).  One workgroup per (b, c) slab: the slab is
// staged in LDS with a padded row pitch (the column reads of different rows
// fall in different banks), then written out one contiguous row of 64 x per k.
__global__ __launch_bounds__(256) void k_head_gather(int C, int K, const uint16_t* __restrict__ v, const int* __restrict__ cols,
                                                     uint16_t* __restrict__ sel) {
  constexpr int LDP = DN + 2;
  __shared__ uint16_t vs[DN * LDP];
  const int64_t slab = blockIdx.x;  // b * C + c
  const int64_t b = slab / C, c = slab - b * C;
  const int tid = threadIdx.x;
  const uint16_t* src = v + slab * DN * DN;
  for (int e = tid; e < DN * DN; e += 256) vs[(e >> 6) * LDP + (e & 63)] = src[e];
  __syncthreads();
  for (int e = tid; e < K * DN; e += 256) {
    const int k = e >> 6, xr = e & 63;
    const int y = min(max(cols[k], 0), DN - 1);  // the caller checks the range; this keeps reads in the slab
    sel[((b * K + k) * C + c) * DN + xr] = vs[xr * LDP + y];
  }
}

hipError_t launch_head_gather(int64_t B, int C, int K, const uint16_t* v, const int* cols, uint16_t* sel, hipStream_t st) {
  hipLaunchKernelGGL(k_head_gather, dim3((unsigned)(B * C)), dim3(256), 0, st, C, K, v, cols, sel);
  return hipGetLastError();
}

// Resident workgroups of a 256-thread kernel on the current device (cached per
// device and kernel; 0 when the runtime cannot say).
static int64_t resident_groups(const void* kernel, int slot) {
  constexpr int kDevices = 64;
  static std::mutex mu;
  static int64_t cache[kDevices][2] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (cache[dev][slot] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu < 1 || cus < 1)
      cache[dev][slot] = -1;
    else
      cache[dev][slot] = (int64_t)per_cu * cus;
  }
  return cache[dev][slot] > 0 ? cache[dev][slot] : 0;
}

static dim3 slab_grid(int64_t S, const void* kernel, int slot) {
  const int64_t res = resident_groups(kernel, slot);
  return dim3((unsigned)(res > 0 ? std::min<int64_t>(S, res) : std::min<int64_t>(S, 2048)));
}

hipError_t launch_dft_fwd(int64_t S, const uint16_t* v, const uint16_t* mx, const uint16_t* myt, float* x, hipStream_t st) {
  hipLaunchKernelGGL(k_dft_fwd, slab_grid(S, reinterpret_cast<const void*>(&k_dft_fwd), 0), dim3(256), 0, st, S, v, mx, myt, x);
  return hipGetLastError();
}

hipError_t launch_dft_inv(int64_t S, const float* y, const uint16_t* nyt, const uint16_t* nx, uint16_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_dft_inv, slab_grid(S, reinterpret_cast<const void*>(&k_dft_inv), 1), dim3(256), 0, st, S, y, nyt, nx, out);
  return hipGetLastError();
}

hipError_t launch_mode_gemm(int M, int R, int N, int K, const uint16_t* A, const uint16_t* Bt, float* C, hipStream_t st) {
  hipLaunchKernelGGL(k_mode_gemm, dim3(N / 32, R / 32, M), dim3(64), 0, st, R, N, K, A, Bt, C);
  return hipGetLastError();
}

}  // namespace kw
