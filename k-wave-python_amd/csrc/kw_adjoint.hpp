// Reverse-time (discrete adjoint) kernels of the step, for the
// gradient of a sensor-trace functional with respect to the sound speed.
//
// Forward step t (kw_kernels.hpp; reference kspace_solver.py:691-747):
//   u_i = a_i^2 u_i - a_i beta_i Gop_i p,   rho_i = b_i^2 rho_i - b_i dt rho0 Dop_i u_i + s_i,
//   p = c^2 sum_i rho_i,                    y = S p
// with a = pml_sg, b = pml, beta_i = dt / rho0_sg_i.  The operator tables are
// Hermitian, so Gop_i^T = -Dop_i and Dop_i^T = -Gop_i: the reverse step is the
// forward one with gradient and divergence swapped.  One reverse step:
//   k_adj_scatter   pbar += S^T r_t
//   k_adj_eos       rbar_i += c^2 pbar; cbar += 2 p pbar / c; R = sum_i fac_i rbar_i;
//                   w_i = b_i dt rho0 rbar_i -> F_z;  rbar_i = b_i^2 rbar_i
//   (x/y passes)    Gop_i on w_i
//   k_adj_mom       ubar_i += Gop_i w_i; v_i = a_i beta_i ubar_i -> F_z; ubar_i = a_i^2 ubar_i
//   (x/y passes)    Dop_i on v_i
//   k_adj_sum       pbar = sum_i Dop_i v_i
//   k_adj_srcgrad   cbar += -(scale / c) sig(t) [K R] at the source points
//
// Absorbing media (linear): the equation of state gains c^2 (tau A - eta B),
// A = L1[rho0 d] (Stokes: rho0 d), d = sum_i Dop_i u_i, B = L2[sum rho]; the forward
// replay keeps A and B of every step (Fields::adjA/adjB).  With w = c^2 pbar:
//   cbar += w ((y-1) tau A - y eta B) / c,  rbar_i += w - L2[eta w],
//   and the gradient jobs act on b dt rho0 rbar_i - rho0 L1[tau w] (Stokes: tau w).
// For the power law k_adj_abs1 forms tau w and eta w (forward z-FFT into spectra 0
// and 1) and the speed term; the fractional column passes apply L1, L2; k_adj_eos
// then reads both back.
//
// Dirichlet pressure source: at its points the step overwrote every rho_i with
// fac_i sig / (D c^2), so cbar gains -2 fac_i v rbar_i / c there and rbar_i is
// cleared before it is carried back through the mass update.
//
// B/A (nonlinear): rbar_i also gains w B/A S^t / rho0, the gradient-job input is
// scaled by (2 S^{t-1} + rho0) / rho0, and -2 dt sum_i b_i d_i^t rbar_i goes to every
// rbar_j of step t-1 (S and d_i = Dop_i u_i kept by the replay).
#pragma once
#include "kw_kernels.hpp"

namespace kw {

// ABS: 0 lossless, 1 Stokes, 2 power law (spectra 0 and 1 hold L1[tau w], L2[eta w])
template <int N, int ND, int ABS>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_adj_eos(Geom g, Fields fl, AdjFields af, const Inst* __restrict__ inst,
                                                         const float2* __restrict__ tw_g, int t) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  float2 l1[ABS == 2 ? C::E : 1], l2[ABS == 2 ? C::E : 1];
  if constexpr (ABS == 2) {
    inv_to_regs<N>(lds, fl.spec[0] + b * g.nspec, tw, row0, rows, g.nhp, tid, l1);
    if (af.disp) {
      inv_to_regs<N>(lds, fl.spec[1] + b * g.nspec, tw, row0, rows, g.nhp, tid, l2);
    } else {
#pragma unroll
      for (int m = 0; m < C::E; ++m) l2[m] = make_float2(0.f, 0.f);
    }
  }
  __shared__ float pmlz[N];
  for (int i_ = tid; i_ < N; i_ += C::NT) pmlz[i_] = gp(I.pml[ND - 1])[i_];
  __syncthreads();
  KW_ROWS_CLAMPED
  const int64_t off = b * g.ntot;
  const Src& src = I.srcp;
  const bool srcon = src.mode != SRC_NONE && t < src.len;
  const bool dir = srcon && src.mode == SRC_DIRICHLET;
  float ph[ND][2];
#pragma unroll
  for (int i = 0; i < ND; ++i) pml_rows(I.pml, g, i, rA, ph[i]);
  const auto pb = gp(af.pb + off);
  const auto cb = gw(af.cb + off);
  const auto pt = gp(af.pt + off);
  float2 pbv[C::E], c2v[C::E], rs[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int64_t i0 = iA + m * C::TPL;
    pbv[m] = make_float2(pb[i0], pb[i0 + N]);
    c2v[m] = make_float2(ld(I.c0sq, i0, I.c0sq_s), ld(I.c0sq, i0 + N, I.c0sq_s));
    rs[m] = make_float2(0.f, 0.f);
    if constexpr (ABS != 2) {  // the power law's speed terms are k_adj_abs1's
      // p = c^2 sum rho  ->  dJ/dc += 2 c sum(rho) pbar = 2 p pbar / c
      float ga = 2.f * pt[i0] * pbv[m].x / sqrtf(c2v[m].x);
      float gb = 2.f * pt[i0 + N] * pbv[m].y / sqrtf(c2v[m].y);
      if constexpr (ABS == 1) {  // Stokes: tau = -2 a c, d tau/dc = tau / c, operand A = rho0 d
        ga += c2v[m].x * pbv[m].x * tau_at(I, i0, c2v[m].x) / sqrtf(c2v[m].x) * af.At[off + i0];
        gb += c2v[m].y * pbv[m].y * tau_at(I, i0 + N, c2v[m].y) / sqrtf(c2v[m].y) * af.At[off + i0 + N];
      }
      if (inside) {
        cb[i0] = cb[i0] + ga;
        cb[i0 + N] = cb[i0 + N] + gb;
      }
    }
  }
  // B/A: the adjoint of S^{t-1} through the mass update's factor, -2 dt sum_i b_i d_i rbar_i,
  // goes to every rbar_j of the step before (added after the axis loop)
  float2 sb[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) sb[m] = make_float2(0.f, 0.f);
  const int64_t nf = (int64_t)gridDim.y * g.ntot;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const auto rb = gw(af.rb[i] + off);
    float2 w[C::E];
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      const float sz = (i == ND - 1) ? pmlz[j0 + m * C::TPL] : 1.f;
      float wv[2], rn[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t idx = iA + h * N + m * C::TPL;
        const float s = sz * ph[i][h];
        const float wc = (h ? c2v[m].y : c2v[m].x) * (h ? pbv[m].y : pbv[m].x);  // c^2 pbar
        const float r0 = ld(I.rho0, idx, I.rho0_s);
        float rbt = rb[idx] + wc;
        float nl = 1.f;
        if (af.nonlin) {  // EOS: c^2 B/A S^2 / (2 rho0);  mass: d_i (2 S^{t-1} + rho0) / rho0
          const float bona = ld(I.bona, idx, I.bona_s);
          rbt += wc * bona * af.St[off + idx] / r0;
          nl = 1.f + 2.f * af.Sp[off + idx] / r0;
        }
        float dbar = 0.f;  // adjoint of div u through the absorption term, over rho0
        if constexpr (ABS == 1) dbar = tau_at(I, idx, h ? c2v[m].y : c2v[m].x) * wc;
        if constexpr (ABS == 2) {
          rbt -= h ? l2[m].y : l2[m].x;
          dbar = h ? l1[m].y : l1[m].x;
        }
        if (dir) {  // dirichlet point: rho_i = fac_i v after the mass update, v = sig / (D c^2)
          float v;
          if (csr_lookup(src, rA + h, j0 + m * C::TPL, t, v)) {
            if (inside) cb[idx] += -2.f * src.fac[i] * v * rbt / sqrtf(h ? c2v[m].y : c2v[m].x);
            rbt = 0.f;  // the overwritten value carries no adjoint further back
          }
        }
        if (h) rs[m].y += src.fac[i] * rbt; else rs[m].x += src.fac[i] * rbt;
        if (af.nonlin) {
          const float sv = -2.f * I.dt * s * af.Dt[(int64_t)i * nf + off + idx] * rbt;
          if (h) sb[m].y += sv; else sb[m].x += sv;
        }
        wv[h] = s * I.dt * r0 * nl * rbt - r0 * dbar;
        rn[h] = s * s * rbt;
      }
      if (inside) {
        rb[iA + m * C::TPL] = rn[0];
        rb[iA + N + m * C::TPL] = rn[1];
      }
      w[m] = make_float2(wv[0], wv[1]);
    }
    fwd_from_regs<N>(lds, w, fl.spec[i] + b * g.nspec, tw, row0, rows, g.nhp, tid);
  }
  if (af.nonlin && inside) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      const auto rb = gw(af.rb[i] + off);
#pragma unroll
      for (int m = 0; m < C::E; ++m) {
        rb[iA + m * C::TPL] += sb[m].x;
        rb[iA + N + m * C::TPL] += sb[m].y;
      }
    }
  }
  if (srcon && inside) {
    const auto R = gw(af.R + off);
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      R[iA + m * C::TPL] = rs[m].x;
      R[iA + N + m * C::TPL] = rs[m].y;
    }
  }
}

// Power-law pre-pass of a reverse step: w = c^2 pbar,
//   cbar += 2 p pbar / c + w ((y-1) tau A - y eta B) / c,
// and the forward z-FFTs of tau w (spectrum 0) and eta w (spectrum 1, with dispersion).
template <int N, int ND>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_adj_abs1(Geom g, Fields fl, AdjFields af, const Inst* __restrict__ inst,
                                                          const float2* __restrict__ tw_g) {
  KW_ZPASS_PROLOGUE
  const Inst& I = inst[b];
  KW_ROWS_CLAMPED
  const int64_t off = b * g.ntot;
  const auto pb = gp(af.pb + off);
  const auto pt = gp(af.pt + off);
  const auto cb = gw(af.cb + off);
  float2 tw_v[C::E], ew_v[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    float tv[2], ev[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t idx = iA + h * N + m * C::TPL;
      const float c2 = ld(I.c0sq, idx, I.c0sq_s), c = sqrtf(c2);
      const float pbar = pb[idx];
      const float w = c2 * pbar;
      const float tau = tau_at(I, idx, c2);
      const float eta = af.disp ? eta_at(I, idx, c2) : 0.f;
      float gc = 2.f * pt[idx] * pbar / c + w * af.ym1 * tau / c * af.At[off + idx];
      if (af.disp) gc -= w * af.y * eta / c * af.Bt[off + idx];
      if (inside) cb[idx] = cb[idx] + gc;
      tv[h] = tau * w;
      ev[h] = eta * w;
    }
    tw_v[m] = make_float2(tv[0], tv[1]);
    ew_v[m] = make_float2(ev[0], ev[1]);
  }
  fwd_from_regs<N>(lds, tw_v, fl.spec[0] + b * g.nspec, tw, row0, rows, g.nhp, tid);
  if (af.disp) fwd_from_regs<N>(lds, ew_v, fl.spec[1] + b * g.nspec, tw, row0, rows, g.nhp, tid);
}

template <int N>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_adj_mom(Geom g, Fields fl, AdjFields af, const Inst* __restrict__ inst,
                                                         const float2* __restrict__ tw_g, int t) {
  KW_ZPASS_PROLOGUE
  const int ax = blockIdx.z;
  const Inst& I = inst[b];
  __shared__ float pmlz[N];  // read after the first transform barrier
  const bool zax = ax == g.nd - 1;
  if (zax)
    for (int i_ = tid; i_ < N; i_ += C::NT) pmlz[i_] = gp(I.pmlsg[ax])[i_];
  KW_ROWS_CLAMPED
  float2* sp = fl.spec[ax] + b * g.nspec;
  const auto ub = gw(af.ub[ax] + b * g.ntot);
  // velocity source of this axis at step t: dirichlet points drop their adjoint; for an
  // additive one dJ/du_i^t goes to U (its speed term needs K applied, k_adj_srcgrad)
  const Src& us = I.srcu[ax];
  const bool uon = us.mode != SRC_NONE && t < us.len;
  const bool udir = uon && us.mode == SRC_DIRICHLET;
  float* const Ub = uon && !udir && af.U[ax] ? af.U[ax] + b * g.ntot : nullptr;
  float ph[2];
  pml_rows(I.pmlsg, g, ax, rA, ph);
  float2 uo[C::E], dr[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const int64_t i0 = iA + m * C::TPL;
    uo[m] = make_float2(ub[i0], ub[i0 + N]);
    dr[m] = make_float2(ld(I.dtr0[ax], i0, I.dtr0_s), ld(I.dtr0[ax], i0 + N, I.dtr0_s));
  }
  float2 gv[C::E];
  inv_to_regs<N>(lds, sp, tw, row0, rows, g.nhp, tid, gv);
#pragma unroll
  for (int m = 0; m < C::E; ++m) {
    const float sz = zax ? pmlz[j0 + m * C::TPL] : 1.f;
    float un[2], vv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float s = sz * ph[h];
      float u = (h ? uo[m].y : uo[m].x) + (h ? gv[m].y : gv[m].x);
      if (udir) {
        float v;
        if (csr_lookup(us, rA + h, j0 + m * C::TPL, t, v)) u = 0.f;
      }
      if (Ub && inside) Ub[iA + h * N + m * C::TPL] = u;
      vv[h] = s * (h ? dr[m].y : dr[m].x) * u;
      un[h] = s * s * u;
    }
    if (inside) {
      ub[iA + m * C::TPL] = un[0];
      ub[iA + N + m * C::TPL] = un[1];
    }
    gv[m] = make_float2(vv[0], vv[1]);
  }
  fwd_from_regs<N>(lds, gv, sp, tw, row0, rows, g.nhp, tid);
}

template <int N, int ND>
__global__ __launch_bounds__(ZCfg<N>::NT) void k_adj_sum(Geom g, Fields fl, AdjFields af, const Inst* __restrict__ inst,
                                                         const float2* __restrict__ tw_g) {
  KW_ZPASS_PROLOGUE
  (void)inst;
  KW_ROWS_CLAMPED
  float2 acc[C::E];
#pragma unroll
  for (int m = 0; m < C::E; ++m) acc[m] = make_float2(0.f, 0.f);
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    float2 d[C::E];
    inv_to_regs<N>(lds, fl.spec[i] + b * g.nspec, tw, row0, rows, g.nhp, tid, d);
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      acc[m].x += d[m].x;
      acc[m].y += d[m].y;
    }
  }
  if (inside) {
    const auto pb = gw(af.pb + b * g.ntot);
#pragma unroll
    for (int m = 0; m < C::E; ++m) {
      pb[iA + m * C::TPL] = acc[m].x;
      pb[iA + N + m * C::TPL] = acc[m].y;
    }
  }
}

}  // namespace kw
