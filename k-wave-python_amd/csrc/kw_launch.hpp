// Host-side launchers for the time-loop kernels (size dispatch in kw_dispatch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kw_types.hpp"

namespace kw {
// Any axis length with a transform: the compiled (fused) lengths, plus 2..4096 through the
// generic family of kw_generic.hip.
bool size_supported(int n);
bool compiled_size(int n);
// The contiguous-axis work of this grid goes to the generic family (length not
// compiled, or an odd number of rows).
bool generic_z(const Geom& g);
// Generic-length family: same contracts as the launchers below (kw_generic.hip).
hipError_t launch_zfwd_gen(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw, hipStream_t st);
hipError_t launch_zinv_gen(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale, hipStream_t st);
hipError_t launch_momentum_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st);
hipError_t launch_mass_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st);
hipError_t launch_eos_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st);
hipError_t launch_init_u_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st);
hipError_t launch_src_build_gen(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int which, hipStream_t st);
hipError_t launch_col_gen(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st);
hipError_t launch_adj_eos_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t, hipStream_t st);
hipError_t launch_adj_mom_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t, hipStream_t st);
hipError_t launch_adj_abs1_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, hipStream_t st);
hipError_t launch_adj_sum_gen(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, hipStream_t st);
hipError_t launch_zfwd(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw, hipStream_t st);
hipError_t launch_zinv(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale, hipStream_t st);
hipError_t launch_momentum(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st);
// sreg: some instance has a spectral (CHAIN) or PATTERN pressure source still running
hipError_t launch_mass(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st, bool sreg);
hipError_t launch_eos(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st);
hipError_t launch_init_u(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st);
hipError_t launch_src_build(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int which, hipStream_t st);
hipError_t launch_col(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st);
hipError_t launch_gather(const Geom& g, int nb, int max_sens, const Fields& fl, const Inst* inst, hipStream_t st);
hipError_t launch_tick(const Fields& fl, const Inst* inst, int nb, hipStream_t st);
hipError_t launch_fill(float* p, int64_t n, float v, hipStream_t st);
hipError_t launch_setint(int* p, int n, int v, hipStream_t st);
// surrogate (kw_fno.hip)
hipError_t launch_mode_gemm(int M, int R, int N, int K, const uint16_t* A, const uint16_t* Bt, float* C, hipStream_t st);
hipError_t launch_head_gather(int64_t B, int C, int K, const uint16_t* v, const int* cols, uint16_t* sel, hipStream_t st);
hipError_t launch_fno_lift(int64_t B, int C, int N, const float* cmap, float cmid, float chalf, const float* lc, const uint16_t* g,
                           uint16_t* v, hipStream_t st);
hipError_t launch_dft_fwd(int64_t S, const uint16_t* v, const uint16_t* mx, const uint16_t* myt, float* x, hipStream_t st);
hipError_t launch_dft_inv(int64_t S, const float* y, const uint16_t* nyt, const uint16_t* nx, uint16_t* out, hipStream_t st);
hipError_t launch_fno_pointwise(int B, int C, int64_t P, const uint16_t* v, const uint16_t* s, const float* pw, const float* pb,
                                uint16_t* out, int gelu, hipStream_t st);
hipError_t launch_mode_contract(int B, int K, int md, int NX, const float2* X, const int64_t xs[4], const uint16_t* Bt, float2* Y,
                                const int64_t ys[4], hipStream_t st);
// reverse sweep (kw_adjoint.hpp)
hipError_t launch_adj_scatter(const Geom& g, int nb, int max_sens, const AdjFields& af, const Inst* inst, int t, hipStream_t st);
hipError_t launch_adj_srcgrad(const Geom& g, int nb, const AdjFields& af, int which, const float* Zn, const float* Zk, const Inst* inst, int t,
                              hipStream_t st);
hipError_t launch_adj_p0(int64_t n, int D, const float* p0, const float* c0sq, int c0sq_s, const float* const rb[3], const float w[3],
                         float* cb, hipStream_t st);
hipError_t launch_adj_eos(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t, hipStream_t st);
hipError_t launch_adj_mom(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t, hipStream_t st);
hipError_t launch_adj_abs1(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, hipStream_t st);
hipError_t launch_adj_sum(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, hipStream_t st);
}  // namespace kw
