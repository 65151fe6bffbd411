// Size dispatch: each launcher is defined once per size family (f_g0 .. f_g3,
// see kw_kernels.hpp); the entry points below hand a call to the family that
// holds the line length, and every other grid to the generic-length family
// (kw_generic.hip): contiguous-axis work whose length is not compiled or whose
// row count is odd, column passes whose length is not compiled.
#include "kw_launch.hpp"

namespace kw {

#define KW_DECL4(RET, NAME, ...) \
  RET NAME##_g0(__VA_ARGS__, bool*); RET NAME##_g1(__VA_ARGS__, bool*); RET NAME##_g2(__VA_ARGS__, bool*); RET NAME##_g3(__VA_ARGS__, bool*);
#define KW_TRY4(NAME, ...)                                    \
  do {                                                        \
    bool h = false;                                           \
    hipError_t e = NAME##_g0(__VA_ARGS__, &h); if (h) return e; \
    e = NAME##_g1(__VA_ARGS__, &h); if (h) return e;            \
    e = NAME##_g2(__VA_ARGS__, &h); if (h) return e;            \
    e = NAME##_g3(__VA_ARGS__, &h); if (h) return e;            \
    return hipErrorInvalidValue;                              \
  } while (0)

KW_DECL4(hipError_t, launch_zfwd, const Geom&, int, const float*, int64_t, float2*, int64_t, const float2*, hipStream_t)
KW_DECL4(hipError_t, launch_zinv, const Geom&, int, const float2*, int64_t, float*, int64_t, const float2*, float, hipStream_t)
KW_DECL4(hipError_t, launch_init_u, const Geom&, int, const Fields&, const Inst*, const float2*, hipStream_t)
KW_DECL4(hipError_t, launch_src_build, const Geom&, int, const Fields&, const Inst*, const float2*, int, hipStream_t)
KW_DECL4(hipError_t, launch_momentum, const Geom&, int, const Fields&, const Inst*, const float2*, hipStream_t)
KW_DECL4(hipError_t, launch_mass, const Geom&, int, const Fields&, const Inst*, const float2*, int, hipStream_t, bool)
KW_DECL4(hipError_t, launch_eos, const Geom&, int, const Fields&, const Inst*, const float2*, int, hipStream_t)
KW_DECL4(hipError_t, launch_col, const Geom&, int, const Fields&, const Ops&, const Inst*, const ColArgs&, int, hipStream_t)
KW_DECL4(hipError_t, launch_adj_eos, const Geom&, int, const Fields&, const AdjFields&, const Inst*, const float2*, int, hipStream_t)
KW_DECL4(hipError_t, launch_adj_mom, const Geom&, int, const Fields&, const AdjFields&, const Inst*, const float2*, int, hipStream_t)
KW_DECL4(hipError_t, launch_adj_abs1, const Geom&, int, const Fields&, const AdjFields&, const Inst*, const float2*, hipStream_t)
KW_DECL4(hipError_t, launch_adj_sum, const Geom&, int, const Fields&, const AdjFields&, const Inst*, const float2*, hipStream_t)

hipError_t launch_zfwd(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw, hipStream_t st) {
  if (generic_z(g)) return launch_zfwd_gen(g, nb, f, fs, spec, ss, tw, st);
  KW_TRY4(launch_zfwd, g, nb, f, fs, spec, ss, tw, st);
}
hipError_t launch_zinv(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale, hipStream_t st) {
  if (generic_z(g)) return launch_zinv_gen(g, nb, spec, ss, f, fs, tw, scale, st);
  KW_TRY4(launch_zinv, g, nb, spec, ss, f, fs, tw, scale, st);
}
hipError_t launch_init_u(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st) {
  if (generic_z(g)) return launch_init_u_gen(g, nb, fl, inst, tw, st);
  KW_TRY4(launch_init_u, g, nb, fl, inst, tw, st);
}
hipError_t launch_src_build(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int which, hipStream_t st) {
  if (generic_z(g)) return launch_src_build_gen(g, nb, fl, inst, tw, which, st);
  KW_TRY4(launch_src_build, g, nb, fl, inst, tw, which, st);
}
hipError_t launch_momentum(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st) {
  if (generic_z(g)) return launch_momentum_gen(g, nb, fl, inst, tw, st);
  KW_TRY4(launch_momentum, g, nb, fl, inst, tw, st);
}
hipError_t launch_mass(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st, bool sreg) {
  if (generic_z(g)) return launch_mass_gen(g, nb, fl, inst, tw, flags, st);
  KW_TRY4(launch_mass, g, nb, fl, inst, tw, flags, st, sreg);
}
hipError_t launch_eos(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st) {
  if (generic_z(g)) return launch_eos_gen(g, nb, fl, inst, tw, flags, st);
  KW_TRY4(launch_eos, g, nb, fl, inst, tw, flags, st);
}
hipError_t launch_col(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st) {
  if (!compiled_size(nline)) return launch_col_gen(g, nb, fl, ops, inst, a, nline, st);
  KW_TRY4(launch_col, g, nb, fl, ops, inst, a, nline, st);
}

hipError_t launch_adj_eos(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t, hipStream_t st) {
  if (generic_z(g)) return launch_adj_eos_gen(g, nb, fl, af, inst, tw, t, st);
  KW_TRY4(launch_adj_eos, g, nb, fl, af, inst, tw, t, st);
}
hipError_t launch_adj_mom(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t, hipStream_t st) {
  if (generic_z(g)) return launch_adj_mom_gen(g, nb, fl, af, inst, tw, t, st);
  KW_TRY4(launch_adj_mom, g, nb, fl, af, inst, tw, t, st);
}
hipError_t launch_adj_abs1(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, hipStream_t st) {
  if (generic_z(g)) return launch_adj_abs1_gen(g, nb, fl, af, inst, tw, st);
  KW_TRY4(launch_adj_abs1, g, nb, fl, af, inst, tw, st);
}
hipError_t launch_adj_sum(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, hipStream_t st) {
  if (generic_z(g)) return launch_adj_sum_gen(g, nb, fl, af, inst, tw, st);
  KW_TRY4(launch_adj_sum, g, nb, fl, af, inst, tw, st);
}

}  // namespace kw
