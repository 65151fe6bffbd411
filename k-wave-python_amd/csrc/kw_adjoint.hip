// Reverse-sweep launchers of one size family (compiled once per family, -DKW_GROUP=0..3).
#include "kw_adjoint.hpp"

namespace kw {

template <int ND, int ABS>
static hipError_t adj_eos_nd(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t,
                             hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_eos<S, ND, ABS>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw, t); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_adj_eos)(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t,
                                    hipStream_t st, bool* handled) {
  if (g.nd == 2) {
    if (af.absorb == 1) return adj_eos_nd<2, 1>(g, nb, fl, af, inst, tw, t, st, handled);
    if (af.absorb == 2) return adj_eos_nd<2, 2>(g, nb, fl, af, inst, tw, t, st, handled);
    return adj_eos_nd<2, 0>(g, nb, fl, af, inst, tw, t, st, handled);
  }
  if (af.absorb == 1) return adj_eos_nd<3, 1>(g, nb, fl, af, inst, tw, t, st, handled);
  if (af.absorb == 2) return adj_eos_nd<3, 2>(g, nb, fl, af, inst, tw, t, st, handled);
  return adj_eos_nd<3, 0>(g, nb, fl, af, inst, tw, t, st, handled);
}

hipError_t KW_GNAME(launch_adj_abs1)(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw,
                                     hipStream_t st, bool* handled) {
  if (g.nd == 2) {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_abs1<S, 2>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  } else {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_abs1<S, 3>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  }
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_adj_mom)(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw,
                                    int t, hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_adj_mom<S>, zgrid(g, ZCfg<S>::L, nb, g.nd), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw, t); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_adj_sum)(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw,
                                    hipStream_t st, bool* handled) {
  if (g.nd == 2) {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_sum<S, 2>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  } else {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_sum<S, 3>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  }
  return hipGetLastError();
}

}  // namespace kw
