// Reverse-sweep launchers of one size family (compiled once per family, -DKW_GROUP=0..3).
#include "kw_adjoint.hpp"

namespace kw {

hipError_t KW_GNAME(launch_adj_eos)(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw, int t,
                                    hipStream_t st, bool* handled) {
  if (g.nd == 2) {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_eos<S, 2>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw, t); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  } else {
#define KW_C(S) case S: hipLaunchKernelGGL((k_adj_eos<S, 3>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw, t); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  }
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_adj_mom)(const Geom& g, int nb, const Fields& fl, const AdjFields& af, const Inst* inst, const float2* tw,
                                    hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_adj_mom<S>, zgrid(g, ZCfg<S>::L, nb, g.nd), dim3(ZCfg<S>::NT), 0, st, g, fl, af, inst, tw); break;
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
