// Mass-conservation launcher of one size family (compiled once per family, -DKW_GROUP=0..3).
#include "kw_kernels.hpp"

namespace kw {

hipError_t KW_GNAME(launch_mass)(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st, bool sreg, bool* handled) {
  if (g.nd == 2) {
#define KW_C(S) case S: if (sreg) hipLaunchKernelGGL((k_mass<S, 2, true>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, flags); else hipLaunchKernelGGL((k_mass<S, 2, false>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, flags); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  } else {
#define KW_C(S) case S: if (sreg) hipLaunchKernelGGL((k_mass<S, 3, true>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, flags); else hipLaunchKernelGGL((k_mass<S, 3, false>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, flags); break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  }
  return hipGetLastError();
}

}  // namespace kw
