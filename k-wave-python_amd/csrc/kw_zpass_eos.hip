// Equation-of-state launcher of one size family (compiled once per family, -DKW_GROUP=0..3).
#include "kw_kernels.hpp"

namespace kw {

hipError_t KW_GNAME(launch_eos)(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_eos<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, flags); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

}  // namespace kw
