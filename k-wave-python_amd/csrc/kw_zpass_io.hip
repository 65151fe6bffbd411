// Launchers for the contiguous-axis transform kernels of one size family
// (compiled once per family, -DKW_GROUP=0..3; see kw_kernels.hpp).
#include "kw_kernels.hpp"

namespace kw {

hipError_t KW_GNAME(launch_zfwd)(const Geom& g, int nb, const float* f, int64_t fs, float2* spec, int64_t ss, const float2* tw, hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_zfwd<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, f, fs, spec, ss, tw); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_zinv)(const Geom& g, int nb, const float2* spec, int64_t ss, float* f, int64_t fs, const float2* tw, float scale, hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_zinv<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, spec, ss, f, fs, tw, scale); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_init_u)(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_init_u<S>, zgrid(g, ZCfg<S>::L, nb, g.nd), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

hipError_t KW_GNAME(launch_src_build)(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int which, hipStream_t st, bool* handled) {
#define KW_C(S) case S: hipLaunchKernelGGL(k_src_build<S>, zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, which); break;
  KW_ZDISPATCH(KW_C)
#undef KW_C
  return hipGetLastError();
}

}  // namespace kw
