// Mass-conservation launcher of one size family (compiled once per family, -DKW_GROUP=0..3).
#include "kw_kernels.hpp"

namespace kw {

hipError_t KW_GNAME(launch_mass)(const Geom& g, int nb, const Fields& fl, const Inst* inst, const float2* tw, int flags, hipStream_t st, bool sreg, bool* handled) {
#define KW_M(S, D, R, P) hipLaunchKernelGGL((k_mass<S, D, R, P>), zgrid(g, ZCfg<S>::L, nb, 1), dim3(ZCfg<S>::NT), 0, st, g, fl, inst, tw, flags)
#define KW_MD(S, D)                                        \
  if (fl.adjD) {                                           \
    if (sreg) KW_M(S, D, true, true); else KW_M(S, D, false, true);   \
  } else {                                                 \
    if (sreg) KW_M(S, D, true, false); else KW_M(S, D, false, false); \
  }
  if (g.nd == 2) {
#define KW_C(S) case S: KW_MD(S, 2) break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  } else {
#define KW_C(S) case S: KW_MD(S, 3) break;
    KW_ZDISPATCH(KW_C)
#undef KW_C
  }
#undef KW_MD
#undef KW_M
  return hipGetLastError();
}

}  // namespace kw
