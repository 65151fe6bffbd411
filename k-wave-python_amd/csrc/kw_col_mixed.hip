// Column-pass launchers (see kw_kernels.hpp).
#include "kw_kernels.hpp"

namespace kw {

#define KW_SUBSET(X) X(24) X(48) X(96) X(192) X(384) X(768) X(40) X(80) X(160) X(320) X(640) X(56) X(112) X(224) X(448)

hipError_t launch_col_mixed(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st, bool* handled) {
#define KW_C(S)                                                                                                                \
  case S: {                                                                                                                    \
    constexpr int LB = col_lbig<S>(), LS = col_lsmall<S>();                                                                    \
    const int64_t wide_blocks = (int64_t)a.nouter * nb * ((a.ncols + LB - 1) / LB);                                             \
    if (wide_blocks >= 1024 || LB == LS) {                                                                                     \
      const int nt = (a.ncols + LB - 1) / LB;                                                                                  \
      if (a.mode == 2) hipLaunchKernelGGL((k_col<S, LB, true>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, LB>::NT), 0, st, g, fl, ops, inst, a); \
      else hipLaunchKernelGGL((k_col<S, LB, false>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, LB>::NT), 0, st, g, fl, ops, inst, a);         \
    } else {                                                                                                                   \
      const int nt = (a.ncols + LS - 1) / LS;                                                                                  \
      if (a.mode == 2) hipLaunchKernelGGL((k_col<S, LS, true>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, LS>::NT), 0, st, g, fl, ops, inst, a); \
      else hipLaunchKernelGGL((k_col<S, LS, false>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, LS>::NT), 0, st, g, fl, ops, inst, a);         \
    }                                                                                                                          \
    break;                                                                                                                     \
  }
  *handled = true;
  switch (nline) {
    KW_SUBSET(KW_C)
    default: *handled = false; return hipSuccess;
  }
#undef KW_C
  return hipGetLastError();
}

}  // namespace kw
