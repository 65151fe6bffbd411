// Strided-axis (column) pass launcher of one size family (compiled once per
// family, -DKW_GROUP=0..3; see kw_kernels.hpp).
#include "kw_kernels.hpp"

namespace kw {

// Wide tiles (col_lbig columns per workgroup, col_lops for the pass with
// spectral jobs) when they still give at least 1024 workgroups, narrow ones
// otherwise (small grids).
hipError_t KW_GNAME(launch_col)(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st, bool* handled) {
#define KW_COLL(S, OPS, LW)                                                                                                    \
  {                                                                                                                            \
    constexpr int LS = col_lsmall<S>();                                                                                        \
    const int ntw = (a.ncols + LW - 1) / LW;                                                                                   \
    if ((int64_t)a.nouter * nb * ntw >= 1024 || LW == LS) {                                                                    \
      hipLaunchKernelGGL((k_col<S, LW, OPS>), dim3(a.nouter * ntw, nb), dim3(ColCfg<S, LW>::NT), 0, st, g, fl, ops, inst, a); \
    } else {                                                                                                                   \
      const int nt = (a.ncols + LS - 1) / LS;                                                                                  \
      hipLaunchKernelGGL((k_col<S, LS, OPS>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, LS>::NT), 0, st, g, fl, ops, inst, a);  \
    }                                                                                                                          \
  }
#define KW_C(S)                                                \
  case S:                                                      \
    if (a.mode == 2) KW_COLL(S, true, col_lops<S>())           \
    else KW_COLL(S, false, col_lbig<S>())                      \
    break;
  *handled = true;
  switch (nline) {
    KW_GROUP_SIZES(KW_C)
    default: *handled = false; return hipSuccess;
  }
#undef KW_C
#undef KW_COLL
  return hipGetLastError();
}

}  // namespace kw
