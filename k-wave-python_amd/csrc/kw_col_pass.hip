// Strided-axis (column) pass launcher of one size family (compiled once per
// family, -DKW_GROUP=0..3; see kw_kernels.hpp).
#include "kw_kernels.hpp"

namespace kw {

// One launch at tile width LW: wide tiles when they still give at least 1024
// workgroups (or a.wide is set), the narrow col_lsmall width otherwise (small grids).
template <int S, int LW, bool OPS>
static hipError_t col_launch(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a,
                             hipStream_t st) {
  constexpr int LS = col_lsmall<S>();
  const int ntw = (a.ncols + LW - 1) / LW;
  if ((int64_t)a.nouter * nb * ntw >= 1024 || LW == LS || a.wide) {
    hipLaunchKernelGGL((k_col<S, LW, OPS>), dim3(a.nouter * ntw, nb), dim3(ColCfg<S, LW>::NT), 0, st, g, fl, ops, inst, a);
  } else {
    const int nt = (a.ncols + LS - 1) / LS;
    hipLaunchKernelGGL((k_col<S, LS, OPS>), dim3(a.nouter * nt, nb), dim3(ColCfg<S, LS>::NT), 0, st, g, fl, ops, inst, a);
  }
  return hipGetLastError();
}

template <int W>
constexpr int half_width() { return W / 2 < 1 ? 1 : W / 2; }

// Tile widths: FFT-only passes col_lbig or half of it, passes with spectral jobs
// col_lops or half of it (a.tile selects; the plan picks by timing both).
hipError_t KW_GNAME(launch_col)(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st, bool* handled) {
#define KW_C(S)                                                                                                          \
  case S:                                                                                                                \
    if (a.mode == 2)                                                                                                     \
      return a.tile ? col_launch<S, half_width<col_lops<S>()>(), true>(g, nb, fl, ops, inst, a, st)                      \
                    : col_launch<S, col_lops<S>(), true>(g, nb, fl, ops, inst, a, st);                                   \
    return a.tile ? col_launch<S, half_width<col_lbig<S>()>(), false>(g, nb, fl, ops, inst, a, st)                       \
                  : col_launch<S, col_lbig<S>(), false>(g, nb, fl, ops, inst, a, st);
  *handled = true;
  switch (nline) {
    KW_GROUP_SIZES(KW_C)
    default: *handled = false; return hipSuccess;
  }
#undef KW_C
  return hipGetLastError();
}

}  // namespace kw
