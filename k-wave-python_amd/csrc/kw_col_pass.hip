// Strided-axis (column) pass launcher of one size family (compiled once per
// family, -DKW_GROUP=0..3; see kw_kernels.hpp).
#include <algorithm>
#include <mutex>

#include "kw_kernels.hpp"

namespace kw {

constexpr int kColDevices = 64;

// Workgroups of k_col<S, L, OPS> resident on the current device at once (0 if
// the runtime cannot say).  Queried once per kernel and device.
template <int S, int L, bool OPS, int PV>
static int col_resident() {
  static std::mutex mu;
  static int per_dev[kColDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kColDevices) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (per_dev[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_col<S, L, OPS, 1, PV>),
                                                     ColCfg<S, L, 1, PV>::NT, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      per_dev[dev] = -1;
    else
      per_dev[dev] = per_cu > 0 && cus > 0 ? per_cu * cus : -1;
  }
  return per_dev[dev] > 0 ? per_dev[dev] : 0;
}

// Workgroups per instance of a column pass: every tile its own workgroup, or -
// when the tiles of all instances exceed what is resident at once - a resident
// grid (a multiple of the 8 XCDs) whose workgroups loop over tiles and load the
// next tile's input during the current one.  a.persist selects the passes that
// loop (bit 0 those with spectral jobs, bit 1 the FFT-only ones); a.grid > 0
// sets the looping grid (tests).  Results do not depend on either: every column
// is transformed the same way.
template <int S, int L, bool OPS, int PV>
static int col_grid(const ColArgs& a, int tiles, int nb) {
  if (!(a.persist & (OPS ? 1 : 2))) return tiles;
  if (a.grid > 0) return std::min(tiles, a.grid);
  const int res = col_resident<S, L, OPS, PV>();
  if (res <= 0 || (int64_t)tiles * nb <= res) return tiles;
  const int per = std::max(8, (res / nb) & ~7);
  return std::min(tiles, per);
}

// One launch at tile width LW: wide tiles when they still give at least 1024
// workgroups (or a.wide is set), the narrow col_lsmall width otherwise (small grids).
template <int S, int LW, bool OPS, int PV>
static hipError_t col_launch_v(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a,
                               hipStream_t st) {
  constexpr int LS = col_lsmall<S>();
  const int ntw = (a.ncols + LW - 1) / LW;
  if ((int64_t)a.nouter * nb * ntw >= 1024 || LW == LS || a.wide) {
    const int gx = col_grid<S, LW, OPS, PV>(a, a.nouter * ntw, nb);
    hipLaunchKernelGGL((k_col<S, LW, OPS, 1, PV>), dim3(gx, nb), dim3(ColCfg<S, LW, 1, PV>::NT), 0, st, g, fl, ops, inst, a);
  } else {
    const int nt = (a.ncols + LS - 1) / LS;
    const int gx = col_grid<S, LS, OPS, PV>(a, a.nouter * nt, nb);
    hipLaunchKernelGGL((k_col<S, LS, OPS, 1, PV>), dim3(gx, nb), dim3(ColCfg<S, LS, 1, PV>::NT), 0, st, g, fl, ops, inst, a);
  }
  return hipGetLastError();
}

// the plan variant a.pv where the length has one, variant 0 otherwise
template <int S, int LW, bool OPS>
static hipError_t col_launch(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a,
                             hipStream_t st) {
  if constexpr (has_plan_variant(S)) {
    if (a.pv) return col_launch_v<S, LW, OPS, 1>(g, nb, fl, ops, inst, a, st);
  }
  return col_launch_v<S, LW, OPS, 0>(g, nb, fl, ops, inst, a, st);
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
