// Column-pass launchers (see kw_kernels.hpp).
#include "kw_kernels.hpp"

namespace kw {

hipError_t launch_col_p2(const Geom&, int, const Fields&, const Ops&, const Inst*, const ColArgs&, int, hipStream_t, bool*);
hipError_t launch_col_mixed(const Geom&, int, const Fields&, const Ops&, const Inst*, const ColArgs&, int, hipStream_t, bool*);

hipError_t launch_col(const Geom& g, int nb, const Fields& fl, const Ops& ops, const Inst* inst, const ColArgs& a, int nline, hipStream_t st) {
  bool handled = false;
  hipError_t e = launch_col_p2(g, nb, fl, ops, inst, a, nline, st, &handled);
  if (handled) return e;
  e = launch_col_mixed(g, nb, fl, ops, inst, a, nline, st, &handled);
  return handled ? e : hipErrorInvalidValue;
}

}  // namespace kw
