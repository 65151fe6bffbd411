// Allocation-free spectral operator: kw_spectral_apply_static (include/kwgpu.h).
//
// The same transform sequence as kw_spectral_apply (forward real-to-half-complex
// transform along the contiguous axis, column passes with the operator applied
// between the forward and inverse transforms of the outermost axis, inverse
// transform back to real values), for callers that cannot allocate device
// memory: the input and output are the caller's device arrays, and everything
// else -- the half spectrum, the 1-D operator and twiddle tables and the
// instance record the column kernels read -- lives in `scratch`, a file-scope
// device array of kScratchBytes per GPU that is held until the process ends.
//
// The 1-D tables are built on the host (the same FP64 construction as a plan's)
// into fixed host arrays in static storage, under a lock, and copied into the
// scratch array before the lock is released; no heap memory is taken.
//
// Host code never uses the array's name as an address (in host code it is a
// stand-in symbol, not device memory): the device address of each GPU's copy is
// looked up once per device with hipGetSymbolAddress, with that device current,
// and kept in g_scratch[device].
//
// Every call on one GPU uses that GPU's one array, so calls that overlap on a
// GPU (two host threads, or two streams in flight) race on it.  Nothing in the
// array has to start at zero: each call writes every value it reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/kwgpu.h"
#include "kw_host.hpp"
#include "kw_launch.hpp"
#include "kw_types.hpp"

using namespace kw;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Host staging of the tables of one call (static storage, guarded by g_table_mu).
constexpr int kMaxLen = 4096;  // longest axis with a transform
std::mutex g_table_mu;
float2 h_tab[3][4][kMaxLen];   // per axis: grad, div, unst, tw
float h_k[3][kMaxLen];
Inst h_inst;

// Device address of this GPU's scratch array (looked up on the first call per device).
int scratch_on(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(KW_EINVAL, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(KW_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return KW_OK;
}

constexpr size_t up256(size_t v) { return (v + 255) & ~size_t(255); }

// Bump allocation inside the scratch array (offsets only; no device call).
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += up256(count * sizeof(T));
    return p;
  }
};

// Bytes of scratch one call needs: the half spectrum (twice when a column length
// goes through the generic transforms, which stash a forward transform), five 1-D
// tables per axis, one Inst.
bool needs_stash(const Geom& g) {
  for (int ax = 0; ax < g.nd - 1; ++ax)
    if (!compiled_size(g.n[ax])) return true;
  return false;
}
size_t scratch_need(const Geom& g) {
  size_t b = (needs_stash(g) ? 2 : 1) * up256(sizeof(float2) * (size_t)g.nspec) + up256(sizeof(Inst));
  for (int ax = 0; ax < g.nd; ++ax) b += 4 * up256(sizeof(float2) * g.n[ax]) + up256(sizeof(float) * g.n[ax]);
  return b;
}

#define KW_S_HIP(call)                                                                                   \
  do {                                                                                                   \
    const hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return set_error(KW_EHIP, std::string(#call ": ") + hipGetErrorString(e_));    \
  } while (0)

}  // namespace

extern "C" int kw_spectral_apply_static(int ndim, const int n[3], const double spacing[3], int use_sg, double kap, int kind,
                                        int axis, double power, const float* f, float* out, void* stream) {
  if (!n || !spacing || !f || !out) return set_error(KW_EINVAL, "NULL argument");
  if (ndim != 2 && ndim != 3) return set_error(KW_EUNSUPPORTED, "only 2-D and 3-D grids run on the GPU path");
  if (kind < 0 || kind > 6 || kind == 4 || axis < 0 || axis >= ndim) return set_error(KW_EINVAL, "bad operator kind or axis");
  for (int i = 0; i < ndim; ++i) {
    if (n[i] < 2) return set_error(KW_EINVAL, "every axis needs at least 2 points");
    if (!size_supported(n[i])) return set_error(KW_EUNSUPPORTED, "axis length " + std::to_string(n[i]) + " has no transform");
    if (!(spacing[i] > 0)) return set_error(KW_EINVAL, "grid spacing must be positive");
  }
  const Geom g = make_geom(ndim, n);
  const size_t need = scratch_need(g);
  if (need > kScratchBytes)
    return set_error(KW_EINVAL, "grid needs " + std::to_string(need) + " bytes of scratch; the allocation-free path holds " +
                                    std::to_string(kScratchBytes));
  int dev = 0;
  KW_S_HIP(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  if (const int rc = scratch_on(dev, &base)) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);

  // host tables (FP64-built, as for a plan), copied into the carved scratch regions
  Carve cv{base};
  Fields fl{};
  fl.spec[0] = cv.take<float2>((size_t)g.nspec);
  if (needs_stash(g)) fl.gstash = cv.take<float2>((size_t)g.nspec);
  Ops ops{};
  Inst* di = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_table_mu);
    for (int ax = 0; ax < ndim; ++ax) {
      float2* gd = cv.take<float2>(n[ax]);
      float2* dv = cv.take<float2>(n[ax]);
      float2* un = cv.take<float2>(n[ax]);
      float2* tw = cv.take<float2>(n[ax]);
      float* kv = cv.take<float>(n[ax]);
      build_axis_into(n[ax], spacing[ax], use_sg != 0, h_tab[ax][0], h_tab[ax][1], h_tab[ax][2], h_tab[ax][3], h_k[ax]);
      KW_S_HIP(hipMemcpyAsync(gd, h_tab[ax][0], sizeof(float2) * n[ax], hipMemcpyHostToDevice, st));
      KW_S_HIP(hipMemcpyAsync(dv, h_tab[ax][1], sizeof(float2) * n[ax], hipMemcpyHostToDevice, st));
      KW_S_HIP(hipMemcpyAsync(un, h_tab[ax][2], sizeof(float2) * n[ax], hipMemcpyHostToDevice, st));
      KW_S_HIP(hipMemcpyAsync(tw, h_tab[ax][3], sizeof(float2) * n[ax], hipMemcpyHostToDevice, st));
      KW_S_HIP(hipMemcpyAsync(kv, h_k[ax], sizeof(float) * n[ax], hipMemcpyHostToDevice, st));
      ops.grad[ax] = gd; ops.div[ax] = dv; ops.unst[ax] = un; ops.tw[ax] = tw; ops.k[ax] = kv;
    }
    h_inst = Inst{};
    h_inst.nt = 1;
    h_inst.kap = (float)kap;
    di = cv.take<Inst>(1);
    KW_S_HIP(hipMemcpyAsync(di, &h_inst, sizeof(Inst), hipMemcpyHostToDevice, st));
    // the staging arrays are reused by the next call: the copies finish under the lock
    KW_S_HIP(hipStreamSynchronize(st));
  }
  ops.inv_ntot = (float)(1.0 / (double)g.ntot);
  ops.frac1 = (float)power;        // |k|^power for kind 3
  ops.frac2 = (float)(power + 1.0);

  static const int kinds[7] = {XK_COPY, XK_GRAD, XK_DIV, XK_FRAC1, XK_FRAC2, XK_SRCKAPPA, XK_UNSTAGGER};
  const float2* twz = ops.tw[ndim - 1];
  KW_S_HIP(launch_zfwd(g, 1, f, 0, fl.spec[0], 0, twz, st));
  if (ndim == 3) {
    ColArgs a = col_args(g, 1, 0);
    a.nbuf = 1;
    a.bufs[0] = 0;
    KW_S_HIP(launch_col(g, 1, fl, ops, di, a, g.n[1], st));
  }
  {
    ColArgs a = col_args(g, 0, 2);
    a.njob = 1;
    a.job[0] = {0, 0, kinds[kind], axis};
    KW_S_HIP(launch_col(g, 1, fl, ops, di, a, g.n[0], st));
  }
  if (ndim == 3) {
    ColArgs a = col_args(g, 1, 1);
    a.nbuf = 1;
    a.bufs[0] = 0;
    KW_S_HIP(launch_col(g, 1, fl, ops, di, a, g.n[1], st));
  }
  KW_S_HIP(launch_zinv(g, 1, fl.spec[0], 0, out, 0, twz, 1.f, st));
  KW_S_HIP(hipStreamSynchronize(st));
  return KW_OK;
}
