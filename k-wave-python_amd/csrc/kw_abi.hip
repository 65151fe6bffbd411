// C ABI of the solver: plan lifetime, host-side table construction (FP64 ->
// FP32), step sequencing and hipGraph capture.  Declared in include/kwgpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/kwgpu.h"
#include "kw_fft.hpp"
#include "kw_host.hpp"
#include "kw_launch.hpp"
#include "kw_types.hpp"

using namespace kw;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define KW_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) return fail(KW_EHIP, std::string(#call ": ") + hipGetErrorString(_e)); \
  } while (0)

struct DevBuf {
  std::vector<void*> ptrs;
  ~DevBuf() { release(); }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
  template <class T>
  int alloc(T** out, size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(e == hipErrorOutOfMemory ? KW_ENOMEM : KW_EHIP,
                  std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    ptrs.push_back(p);
    *out = static_cast<T*>(p);
    return KW_OK;
  }
  template <class T>
  int upload(T** out, const T* host, size_t count) {
    int rc = alloc(out, count);
    if (rc) return rc;
    if (count) KW_HIP(hipMemcpy(*out, host, count * sizeof(T), hipMemcpyHostToDevice));
    return KW_OK;
  }
};

}  // namespace

namespace kw {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace kw

struct kw_plan {
  int dev = 0;
  hipStream_t st = nullptr;
  Geom g{};
  int nb = 0;
  int flags = 0;
  double alpha_power = 1.5;
  double spacing[3] = {0, 0, 0};
  bool use_sg = true;
  DevBuf mem;
  Fields fl{};
  Ops ops{};
  std::vector<Inst> h_inst;
  std::vector<char> inst_set;
  std::vector<float*> p0;
  std::vector<char> alpha_scalar;
  Inst* d_inst = nullptr;
  int nspecbuf = 0;       // allocated spectrum buffers
  bool chain_p = false;
  bool pattern_p = false;  // some instance has a single-row additive pressure source
  bool chain_u[3] = {false, false, false};
  bool rec_uc[3] = {false, false, false};  // some instance records the collocated velocity
  int max_sens = 0;
  std::vector<int> nrec, nsens;
  std::vector<unsigned> rec_flags;
  hipGraphExec_t g1 = nullptr, gk = nullptr;  // one step, graph_k steps
  hipGraphExec_t q1 = nullptr, qk = nullptr;  // the lean forms, once every filtered source has ended
  int graph_k = 8;
  int64_t t_host = -1;  // steps taken since kw_plan_init (-1: not known on the host)
  bool lean = true;      // use the lean step form once filtered sources have ended
  bool graphs = true;
  bool skip_gather = false;  // replayed steps of the reverse sweep do not record
  int col_tile[2] = {0, 0};  // tile variant of the FFT-only and the spectral-job column passes
  int col_pv[2] = {0, 0};    // radix-plan variant of those passes (lengths with a factor 5 or 7)
  bool tuned = false;
  bool col_wide = false;  // KW_COL_WIDE=1: no narrow-tile fallback on small grids
  int col_persist = 3;    // KW_COL_PERSIST: column passes that loop over tiles (ColArgs::persist)
  int col_grid = 0;       // KW_COL_GRID: workgroups per instance of looping passes (0: resident grid)
  bool inst_dirty = true;
  // per-kernel timing (kw_plan_run_profiled)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Rec { int kind; size_t e0, e1; double bytes; };
  std::vector<Rec> recs;

  ~kw_plan() {
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (hipGraphExec_t x : {g1, gk, q1, qk})
      if (x) (void)hipGraphExecDestroy(x);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

void drop_graphs(kw_plan* P) {
  for (hipGraphExec_t* x : {&P->g1, &P->gk, &P->q1, &P->qk})
    if (*x) { (void)hipGraphExecDestroy(*x); *x = nullptr; }
}

// First step at which no instance injects a source through a spectrum buffer
// (mode CHAIN) or a pressure source through a dense pattern (mode PATTERN);
// -1 when the plan has none.  From that step on, the step skips the source
// transforms and the register-staged source term of the mass update, as the
// per-step loop skips a source whose signal has ended (kspace_solver.py:96-98).
int64_t source_end(const kw_plan* P) {
  bool any = false;
  int64_t end = 0;
  for (const Inst& I : P->h_inst) {
    const Src* ss[4] = {&I.srcp, &I.srcu[0], &I.srcu[1], &I.srcu[2]};
    for (const Src* s : ss)
      if (s->mode == SRC_CHAIN || (s == &I.srcp && s->mode == SRC_PATTERN)) { any = true; end = std::max<int64_t>(end, s->len); }
  }
  return any ? end : -1;
}

int sync_inst(kw_plan* P) {
  if (!P->inst_dirty) return KW_OK;
  KW_HIP(hipMemcpy(P->d_inst, P->h_inst.data(), sizeof(Inst) * P->nb, hipMemcpyHostToDevice));
  P->inst_dirty = false;
  return KW_OK;
}

int ensure_spec(kw_plan* P, int idx) {
  while (P->nspecbuf <= idx) {
    float2* s = nullptr;
    int rc = P->mem.alloc(&s, (size_t)P->nb * P->g.nspec);
    if (rc) return rc;
    P->fl.spec[P->nspecbuf++] = s;
  }
  return KW_OK;
}

#define KW_LAUNCH(x)                                                              \
  do {                                                                            \
    hipError_t _e = (x);                                                          \
    if (_e != hipSuccess) return fail(KW_EHIP, std::string(#x ": ") + hipGetErrorString(_e)); \
  } while (0)

enum { KK_YPASS = 0, KK_XGRAD, KK_XDIV, KK_XFRAC, KK_MOMENTUM, KK_MASS, KK_EOS, KK_SRC, KK_GATHER, KK_TICK, KK_COUNT };
const char* const kKindNames[KK_COUNT] = {"col_y", "col_x_grad", "col_x_div", "col_x_frac", "z_momentum",
                                          "z_mass",  "z_eos",      "z_src",     "gather",     "tick"};

int ev_get(kw_plan* P, size_t* idx) {
  if (P->ev_used == P->ev_pool.size()) {
    hipEvent_t e;
    KW_HIP(hipEventCreate(&e));
    P->ev_pool.push_back(e);
  }
  *idx = P->ev_used++;
  return KW_OK;
}

// Launch `x`; in profiling mode bracket it with events on the plan stream and
// remember the algorithmic bytes the launch moves.
#define KW_TLAUNCH(kind, bytes, x)                                                 \
  do {                                                                             \
    size_t _i0 = 0, _i1 = 0;                                                       \
    if (P->prof) {                                                                 \
      int _r = ev_get(P, &_i0);                                                    \
      if (_r) return _r;                                                           \
      KW_HIP(hipEventRecord(P->ev_pool[_i0], P->st));                              \
    }                                                                              \
    KW_LAUNCH(x);                                                                  \
    if (P->prof) {                                                                 \
      int _r = ev_get(P, &_i1);                                                    \
      if (_r) return _r;                                                           \
      KW_HIP(hipEventRecord(P->ev_pool[_i1], P->st));                              \
      P->recs.push_back({(kind), _i0, _i1, (double)(bytes)});                      \
    }                                                                              \
  } while (0)

// y-pass over a list of buffers (3-D only)
int ypass(kw_plan* P, int mode, const std::vector<int>& bufs) {
  if (P->g.nd != 3 || bufs.empty()) return KW_OK;
  ColArgs a = col_args(P->g, 1, mode);
  a.tile = P->col_tile[0];
  a.pv = P->col_pv[0];
  a.wide = P->col_wide;
  a.persist = P->col_persist;
  a.grid = P->col_grid;
  a.nbuf = (int)bufs.size();
  for (size_t i = 0; i < bufs.size(); ++i) a.bufs[i] = bufs[i];
  const double spec_bytes = 8.0 * P->nb * (double)P->g.rows * P->g.nh;
  KW_TLAUNCH(KK_YPASS, 2.0 * spec_bytes * bufs.size(), launch_col(P->g, P->nb, P->fl, P->ops, P->d_inst, a, P->g.n[1], P->st));
  return KW_OK;
}

int xpass(kw_plan* P, const std::vector<XJob>& jobs, int kind) {
  ColArgs a = col_args(P->g, 0, 2);
  a.tile = P->col_tile[1];
  a.pv = P->col_pv[1];
  a.wide = P->col_wide;
  a.persist = P->col_persist;
  a.grid = P->col_grid;
  a.njob = (int)jobs.size();
  int nin = 0, last = -1;
  for (size_t i = 0; i < jobs.size(); ++i) {
    a.job[i] = jobs[i];
    if (jobs[i].in != last) { ++nin; last = jobs[i].in; }
  }
  const double spec_bytes = 8.0 * P->nb * (double)P->g.rows * P->g.nh;
  KW_TLAUNCH(kind, spec_bytes * (nin + jobs.size()), launch_col(P->g, P->nb, P->fl, P->ops, P->d_inst, a, P->g.n[0], P->st));
  return KW_OK;
}

// algorithmic bytes of the z-pass kernels (state arrays, medium arrays, spectra)
double zbytes_momentum(const kw_plan* P) {
  const double nr = (double)P->nb * P->g.ntot, ns = (double)P->nb * P->g.rows * P->g.nh;
  const double dtr = P->h_inst[0].dtr0_s ? 4.0 * nr : 0.0;
  return P->g.nd * (16.0 * ns + 8.0 * nr + dtr);
}
double zbytes_mass(const kw_plan* P) {
  const kw::Inst& I = P->h_inst[0];
  const double nr = (double)P->nb * P->g.ntot, ns = (double)P->nb * P->g.rows * P->g.nh;
  const int D = P->g.nd;
  double b = 8.0 * ns * D + 8.0 * nr * D + (I.rho0_s ? 4.0 * nr : 0.0);
  if (P->fl.rhot) b += 4.0 * nr;                      // write the density sum
  if (P->flags & KW_NONLINEAR) b += 4.0 * nr;         // read the previous sum
  if (P->flags & (KW_POWERLAW_ABS | KW_POWERLAW_DISP)) {
    b += 8.0 * ns * (((P->flags & KW_POWERLAW_ABS) ? 1 : 0) + ((P->flags & KW_POWERLAW_DISP) ? 1 : 0));
  } else {
    b += 4.0 * nr + 8.0 * ns + (I.c0sq_s ? 4.0 * nr : 0.0);
    if ((P->flags & KW_STOKES) && I.tau && I.tau_s) b += 4.0 * nr;
    if ((P->flags & KW_NONLINEAR) && I.bona_s) b += 4.0 * nr;
  }
  return b;
}
double zbytes_eos(const kw_plan* P) {
  const kw::Inst& I = P->h_inst[0];
  const double nr = (double)P->nb * P->g.ntot, ns = (double)P->nb * P->g.rows * P->g.nh;
  double b = 4.0 * nr + 4.0 * nr + 8.0 * ns;  // rho_total, p, P_z
  if (P->flags & KW_POWERLAW_ABS) b += 8.0 * ns + ((I.tau && I.tau_s) ? 4.0 * nr : 0.0);
  if (P->flags & KW_POWERLAW_DISP) b += 8.0 * ns + ((I.eta && I.eta_s) ? 4.0 * nr : 0.0);
  if (I.c0sq_s) b += 4.0 * nr;
  if (P->flags & KW_NONLINEAR) b += (I.bona_s ? 4.0 * nr : 0.0) + (I.rho0_s ? 4.0 * nr : 0.0);
  return b;
}

// One time step.  lean: every spectral (CHAIN) and pattern pressure source
// has ended, so no source spectrum is built or carried through the column
// passes and the mass update takes its form without a staged source term.
int enqueue_step(kw_plan* P, bool lean = false) {
  const Geom& g = P->g;
  const int D = g.nd;
  const float2* twz = P->ops.tw[D - 1];
  int rc;
  // --- momentum: grad p, u update, F_z[u_i] ---
  std::vector<int> ub;
  std::vector<XJob> jobs;
  for (int i = 0; i < D; ++i) jobs.push_back({0, i, XK_GRAD, i});
  for (int i = 0; i < D; ++i)
    if (P->chain_u[i] && !lean) {
      KW_TLAUNCH(KK_SRC, 0.0, launch_src_build(g, P->nb, P->fl, P->d_inst, twz, i, P->st));
      ub.push_back(D + 1 + i);
      jobs.push_back({D + 1 + i, D + 1 + i, XK_SRCKAPPA, 0});
    }
  std::vector<int> y0 = {0};
  y0.insert(y0.end(), ub.begin(), ub.end());
  if ((rc = ypass(P, 0, y0))) return rc;
  if ((rc = xpass(P, jobs, KK_XGRAD))) return rc;
  std::vector<int> y1;
  for (int i = 0; i < D; ++i) y1.push_back(i);
  y1.insert(y1.end(), ub.begin(), ub.end());
  if ((rc = ypass(P, 1, y1))) return rc;
  KW_TLAUNCH(KK_MOMENTUM, zbytes_momentum(P), launch_momentum(g, P->nb, P->fl, P->d_inst, twz, P->st));
  // --- mass: div u, rho update, EOS ---
  jobs.clear();
  std::vector<int> yb;
  for (int i = 0; i < D; ++i) {
    jobs.push_back({i, i, XK_DIV, i});
    yb.push_back(i);
    if (P->rec_uc[i]) { jobs.push_back({i, 2 * D + 1 + i, XK_UNSTAGGER, i}); yb.push_back(2 * D + 1 + i); }
  }
  const bool chain_p = P->chain_p && !lean;
  if (chain_p) {
    KW_TLAUNCH(KK_SRC, 0.0, launch_src_build(g, P->nb, P->fl, P->d_inst, twz, -1, P->st));
    jobs.push_back({D, D, XK_SRCKAPPA, 0});
    yb.push_back(D);
  }
  if ((rc = ypass(P, 0, yb))) return rc;
  if ((rc = xpass(P, jobs, KK_XDIV))) return rc;
  if ((rc = ypass(P, 1, yb))) return rc;
  for (int i = 0; i < D; ++i)
    if (P->rec_uc[i])
      KW_TLAUNCH(KK_SRC, 0.0, launch_zinv(g, P->nb, P->fl.spec[2 * D + 1 + i], g.nspec, P->fl.uc[i], g.ntot, twz, 1.f, P->st));
  KW_TLAUNCH(KK_MASS, zbytes_mass(P), launch_mass(g, P->nb, P->fl, P->d_inst, twz, P->flags, P->st, (P->chain_p || P->pattern_p) && !lean));
  if (P->flags & (KW_POWERLAW_ABS | KW_POWERLAW_DISP)) {
    jobs.clear();
    yb.clear();
    if (P->flags & KW_POWERLAW_ABS) { jobs.push_back({0, 0, XK_FRAC1, 0}); yb.push_back(0); }
    if (P->flags & KW_POWERLAW_DISP) { jobs.push_back({1, 1, XK_FRAC2, 0}); yb.push_back(1); }
    if ((rc = ypass(P, 0, yb))) return rc;
    if ((rc = xpass(P, jobs, KK_XFRAC))) return rc;
    if ((rc = ypass(P, 1, yb))) return rc;
    KW_TLAUNCH(KK_EOS, zbytes_eos(P), launch_eos(g, P->nb, P->fl, P->d_inst, twz, P->flags, P->st));
  }
  if (!P->skip_gather) KW_TLAUNCH(KK_GATHER, 0.0, launch_gather(g, P->nb, P->max_sens, P->fl, P->d_inst, P->st));
  KW_TLAUNCH(KK_TICK, 0.0, launch_tick(P->fl, P->d_inst, P->nb, P->st));
  return KW_OK;
}

int build_graphs(kw_plan* P) {
  const bool lean_too = P->lean && source_end(P) >= 0;
  for (int pass = 0; pass < (lean_too ? 4 : 2); ++pass) {
    const int nsteps = pass % 2 == 0 ? 1 : P->graph_k;
    KW_HIP(hipStreamBeginCapture(P->st, hipStreamCaptureModeThreadLocal));
    int rc = KW_OK;
    for (int s = 0; s < nsteps && rc == KW_OK; ++s) rc = enqueue_step(P, pass >= 2);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(P->st, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return fail(KW_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(KW_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    hipGraphExec_t* slot[4] = {&P->g1, &P->gk, &P->q1, &P->qk};
    *slot[pass] = ex;
  }
  return KW_OK;
}

// Enqueue nsteps steps: the lean form once the host knows every spectral
// source has ended, the full form otherwise (always correct: the kernels test
// the signal length themselves).
int launch_steps(kw_plan* P, int64_t nsteps) {
  int rc;
  const int64_t qend = P->lean ? source_end(P) : -1;
  if (P->graphs && !P->g1 && (rc = build_graphs(P))) return rc;
  const bool can_lean = qend >= 0 && P->t_host >= 0 && (!P->graphs || P->q1);
  auto lean_at = [&](int64_t s) { return can_lean && P->t_host + s >= qend; };
  for (int64_t s = 0; s < nsteps;) {
    const bool q = lean_at(s);
    if (!P->graphs) {
      if ((rc = enqueue_step(P, q))) return rc;
      ++s;
      continue;
    }
    // a graph_k block runs in one form: a full block that would reach the
    // lean steps is split into single steps up to them
    const bool block = s + P->graph_k <= nsteps && (q || !lean_at(s + P->graph_k - 1));
    KW_HIP(hipGraphLaunch(block ? (q ? P->qk : P->gk) : (q ? P->q1 : P->g1), P->st));
    s += block ? P->graph_k : 1;
  }
  if (P->t_host >= 0) P->t_host += nsteps;
  return KW_OK;
}

bool valid_field_n(const kw_plan* P, int64_t n) { return n == 1 || n == P->g.ntot; }

}  // namespace

extern "C" {

int kw_abi_version(void) { return KW_ABI_VERSION; }

const char* kw_last_error(void) { return g_err.c_str(); }

int kw_device_count(int* count) {
  if (!count) return fail(KW_EINVAL, "count is NULL");
  KW_HIP(hipGetDeviceCount(count));
  return KW_OK;
}

int kw_size_supported(int n) { return size_supported(n) ? 1 : 0; }

int kw_plan_create(const kw_plan_desc* d, kw_plan** out) {
  if (!d || !out) return fail(KW_EINVAL, "NULL argument");
  *out = nullptr;
  if (d->ndim != 2 && d->ndim != 3) return fail(KW_EUNSUPPORTED, "only 2-D and 3-D grids run on the GPU path");
  if (d->batch < 1) return fail(KW_EINVAL, "batch must be >= 1");
  for (int i = 0; i < d->ndim; ++i) {
    if (d->n[i] < 2) return fail(KW_EINVAL, "every axis needs at least 2 points");
    if (!size_supported(d->n[i]))
      return fail(KW_EUNSUPPORTED, "axis length " + std::to_string(d->n[i]) + " has no transform (lengths 2..4096 run)");
    if (!(d->spacing[i] > 0)) return fail(KW_EINVAL, "grid spacing must be positive");
  }
  int ndev = 0;
  KW_HIP(hipGetDeviceCount(&ndev));
  if (d->device < 0 || d->device >= ndev) return fail(KW_EINVAL, "device ordinal out of range");
  KW_HIP(hipSetDevice(d->device));
  auto* P = new kw_plan();
  P->dev = d->device;
  P->g = make_geom(d->ndim, d->n);
  P->nb = d->batch;
  P->flags = d->flags;
  P->alpha_power = d->alpha_power;
  P->use_sg = d->use_sg != 0;
  for (int i = 0; i < 3; ++i) P->spacing[i] = d->spacing[i];
  P->graphs = std::getenv("KW_NO_GRAPH") == nullptr;
  P->lean = std::getenv("KW_NO_LEAN") == nullptr;
  {
    const char* w = std::getenv("KW_COL_WIDE");
    P->col_wide = w && std::atoi(w) != 0;
    const char* pe = std::getenv("KW_COL_PERSIST");
    if (pe && pe[0] >= '0' && pe[0] <= '3') P->col_persist = pe[0] - '0';
    const char* ge = std::getenv("KW_COL_GRID");
    P->col_grid = ge ? std::max(0, std::atoi(ge)) : 0;
  }
  int rc = KW_OK;
  auto cleanup = [&](int code) { delete P; return code; };
  if (hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking) != hipSuccess) return cleanup(fail(KW_EHIP, "stream creation failed"));
  const size_t nf = (size_t)P->nb * P->g.ntot;
  if ((rc = P->mem.alloc(&P->fl.p, nf))) return cleanup(rc);
  for (int i = 0; i < d->ndim; ++i) {
    if ((rc = P->mem.alloc(&P->fl.u[i], nf))) return cleanup(rc);
    if ((rc = P->mem.alloc(&P->fl.rho[i], nf))) return cleanup(rc);
  }
  // sum of the split densities: EOS input (power law) and nonlinear factor of the next step
  if (P->flags & (KW_POWERLAW_ABS | KW_POWERLAW_DISP | KW_NONLINEAR))
    if ((rc = P->mem.alloc(&P->fl.rhot, nf))) return cleanup(rc);
  if ((rc = ensure_spec(P, d->ndim - 1))) return cleanup(rc);
  // generic-length family (kw_generic.hip): real temporaries for the unfused
  // contiguous-axis updates, a stash spectrum for column passes with jobs
  if (generic_z(P->g))
    for (int i = 0; i < 4; ++i)
      if ((rc = P->mem.alloc(&P->fl.tmp[i], nf))) return cleanup(rc);
  for (int ax = 0; ax < d->ndim - 1; ++ax)
    if (!compiled_size(d->n[ax]) && !P->fl.gstash)
      if ((rc = P->mem.alloc(&P->fl.gstash, (size_t)P->nb * P->g.nspec))) return cleanup(rc);
  if ((rc = P->mem.alloc(&P->fl.tdev, P->nb))) return cleanup(rc);
  if ((rc = P->mem.alloc(&P->d_inst, P->nb))) return cleanup(rc);
  // operator tables
  for (int ax = 0; ax < d->ndim; ++ax) {
    AxisTables t = build_axis(d->n[ax], d->spacing[ax], P->use_sg);
    float2 *gp, *dp, *up, *tp;
    float* kp;
    if ((rc = P->mem.upload(&gp, t.grad.data(), t.grad.size())) || (rc = P->mem.upload(&dp, t.div.data(), t.div.size())) ||
        (rc = P->mem.upload(&up, t.unst.data(), t.unst.size())) || (rc = P->mem.upload(&tp, t.tw.data(), t.tw.size())) ||
        (rc = P->mem.upload(&kp, t.k.data(), t.k.size())))
      return cleanup(rc);
    P->ops.grad[ax] = gp; P->ops.div[ax] = dp; P->ops.unst[ax] = up; P->ops.tw[ax] = tp; P->ops.k[ax] = kp;
  }
  P->ops.inv_ntot = (float)(1.0 / (double)P->g.ntot);
  P->ops.frac1 = (float)(d->alpha_power - 2.0);
  P->ops.frac2 = (float)(d->alpha_power - 1.0);
  P->h_inst.assign(P->nb, Inst{});
  P->inst_set.assign(P->nb, 0);
  P->p0.assign(P->nb, nullptr);
  P->alpha_scalar.assign(P->nb, 0);
  P->nrec.assign(P->nb, 0);
  P->nsens.assign(P->nb, 0);
  P->rec_flags.assign(P->nb, 0);
  *out = P;
  return KW_OK;
}

int kw_plan_set_instance(kw_plan* P, int b, const kw_inst_desc* d) {
  if (!P || !d || b < 0 || b >= P->nb) return fail(KW_EINVAL, "bad plan, instance or descriptor");
  if (d->nt < 1 || !(d->dt > 0) || !(d->c_ref > 0)) return fail(KW_EINVAL, "nt, dt and c_ref must be positive");
  KW_HIP(hipSetDevice(P->dev));
  Inst& I = P->h_inst[b];
  I.nt = d->nt;
  I.dt = (float)d->dt;
  I.kap = d->use_kspace ? (float)(d->c_ref * d->dt / 2.0) : 0.f;
  for (int ax = 0; ax < 3; ++ax) I.p0div[ax] = (float)P->g.nd;
  for (int ax = 0; ax < P->g.nd; ++ax) {
    if (!d->pml[ax] || !d->pml_sg[ax]) return fail(KW_EINVAL, "pml tables required for every axis");
    float *a, *s;
    int rc;
    if ((rc = P->mem.upload(&a, d->pml[ax], P->g.n[ax])) || (rc = P->mem.upload(&s, d->pml_sg[ax], P->g.n[ax]))) return rc;
    I.pml[ax] = a;
    I.pmlsg[ax] = s;
  }
  P->inst_set[b] = 1;
  P->inst_dirty = true;
  return sync_inst(P);
}

int kw_plan_set_density_split(kw_plan* P, int b, const double share[3]) {
  if (!P || !share || b < 0 || b >= P->nb) return fail(KW_EINVAL, "bad plan, instance or shares");
  if (!P->inst_set[b]) return fail(KW_EINVAL, "kw_plan_set_instance must precede kw_plan_set_density_split");
  double sum = 0.0;
  for (int ax = 0; ax < P->g.nd; ++ax) {
    if (!(share[ax] >= 0.0 && share[ax] <= 1.0)) return fail(KW_EINVAL, "density shares must lie in [0, 1]");
    sum += share[ax];
  }
  if (std::fabs(sum - 1.0) > 1e-9) return fail(KW_EINVAL, "density shares must sum to 1");
  KW_HIP(hipSetDevice(P->dev));
  Inst& I = P->h_inst[b];
  for (int ax = 0; ax < P->g.nd; ++ax) I.p0div[ax] = share[ax] > 0.0 ? (float)(1.0 / share[ax]) : INFINITY;
  P->inst_dirty = true;
  return sync_inst(P);
}

int kw_plan_set_field(kw_plan* P, int b, int which, const float* data, int64_t n) {
  if (!P || !data || b < 0 || b >= P->nb) return fail(KW_EINVAL, "bad plan, instance or data");
  if (!valid_field_n(P, n)) return fail(KW_EINVAL, "field must have 1 or grid-size elements");
  KW_HIP(hipSetDevice(P->dev));
  Inst& I = P->h_inst[b];
  float* dp;
  int rc = P->mem.upload(&dp, data, (size_t)n);
  if (rc) return rc;
  const int s = n == 1 ? 0 : 1;
  switch (which) {
    case KW_F_C0SQ: I.c0sq = dp; I.c0sq_s = s; break;
    case KW_F_RHO0: I.rho0 = dp; I.rho0_s = s; break;
    case KW_F_DTR0_X: case KW_F_DTR0_X + 1: case KW_F_DTR0_X + 2: {
      const int ax = which - KW_F_DTR0_X;
      if (ax >= P->g.nd) return fail(KW_EINVAL, "dt/rho0 axis out of range");
      if (I.dtr0_s != s && I.dtr0[0] && ax > 0) return fail(KW_EINVAL, "dt/rho0 arrays must all be scalar or all grid-sized");
      I.dtr0[ax] = dp; I.dtr0_s = s; break;
    }
    case KW_F_BONA: I.bona = dp; I.bona_s = s; break;
    case KW_F_TAU: I.tau = dp; I.tau_s = s; break;
    case KW_F_ETA: I.eta = dp; I.eta_s = s; break;
    case KW_F_ALPHA_NP: {
      if (n != 1) return fail(KW_EINVAL, "alpha (Np) is a scalar field");
      const double y = P->alpha_power, a_np = data[0];
      I.tau_c = (float)(-2.0 * a_np);
      I.eta_c = (float)(2.0 * a_np * std::tan(M_PI * y / 2.0));
      I.ym1h = (float)((y - 1.0) / 2.0);
      I.yh = (float)(y / 2.0);
      I.tau = nullptr; I.eta = nullptr;
      P->alpha_scalar[b] = 1;
      break;
    }
    case KW_F_P0:
      if (n != P->g.ntot) return fail(KW_EINVAL, "p0 must be grid-sized");
      P->p0[b] = dp; break;
    default: return fail(KW_EINVAL, "unknown field id");
  }
  P->inst_dirty = true;
  return sync_inst(P);
}

int kw_plan_set_source(kw_plan* P, int b, int which, int mode, int64_t npts, const int64_t* idx, const float* scale,
                       const int32_t* sigrow, const float* sig, int nsigrow, int len, const double fac[3]) {
  if (!P || b < 0 || b >= P->nb || which < -1 || which >= P->g.nd) return fail(KW_EINVAL, "bad plan, instance or source slot");
  if (!P->inst_set[b]) return fail(KW_EINVAL, "kw_plan_set_instance must precede kw_plan_set_source");
  if (npts < 1 || !idx || !scale || !sigrow || !sig || nsigrow < 1 || len < 1 || !fac)
    return fail(KW_EINVAL, "source needs points, scales, signal rows and a signal");
  if (mode != KW_SRC_DIRICHLET && mode != KW_SRC_ADDITIVE && mode != KW_SRC_ADDITIVE_NOCORR)
    return fail(KW_EINVAL, "unknown source mode");
  KW_HIP(hipSetDevice(P->dev));
  const Geom& g = P->g;
  for (int64_t i = 0; i < npts; ++i) {
    if (idx[i] < 0 || idx[i] >= g.ntot || (i && idx[i] <= idx[i - 1])) return fail(KW_EINVAL, "source indices must be ascending and inside the grid");
    if (sigrow[i] < 0 || sigrow[i] >= nsigrow) return fail(KW_EINVAL, "source signal row out of range");
  }
  Src s{};
  s.len = len;
  s.nsigrow = nsigrow;
  for (int i = 0; i < 3; ++i) s.fac[i] = (float)fac[i];
  int rc;
  float* dsig;
  if ((rc = P->mem.upload(&dsig, sig, (size_t)nsigrow * len))) return rc;
  s.sig = dsig;
  // CSR over grid rows
  std::vector<int> row_ptr(g.rows + 1, 0), col(npts);
  for (int64_t i = 0; i < npts; ++i) {
    row_ptr[idx[i] / g.nz + 1]++;
    col[i] = (int)(idx[i] % g.nz);
  }
  for (int r = 0; r < g.rows; ++r) row_ptr[r + 1] += row_ptr[r];
  int *drp, *dcol, *dsr;
  float* dsc;
  if ((rc = P->mem.upload(&drp, row_ptr.data(), row_ptr.size())) || (rc = P->mem.upload(&dcol, col.data(), col.size())) ||
      (rc = P->mem.upload(&dsc, scale, (size_t)npts)) || (rc = P->mem.upload(&dsr, sigrow, (size_t)npts)))
    return rc;
  s.row_ptr = drp; s.col = dcol; s.scale = dsc; s.sigrow = dsr;
  if (mode == KW_SRC_DIRICHLET) s.mode = SRC_DIRICHLET;
  else if (mode == KW_SRC_ADDITIVE_NOCORR) s.mode = SRC_NOCORR;
  else if (nsigrow == 1) {
    // single signal row: the filtered source is sig(t) * F^-1[kappa_s F[scale*mask]]
    s.mode = SRC_PATTERN;
    std::vector<float> dense((size_t)g.ntot, 0.f);
    for (int64_t i = 0; i < npts; ++i) dense[idx[i]] = scale[i];
    float* pat;
    if ((rc = P->mem.upload(&pat, dense.data(), dense.size()))) return rc;
    float2* tmp = nullptr;
    if (hipMalloc(&tmp, sizeof(float2) * g.nspec) != hipSuccess) return fail(KW_ENOMEM, "pattern scratch allocation failed");
    Fields f{};
    f.spec[0] = tmp;
    f.gstash = P->fl.gstash;  // generic-length x-pass: its forward-transform stash (one instance)
    const Inst* di = P->d_inst + b;
    const float2* twz = P->ops.tw[g.nd - 1];
    hipError_t e = launch_zfwd(g, 1, pat, 0, tmp, 0, twz, P->st);
    if (e == hipSuccess && g.nd == 3) {
      ColArgs a = col_args(g, 1, 0);
      a.nbuf = 1; a.bufs[0] = 0;
      e = launch_col(g, 1, f, P->ops, di, a, g.n[1], P->st);
    }
    if (e == hipSuccess) {
      ColArgs a = col_args(g, 0, 2);
      a.njob = 1; a.job[0] = {0, 0, XK_SRCKAPPA, 0};
      e = launch_col(g, 1, f, P->ops, di, a, g.n[0], P->st);
    }
    if (e == hipSuccess && g.nd == 3) {
      ColArgs a = col_args(g, 1, 1);
      a.nbuf = 1; a.bufs[0] = 0;
      e = launch_col(g, 1, f, P->ops, di, a, g.n[1], P->st);
    }
    if (e == hipSuccess) e = launch_zinv(g, 1, tmp, 0, pat, 0, twz, 1.f, P->st);
    if (e == hipSuccess) e = hipStreamSynchronize(P->st);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(KW_EHIP, std::string("source pattern: ") + hipGetErrorString(e));
    s.pattern = pat;
    if (which < 0 && !P->pattern_p) {
      P->pattern_p = true;
      drop_graphs(P);
    }
  } else {
    s.mode = SRC_CHAIN;
    s.chain_buf = which < 0 ? g.nd : g.nd + 1 + which;
    if ((rc = ensure_spec(P, s.chain_buf))) return rc;
    if (which < 0) P->chain_p = true; else P->chain_u[which] = true;
    drop_graphs(P);
  }
  if (which < 0) P->h_inst[b].srcp = s; else P->h_inst[b].srcu[which] = s;
  P->inst_dirty = true;
  return sync_inst(P);
}

int kw_plan_set_sensor(kw_plan* P, int b, int nsens, int ncorner, const int64_t* idx, const float* w, int rec_start, unsigned rec_flags) {
  if (!P || b < 0 || b >= P->nb) return fail(KW_EINVAL, "bad plan or instance");
  if (!P->inst_set[b]) return fail(KW_EINVAL, "kw_plan_set_instance must precede kw_plan_set_sensor");
  Inst& I = P->h_inst[b];
  if (nsens < 1 || ncorner < 1 || !idx) return fail(KW_EINVAL, "sensor needs points");
  if (rec_start < 0 || rec_start >= I.nt) return fail(KW_EINVAL, "record start outside [0, Nt)");
  for (int64_t i = 0; i < (int64_t)nsens * ncorner; ++i)
    if (idx[i] < 0 || idx[i] >= P->g.ntot) return fail(KW_EINVAL, "sensor index outside the grid");
  KW_HIP(hipSetDevice(P->dev));
  int rc;
  int64_t* di;
  if ((rc = P->mem.upload(&di, idx, (size_t)nsens * ncorner))) return rc;
  I.sidx = di;
  I.sens_kind = 1;
  if (w) {
    float* dw;
    if ((rc = P->mem.upload(&dw, w, (size_t)nsens * ncorner))) return rc;
    I.sw = dw;
    I.sens_kind = 2;
  }
  I.nsens = nsens;
  I.ncorner = ncorner;
  I.rec_start = rec_start;
  const int nrec = I.nt - rec_start;
  I.rec_p = nullptr; I.agg_p = nullptr;
  for (int i = 0; i < 3; ++i) I.rec_u[i] = nullptr;
  if (rec_flags & KW_REC_P)
    if ((rc = P->mem.alloc(&I.rec_p, (size_t)nrec * nsens))) return rc;
  if (rec_flags & KW_REC_P_AGG)
    if ((rc = P->mem.alloc(&I.agg_p, (size_t)3 * nsens))) return rc;
  for (int i = 0; i < P->g.nd; ++i) {
    I.rec_uc[i] = nullptr;
    if (rec_flags & (KW_REC_UX << i))
      if ((rc = P->mem.alloc(&I.rec_u[i], (size_t)nrec * nsens))) return rc;
    if (rec_flags & (KW_REC_UCX << i)) {
      if ((rc = P->mem.alloc(&I.rec_uc[i], (size_t)nrec * nsens))) return rc;
      if (!P->rec_uc[i]) {
        if ((rc = ensure_spec(P, 2 * P->g.nd + 1 + i))) return rc;
        if ((rc = P->mem.alloc(&P->fl.uc[i], (size_t)P->nb * P->g.ntot))) return rc;
        P->rec_uc[i] = true;
        drop_graphs(P);
      }
    }
  }
  P->nrec[b] = nrec;
  P->nsens[b] = nsens;
  P->rec_flags[b] = rec_flags;
  if (nsens > P->max_sens) {
    P->max_sens = nsens;
    drop_graphs(P);
  }
  P->inst_dirty = true;
  return sync_inst(P);
}

namespace {

// Pick, per plan, between the two compiled tile widths of the FFT-only column
// pass and of the pass with spectral jobs, by timing each on the plan's own
// (zeroed) spectrum buffers: a forward transform of the D working spectra along y
// (along x in 2-D), and the gradient plus the divergence pass.  The alternative
// replaces the default only when at least 3 % faster (the timing of a few passes
// is noisy, and the default widths were measured best at 512).  Runs once, from
// kw_plan_init before it clears the spectra; KW_AUTOTUNE=0 keeps the defaults and
// KW_COL_TILES="ab" forces the variants (a: FFT-only pass, b: job pass).
int autotune_tiles(kw_plan* P) {
  if (P->tuned) return KW_OK;
  P->tuned = true;
  const char* env = std::getenv("KW_AUTOTUNE");
  if (env && std::atoi(env) == 0) return KW_OK;
  const char* force = std::getenv("KW_COL_TILES");  // two digits, e.g. "10": FFT-only variant, job-pass variant
  // KW_COL_PLANS: two digits force the radix-plan variants (same order); "tune" adds them to
  // the timed candidates.  By default plan 0 is kept: the plans agree only to FP32 rounding,
  // and a choice by timing could then differ between a batch and its instances run alone.
  const char* fplan = std::getenv("KW_COL_PLANS");
  const bool tune_pv = fplan && std::strcmp(fplan, "tune") == 0;
  const bool ftile = force && std::strlen(force) == 2, fpv = !tune_pv && fplan && std::strlen(fplan) == 2;
  if (ftile) {
    P->col_tile[0] = force[0] == '1';
    P->col_tile[1] = force[1] == '1';
  }
  if (fpv) {
    P->col_pv[0] = fplan[0] == '1';
    P->col_pv[1] = fplan[1] == '1';
  }
  if (ftile && fpv) return KW_OK;
  const Geom& g = P->g;
  const int D = g.nd;
  for (int s = 0; s < D; ++s) KW_LAUNCH(launch_fill(reinterpret_cast<float*>(P->fl.spec[s]), 2 * (int64_t)P->nb * g.nspec, 0.f, P->st));
  int* tdev = P->fl.tdev;
  P->fl.tdev = nullptr;  // time every instance
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = KW_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(KW_EHIP, "autotune: event creation failed");
  std::vector<int> bufs;
  std::vector<XJob> jobs, djobs;
  for (int i = 0; i < D; ++i) {
    bufs.push_back(i);
    jobs.push_back({0, i, XK_GRAD, i});
    djobs.push_back({i, i, XK_DIV, i});
  }
  // candidates per pass kind: both tile widths (and, with KW_COL_PLANS=tune, both radix plans
  // where the length has two); the default (width 0, plan 0) is replaced only by a candidate
  // >= 3 % faster
  for (int kind = 0; kind < 2 && rc == KW_OK; ++kind) {
    const int nline = (kind == 1 || D == 2) ? g.n[0] : g.n[1];
    const int npv = tune_pv && compiled_size(nline) && has_plan_variant(nline) ? 2 : 1;
    // a forced choice is the only candidate of its kind
    const int pv0 = fpv ? P->col_pv[kind] : 0, pv1 = fpv ? P->col_pv[kind] + 1 : npv;
    const int tl0 = ftile ? P->col_tile[kind] : 0, tl1 = ftile ? P->col_tile[kind] + 1 : 2;
    float best_ms = -1.f;
    int best_tile = tl0, best_pv = pv0;
    for (int pv = pv0; pv < pv1 && rc == KW_OK; ++pv) {
      for (int tile = tl0; tile < tl1 && rc == KW_OK; ++tile) {
        P->col_tile[kind] = tile;
        P->col_pv[kind] = pv;
        auto pass = [&]() -> int {
          if (kind == 1) {
            const int r = xpass(P, jobs, KK_XGRAD);
            return r ? r : xpass(P, djobs, KK_XDIV);
          }
          if (D == 3) return ypass(P, 0, bufs);
          ColArgs a = col_args(g, 0, 0);
          a.tile = tile;
          a.pv = pv;
          a.persist = P->col_persist;
          a.grid = P->col_grid;
          a.nbuf = D;
          for (int i = 0; i < D; ++i) a.bufs[i] = i;
          KW_LAUNCH(launch_col(g, P->nb, P->fl, P->ops, P->d_inst, a, g.n[0], P->st));
          return KW_OK;
        };
        if ((rc = pass())) break;  // warm-up
        if (hipEventRecord(e0, P->st) != hipSuccess) { rc = fail(KW_EHIP, "autotune: event record failed"); break; }
        for (int r = 0; r < 3 && rc == KW_OK; ++r) rc = pass();
        if (rc) break;
        float ms = 0.f;
        if (hipEventRecord(e1, P->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = fail(KW_EHIP, "autotune: timing failed"); break; }
        if (best_ms < 0.f || ms < 0.97f * best_ms) {
          best_ms = ms;
          best_tile = tile;
          best_pv = pv;
        }
      }
    }
    P->col_tile[kind] = rc == KW_OK ? best_tile : tl0;
    P->col_pv[kind] = rc == KW_OK ? best_pv : pv0;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  P->fl.tdev = tdev;
  return rc;
}

}  // namespace

int kw_plan_tiles(kw_plan* P, int* fft_tile, int* job_tile) {
  if (!P || !fft_tile || !job_tile) return fail(KW_EINVAL, "NULL argument");
  *fft_tile = P->col_tile[0];
  *job_tile = P->col_tile[1];
  return KW_OK;
}

int kw_plan_radix_plans(kw_plan* P, int* fft_plan, int* job_plan) {
  if (!P || !fft_plan || !job_plan) return fail(KW_EINVAL, "NULL argument");
  *fft_plan = P->col_pv[0];
  *job_plan = P->col_pv[1];
  return KW_OK;
}

int kw_plan_init(kw_plan* P) {
  if (!P) return fail(KW_EINVAL, "NULL plan");
  KW_HIP(hipSetDevice(P->dev));
  P->t_host = -1;
  const Geom& g = P->g;
  int with_p0 = 0;
  for (int b = 0; b < P->nb; ++b) {
    const Inst& I = P->h_inst[b];
    if (!P->inst_set[b]) return fail(KW_EINVAL, "instance " + std::to_string(b) + " not set");
    if (!I.c0sq || !I.rho0) return fail(KW_EINVAL, "c0^2 and rho0 must be set");
    for (int ax = 0; ax < g.nd; ++ax)
      if (!I.dtr0[ax]) return fail(KW_EINVAL, "dt/rho0_sg must be set for every axis");
    if ((P->flags & KW_NONLINEAR) && !I.bona) return fail(KW_EINVAL, "B/A must be set for a nonlinear plan");
    if ((P->flags & (KW_STOKES | KW_POWERLAW_ABS)) && !I.tau && !P->alpha_scalar[b]) return fail(KW_EINVAL, "tau or alpha must be set for an absorbing plan");
    if ((P->flags & KW_POWERLAW_DISP) && !I.eta && !P->alpha_scalar[b]) return fail(KW_EINVAL, "eta or alpha must be set for a dispersive plan");
    with_p0 += P->p0[b] != nullptr;
  }
  if (with_p0 != 0 && with_p0 != P->nb) return fail(KW_EINVAL, "either every instance of a plan has p0 or none has");
  int rc;
  if ((rc = sync_inst(P))) return rc;
  if ((rc = autotune_tiles(P))) return rc;
  const int64_t nf = (int64_t)P->nb * g.ntot;
  KW_LAUNCH(launch_fill(P->fl.p, nf, 0.f, P->st));
  for (int i = 0; i < g.nd; ++i) {
    KW_LAUNCH(launch_fill(P->fl.u[i], nf, 0.f, P->st));
    KW_LAUNCH(launch_fill(P->fl.rho[i], nf, 0.f, P->st));
  }
  if (P->fl.rhot) KW_LAUNCH(launch_fill(P->fl.rhot, nf, 0.f, P->st));
  for (int s = 0; s < P->nspecbuf; ++s) KW_LAUNCH(launch_fill(reinterpret_cast<float*>(P->fl.spec[s]), 2 * (int64_t)P->nb * g.nspec, 0.f, P->st));
  KW_LAUNCH(launch_setint(P->fl.tdev, P->nb, 0, P->st));
  if (with_p0) {
    for (int b = 0; b < P->nb; ++b)
      KW_HIP(hipMemcpyAsync(P->fl.p + (int64_t)b * g.ntot, P->p0[b], sizeof(float) * g.ntot, hipMemcpyDeviceToDevice, P->st));
    const float2* twz = P->ops.tw[g.nd - 1];
    KW_LAUNCH(launch_zfwd(g, P->nb, P->fl.p, g.ntot, P->fl.spec[0], g.nspec, twz, P->st));
    if ((rc = ypass(P, 0, {0}))) return rc;
    std::vector<XJob> jobs;
    std::vector<int> yb;
    for (int i = 0; i < g.nd; ++i) { jobs.push_back({0, i, XK_GRAD, i}); yb.push_back(i); }
    if ((rc = xpass(P, jobs, KK_XGRAD))) return rc;
    if ((rc = ypass(P, 1, yb))) return rc;
    KW_LAUNCH(launch_init_u(g, P->nb, P->fl, P->d_inst, twz, P->st));
    // velocity on the pressure grid for the t = 0 record (kspace_solver.py:741-743)
    std::vector<XJob> uj;
    std::vector<int> us;
    for (int i = 0; i < g.nd; ++i)
      if (P->rec_uc[i]) {
        const int s = 2 * g.nd + 1 + i;
        KW_LAUNCH(launch_zfwd(g, P->nb, P->fl.u[i], g.ntot, P->fl.spec[s], g.nspec, twz, P->st));
        uj.push_back({s, s, XK_UNSTAGGER, i});
        us.push_back(s);
      }
    if (!uj.empty()) {
      if ((rc = ypass(P, 0, us))) return rc;
      if ((rc = xpass(P, uj, KK_XDIV))) return rc;
      if ((rc = ypass(P, 1, us))) return rc;
      for (int i = 0; i < g.nd; ++i)
        if (P->rec_uc[i])
          KW_LAUNCH(launch_zinv(g, P->nb, P->fl.spec[2 * g.nd + 1 + i], g.nspec, P->fl.uc[i], g.ntot, twz, 1.f, P->st));
    }
    KW_LAUNCH(launch_zfwd(g, P->nb, P->fl.p, g.ntot, P->fl.spec[0], g.nspec, twz, P->st));
    KW_LAUNCH(launch_gather(g, P->nb, P->max_sens, P->fl, P->d_inst, P->st));
    KW_LAUNCH(launch_tick(P->fl, P->d_inst, P->nb, P->st));
  }
  KW_HIP(hipStreamSynchronize(P->st));
  int t0 = 0;  // every instance starts at the same step
  KW_HIP(hipMemcpy(&t0, P->fl.tdev, sizeof(int), hipMemcpyDeviceToHost));
  P->t_host = t0;
  return KW_OK;
}

int kw_plan_launch(kw_plan* P, int64_t nsteps) {
  if (!P || nsteps < 0) return fail(KW_EINVAL, "bad plan or step count");
  KW_HIP(hipSetDevice(P->dev));
  int rc;
  if ((rc = sync_inst(P))) return rc;
  if ((rc = launch_steps(P, nsteps))) return rc;
  return KW_OK;
}

int kw_plan_sync(kw_plan* P) {
  if (!P) return fail(KW_EINVAL, "NULL plan");
  KW_HIP(hipSetDevice(P->dev));
  KW_HIP(hipStreamSynchronize(P->st));
  return KW_OK;
}

int kw_plan_run(kw_plan* P, int64_t nsteps) {
  if (!P || nsteps < 0) return fail(KW_EINVAL, "bad plan or step count");
  KW_HIP(hipSetDevice(P->dev));
  int rc;
  if ((rc = sync_inst(P))) return rc;
  if ((rc = launch_steps(P, nsteps))) return rc;
  KW_HIP(hipStreamSynchronize(P->st));
  return KW_OK;
}

int kw_plan_run_profiled(kw_plan* P, int64_t nsteps, double* kernel_ms, double* kernel_bytes, int64_t* launches) {
  if (!P || nsteps < 0 || !kernel_ms || !kernel_bytes || !launches) return fail(KW_EINVAL, "bad arguments");
  KW_HIP(hipSetDevice(P->dev));
  int rc;
  if ((rc = sync_inst(P))) return rc;
  for (int k = 0; k < KK_COUNT; ++k) { kernel_ms[k] = 0.0; kernel_bytes[k] = 0.0; launches[k] = 0; }
  const int64_t qend = P->lean ? source_end(P) : -1;
  // Steps run in chunks: after each chunk the stream is drained, the event pairs
  // are read and the pool is reused, so the run holds at most one chunk's events
  // (a few hundred) whatever its length.
  constexpr int64_t kChunk = 8;
  for (int64_t s0 = 0; s0 < nsteps && rc == KW_OK; s0 += kChunk) {
    P->prof = true;
    P->ev_used = 0;
    P->recs.clear();
    const int64_t s1 = std::min(nsteps, s0 + kChunk);
    for (int64_t s = s0; s < s1 && rc == KW_OK; ++s) rc = enqueue_step(P, qend >= 0 && P->t_host >= 0 && P->t_host + s >= qend);
    P->prof = false;
    if (rc) break;
    KW_HIP(hipStreamSynchronize(P->st));
    for (const auto& r : P->recs) {
      float ms = 0.f;
      KW_HIP(hipEventElapsedTime(&ms, P->ev_pool[r.e0], P->ev_pool[r.e1]));
      kernel_ms[r.kind] += ms;
      kernel_bytes[r.kind] += r.bytes;
      launches[r.kind] += 1;
    }
  }
  P->prof = false;
  P->recs.clear();
  if (P->t_host >= 0) P->t_host += nsteps;
  return rc;
}

int kw_kernel_kind_count(void) { return KK_COUNT; }

const char* kw_kernel_kind_name(int kind) { return (kind >= 0 && kind < KK_COUNT) ? kKindNames[kind] : ""; }

int kw_plan_field_size(kw_plan* P, int b, int which, int64_t* n) {
  if (!P || !n || b < 0 || b >= P->nb) return fail(KW_EINVAL, "bad plan or instance");
  const Geom& g = P->g;
  if (which == KW_F_P || (which >= KW_F_UX && which < KW_F_UX + g.nd) || (which >= KW_F_RHOX && which < KW_F_RHOX + g.nd)) *n = g.ntot;
  else if (which == KW_F_REC_P || (which >= KW_F_REC_UX && which < KW_F_REC_UX + g.nd) ||
           (which >= KW_F_REC_UCX && which < KW_F_REC_UCX + g.nd)) *n = (int64_t)P->nrec[b] * P->nsens[b];
  else if (which == KW_F_AGG_P) *n = 3 * (int64_t)P->nsens[b];
  else if (which == KW_F_STEP) *n = 1;
  else return fail(KW_EINVAL, "field cannot be read");
  return KW_OK;
}

int kw_plan_read(kw_plan* P, int b, int which, float* out, int64_t n) {
  int64_t need = 0;
  int rc = kw_plan_field_size(P, b, which, &need);
  if (rc) return rc;
  if (!out || n != need) return fail(KW_EINVAL, "output buffer size mismatch: need " + std::to_string(need));
  KW_HIP(hipSetDevice(P->dev));
  KW_HIP(hipStreamSynchronize(P->st));
  const Geom& g = P->g;
  const Inst& I = P->h_inst[b];
  const float* src = nullptr;
  const int64_t off = (int64_t)b * g.ntot;
  if (which == KW_F_P) src = P->fl.p + off;
  else if (which >= KW_F_UX && which < KW_F_UX + 3) src = P->fl.u[which - KW_F_UX] + off;
  else if (which >= KW_F_RHOX && which < KW_F_RHOX + 3) src = P->fl.rho[which - KW_F_RHOX] + off;
  else if (which == KW_F_REC_P) src = I.rec_p;
  else if (which >= KW_F_REC_UX && which < KW_F_REC_UX + 3) src = I.rec_u[which - KW_F_REC_UX];
  else if (which >= KW_F_REC_UCX && which < KW_F_REC_UCX + 3) src = I.rec_uc[which - KW_F_REC_UCX];
  else if (which == KW_F_AGG_P) src = I.agg_p;
  else if (which == KW_F_STEP) {
    int t = 0;
    KW_HIP(hipMemcpy(&t, P->fl.tdev + b, sizeof(int), hipMemcpyDeviceToHost));
    out[0] = (float)t;
    return KW_OK;
  }
  if (!src) return fail(KW_EINVAL, "field was not recorded");
  KW_HIP(hipMemcpy(out, src, sizeof(float) * n, hipMemcpyDeviceToHost));
  return KW_OK;
}

// ---------------------------------------------------------------------------
// gradient of a sensor-trace functional with respect to c0 (kw_adjoint.hpp)
// ---------------------------------------------------------------------------
namespace {

int adjoint_step(kw_plan* P, const AdjFields& af, float* Zbuf, int t) {
  const Geom& g = P->g;
  const int D = g.nd;
  const float2* twz = P->ops.tw[D - 1];
  int rc;
  KW_LAUNCH(launch_adj_scatter(g, P->nb, P->max_sens, af, P->d_inst, t, P->st));
  if (af.absorb == 2) {  // power law: tau w and eta w through the fractional column passes
    KW_LAUNCH(launch_adj_abs1(g, P->nb, P->fl, af, P->d_inst, twz, P->st));
    std::vector<XJob> fj = {{0, 0, XK_FRAC1, 0}};
    std::vector<int> fb = {0};
    if (af.disp) { fj.push_back({1, 1, XK_FRAC2, 0}); fb.push_back(1); }
    if ((rc = ypass(P, 0, fb))) return rc;
    if ((rc = xpass(P, fj, KK_XFRAC))) return rc;
    if ((rc = ypass(P, 1, fb))) return rc;
  }
  KW_LAUNCH(launch_adj_eos(g, P->nb, P->fl, af, P->d_inst, twz, t, P->st));
  bool src_on = false, filt = false;
  for (int b = 0; b < P->nb; ++b) {
    const kw::Src& s = P->h_inst[b].srcp;
    if (s.mode != SRC_NONE && t < s.len) {
      src_on = true;
      filt = filt || s.mode == SRC_PATTERN || s.mode == SRC_CHAIN;
    }
  }
  if (src_on) {
    if (filt) {  // Z = F^-1[kappa_s F[R]] (K is symmetric)
      KW_LAUNCH(launch_zfwd(g, P->nb, af.R, g.ntot, P->fl.spec[D], g.nspec, twz, P->st));
      if ((rc = ypass(P, 0, {D}))) return rc;
      if ((rc = xpass(P, {{D, D, XK_SRCKAPPA, 0}}, KK_XDIV))) return rc;
      if ((rc = ypass(P, 1, {D}))) return rc;
      KW_LAUNCH(launch_zinv(g, P->nb, P->fl.spec[D], g.nspec, Zbuf, g.ntot, twz, 1.f, P->st));
    }
    KW_LAUNCH(launch_adj_srcgrad(g, P->nb, af, -1, af.R, Zbuf, P->d_inst, t, P->st));
  }
  std::vector<int> bufs;
  std::vector<XJob> jobs;
  for (int i = 0; i < D; ++i) { bufs.push_back(i); jobs.push_back({i, i, XK_GRAD, i}); }
  if ((rc = ypass(P, 0, bufs))) return rc;
  if ((rc = xpass(P, jobs, KK_XGRAD))) return rc;
  if ((rc = ypass(P, 1, bufs))) return rc;
  KW_LAUNCH(launch_adj_mom(g, P->nb, P->fl, af, P->d_inst, twz, t, P->st));
  for (int i = 0; i < D; ++i) jobs[i].kind = XK_DIV;
  if ((rc = ypass(P, 0, bufs))) return rc;
  if ((rc = xpass(P, jobs, KK_XDIV))) return rc;
  if ((rc = ypass(P, 1, bufs))) return rc;
  KW_LAUNCH(launch_adj_sum(g, P->nb, P->fl, af, P->d_inst, twz, P->st));
  // velocity sources: the speed term of each axis's additive source, from dJ/du_i^t that
  // k_adj_mom kept in U_i (spectrum buffer D is free again once the divergence passes ran)
  for (int ax = 0; ax < D; ++ax) {
    if (!af.U[ax]) continue;
    bool on = false, ufilt = false;
    for (int b = 0; b < P->nb; ++b) {
      const kw::Src& s = P->h_inst[b].srcu[ax];
      if (s.mode != SRC_NONE && s.mode != SRC_DIRICHLET && t < s.len) {
        on = true;
        ufilt = ufilt || s.mode == SRC_PATTERN || s.mode == SRC_CHAIN;
      }
    }
    if (!on) continue;
    if (ufilt) {  // Z = F^-1[kappa_s F[U_i]]
      KW_LAUNCH(launch_zfwd(g, P->nb, af.U[ax], g.ntot, P->fl.spec[D], g.nspec, twz, P->st));
      if ((rc = ypass(P, 0, {D}))) return rc;
      if ((rc = xpass(P, {{D, D, XK_SRCKAPPA, 0}}, KK_XDIV))) return rc;
      if ((rc = ypass(P, 1, {D}))) return rc;
      KW_LAUNCH(launch_zinv(g, P->nb, P->fl.spec[D], g.nspec, Zbuf, g.ntot, twz, 1.f, P->st));
    }
    KW_LAUNCH(launch_adj_srcgrad(g, P->nb, af, ax, af.U[ax], Zbuf, P->d_inst, t, P->st));
  }
  return KW_OK;
}

}  // namespace

int kw_plan_adjoint_c(kw_plan* P, const float* adj, int ckpt, float* grad_c) {
  if (!P || !adj || !grad_c || ckpt < 1) return fail(KW_EINVAL, "NULL argument or checkpoint interval < 1");
  const Geom& g = P->g;
  const int D = g.nd;
  const int nb = P->nb;
  const kw::Inst& I0 = P->h_inst[0];
  for (int b = 0; b < nb; ++b) {
    if (!P->inst_set[b]) return fail(KW_EINVAL, "instance " + std::to_string(b) + " not set");
    const kw::Inst& I = P->h_inst[b];
    if (!I.rec_p) return fail(KW_EINVAL, "adjoint: every instance must record p");
    if (I.nt != I0.nt || I.nsens != I0.nsens || I.rec_start != I0.rec_start)
      return fail(KW_EINVAL, "adjoint: instances must share Nt, sensor count and record start");
  }
  KW_HIP(hipSetDevice(P->dev));
  int rc;
  if ((rc = ensure_spec(P, D))) return rc;
  if ((rc = sync_inst(P))) return rc;
  const int nt = I0.nt;
  const int nrec = nt - I0.rec_start;
  // with an initial pressure the state after step 0 is set by kw_plan_init (p = p0,
  // rho_i = p0 / (c0^2 p0div_i), u_i from grad p0), so the stepped range starts at 1
  const bool with_p0 = P->p0[0] != nullptr;
  const int tb = with_p0 ? 1 : 0;
  const int nck = (nt - tb + ckpt - 1) / ckpt;
  const int64_t nf = (int64_t)nb * g.ntot;
  // B/A: the density sum of the step before enters the mass update, so it is state too
  const bool nonlin = (P->flags & KW_NONLINEAR) != 0;
  const int nstate = 2 * D + 1 + (nonlin ? 1 : 0);  // p, u_i, rho_i (, sum of rho_i)
  DevBuf tmp;                    // freed when the call returns
  float *ck, *pbuf, *Zbuf, *dadj;
  AdjFields af{};
  if ((rc = tmp.alloc(&ck, (size_t)nck * nstate * nf)) || (rc = tmp.alloc(&pbuf, (size_t)std::min(ckpt, nt) * nf)) ||
      (rc = tmp.alloc(&af.pb, (size_t)nf)) || (rc = tmp.alloc(&af.cb, (size_t)nf)) || (rc = tmp.alloc(&af.R, (size_t)nf)) ||
      (rc = tmp.alloc(&Zbuf, (size_t)nf)) || (rc = tmp.upload(&dadj, adj, (size_t)nb * nrec * I0.nsens)))
    return rc;
  for (int i = 0; i < D; ++i)
    if ((rc = tmp.alloc(&af.ub[i], (size_t)nf)) || (rc = tmp.alloc(&af.rb[i], (size_t)nf))) return rc;
  // axes with an additive velocity source keep dJ/du_i of the step being reversed
  for (int i = 0; i < D; ++i) {
    bool add = false;
    for (int b = 0; b < nb; ++b) add = add || (P->h_inst[b].srcu[i].mode != SRC_NONE && P->h_inst[b].srcu[i].mode != SRC_DIRICHLET);
    if (add && (rc = tmp.alloc(&af.U[i], (size_t)nf))) return rc;
  }
  // absorbing media: the replay keeps each step's absorption / dispersion operands
  af.absorb = (P->flags & KW_STOKES) ? 1 : (P->flags & KW_POWERLAW_ABS) ? 2 : 0;
  if ((P->flags & KW_POWERLAW_DISP) && !(P->flags & KW_POWERLAW_ABS))
    return fail(KW_EUNSUPPORTED, "adjoint: dispersion without absorption is not differentiated");
  af.disp = (P->flags & KW_POWERLAW_DISP) ? 1 : 0;
  af.ym1 = (float)(P->alpha_power - 1.0);
  af.y = (float)P->alpha_power;
  float *abuf = nullptr, *bbuf = nullptr;
  if (af.absorb && (rc = tmp.alloc(&abuf, (size_t)std::min(ckpt, nt) * nf))) return rc;
  if (af.disp && (rc = tmp.alloc(&bbuf, (size_t)std::min(ckpt, nt) * nf))) return rc;
  // generic-length contiguous axis with the power law: L1[tau w] and L2[eta w] on the grid
  if (generic_z(g) && af.absorb == 2 && ((rc = tmp.alloc(&af.gl[0], (size_t)nf)) || (rc = tmp.alloc(&af.gl[1], (size_t)nf)))) return rc;
  // B/A: the replay keeps the density sum before and after every step (ckpt + 1 slots)
  // and each step's per-axis Dop_i u_i
  float *sbuf = nullptr, *dbuf = nullptr;
  af.nonlin = nonlin ? 1 : 0;
  if (nonlin && ((rc = tmp.alloc(&sbuf, (size_t)(std::min(ckpt, nt) + 1) * nf)) ||
                 (rc = tmp.alloc(&dbuf, (size_t)std::min(ckpt, nt) * D * nf))))
    return rc;
  af.adj = dadj;
  af.nrec = nrec;
  af.nsens = I0.nsens;
  std::vector<float*> state = {P->fl.p};
  for (int i = 0; i < D; ++i) state.push_back(P->fl.u[i]);
  for (int i = 0; i < D; ++i) state.push_back(P->fl.rho[i]);
  if (nonlin) state.push_back(P->fl.rhot);
  auto copy_state = [&](int k, bool save) -> int {
    for (int s = 0; s < nstate; ++s) {
      float* slot = ck + ((int64_t)k * nstate + s) * nf;
      KW_HIP(hipMemcpyAsync(save ? slot : state[s], save ? state[s] : slot, sizeof(float) * nf, hipMemcpyDeviceToDevice, P->st));
    }
    return KW_OK;
  };
  // forward pass: records the traces, keeps the state before every ckpt-th step
  if ((rc = kw_plan_init(P))) return rc;
  P->t_host = -1;  // the sweep below sets the step counter itself
  P->skip_gather = false;
  for (int t = tb; t < nt; ++t) {
    if ((t - tb) % ckpt == 0 && (rc = copy_state((t - tb) / ckpt, true))) return rc;
    if ((rc = enqueue_step(P))) return rc;
  }
  // reverse sweep, one checkpoint segment at a time
  for (float* f : {af.pb, af.cb, af.R}) KW_LAUNCH(launch_fill(f, nf, 0.f, P->st));
  for (int i = 0; i < D; ++i) {
    KW_LAUNCH(launch_fill(af.ub[i], nf, 0.f, P->st));
    KW_LAUNCH(launch_fill(af.rb[i], nf, 0.f, P->st));
  }
  int* tdev = P->fl.tdev;
  P->skip_gather = true;
  for (int seg = nck - 1; seg >= 0 && rc == KW_OK; --seg) {
    const int t0 = tb + seg * ckpt, t1 = std::min(nt, t0 + ckpt);
    P->fl.tdev = tdev;
    if ((rc = copy_state(seg, false))) break;
    // a step starts from F_z[p], which the previous step's last kernel leaves in
    // spectrum buffer 0: rebuild it from the restored p (same transform code)
    if (launch_zfwd(g, nb, P->fl.p, g.ntot, P->fl.spec[0], g.nspec, P->ops.tw[D - 1], P->st) != hipSuccess) {
      rc = fail(KW_EHIP, "adjoint: transform of the restored pressure failed");
      break;
    }
    if (launch_setint(tdev, nb, t0, P->st) != hipSuccess) { rc = fail(KW_EHIP, "adjoint: step counter reset failed"); break; }
    if (nonlin && hipMemcpyAsync(sbuf, P->fl.rhot, sizeof(float) * nf, hipMemcpyDeviceToDevice, P->st) != hipSuccess) {
      rc = fail(KW_EHIP, "adjoint: copy of the density sum failed");
      break;
    }
    for (int t = t0; t < t1 && rc == KW_OK; ++t) {
      P->fl.adjA = abuf ? abuf + (int64_t)(t - t0) * nf : nullptr;
      P->fl.adjB = bbuf ? bbuf + (int64_t)(t - t0) * nf : nullptr;
      P->fl.adjD = dbuf ? dbuf + (int64_t)(t - t0) * D * nf : nullptr;
      if ((rc = enqueue_step(P))) break;
      if (hipMemcpyAsync(pbuf + (int64_t)(t - t0) * nf, P->fl.p, sizeof(float) * nf, hipMemcpyDeviceToDevice, P->st) != hipSuccess)
        rc = fail(KW_EHIP, "adjoint: copy of p failed");
      else if (nonlin && hipMemcpyAsync(sbuf + (int64_t)(t - t0 + 1) * nf, P->fl.rhot, sizeof(float) * nf, hipMemcpyDeviceToDevice,
                                        P->st) != hipSuccess)
        rc = fail(KW_EHIP, "adjoint: copy of the density sum failed");
    }
    P->fl.adjA = P->fl.adjB = P->fl.adjD = nullptr;
    P->fl.tdev = nullptr;  // column passes of the reverse step must not skip instances at t = Nt
    for (int t = t1 - 1; t >= t0 && rc == KW_OK; --t) {
      af.pt = pbuf + (int64_t)(t - t0) * nf;
      af.At = abuf ? abuf + (int64_t)(t - t0) * nf : nullptr;
      af.Bt = bbuf ? bbuf + (int64_t)(t - t0) * nf : nullptr;
      af.St = sbuf ? sbuf + (int64_t)(t - t0 + 1) * nf : nullptr;
      af.Sp = sbuf ? sbuf + (int64_t)(t - t0) * nf : nullptr;
      af.Dt = dbuf ? dbuf + (int64_t)(t - t0) * D * nf : nullptr;
      rc = adjoint_step(P, af, Zbuf, t);
    }
  }
  P->fl.tdev = tdev;
  P->skip_gather = false;
  if (rc) { (void)hipStreamSynchronize(P->st); return rc; }
  if (with_p0) {  // the c-dependence of the set state: rho_i = p0 / (c0^2 p0div_i)
    for (int b = 0; b < nb && rc == KW_OK; ++b) {
      const kw::Inst& I = P->h_inst[b];
      const int64_t off = (int64_t)b * g.ntot;
      const float* rb[3] = {af.rb[0] + off, D > 1 ? af.rb[1] + off : nullptr, D > 2 ? af.rb[2] + off : nullptr};
      float w[3] = {0.f, 0.f, 0.f};
      for (int i = 0; i < D; ++i) w[i] = std::isinf(I.p0div[i]) ? 0.f : 1.f / I.p0div[i];
      if (launch_adj_p0(g.ntot, D, P->p0[b], I.c0sq, I.c0sq_s, rb, w, af.cb + off, P->st) != hipSuccess)
        rc = fail(KW_EHIP, "adjoint: initial-pressure term launch failed");
    }
    if (rc) { (void)hipStreamSynchronize(P->st); return rc; }
  }
  KW_HIP(hipMemcpyAsync(grad_c, af.cb, sizeof(float) * nf, hipMemcpyDeviceToHost, P->st));
  KW_HIP(hipStreamSynchronize(P->st));
  return KW_OK;
}

int kw_mode_gemm_bf16(int nmodes, int R, int N, int K, const void* A, const void* Bt, float* C, void* stream) {
  if (!A || !Bt || !C || nmodes < 1 || R < 32 || N < 32 || K < 8) return fail(KW_EINVAL, "mode GEMM: NULL pointer or empty shape");
  if (R % 32 || N % 32 || K % 8) return fail(KW_EINVAL, "mode GEMM: R and N must be multiples of 32, K of 8");
  if (((uintptr_t)A | (uintptr_t)Bt) & 7) return fail(KW_EINVAL, "mode GEMM: bf16 operands must be 8-byte aligned");
  if ((int64_t)nmodes > 65535) return fail(KW_EINVAL, "mode GEMM: at most 65535 modes per call");
  hipError_t e = launch_mode_gemm(nmodes, R, N, K, static_cast<const uint16_t*>(A), static_cast<const uint16_t*>(Bt), C,
                                  static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(KW_EHIP, std::string("mode GEMM launch: ") + hipGetErrorString(e));
  return KW_OK;
}

int kw_fno_head_gather_bf16(int64_t B, int C, int K, const void* v, const int* cols, void* sel, void* stream) {
  if (!v || !cols || !sel || B < 1 || C < 1 || K < 1) return fail(KW_EINVAL, "FNO head gather: NULL pointer or empty shape");
  if (B * C > 0x7fffffff) return fail(KW_EINVAL, "FNO head gather: too many slabs");
  hipError_t e = launch_head_gather(B, C, K, static_cast<const uint16_t*>(v), cols, static_cast<uint16_t*>(sel),
                                    static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(KW_EHIP, std::string("FNO head gather launch: ") + hipGetErrorString(e));
  return KW_OK;
}

int kw_fno_lift_bf16(int64_t B, int C, int N, const float* cmap, float cmid, float chalf, const float* lc, const void* g, void* v,
                     void* stream) {
  if (!cmap || !lc || !g || !v || B < 1 || N < 4) return fail(KW_EINVAL, "FNO lift: NULL pointer or empty shape");
  if (C != 16 && C != 32) return fail(KW_EUNSUPPORTED, "FNO lift: channel count must be 16 or 32");
  if (N % 4) return fail(KW_EINVAL, "FNO lift: the coarse grid edge must be a multiple of 4");
  if (!(chalf != 0.f)) return fail(KW_EINVAL, "FNO lift: the normalising half-range must be nonzero");
  if (B * N * (N / 4) > (int64_t)0xffffffff * 256 / 2) return fail(KW_EINVAL, "FNO lift: too many samples");
  if (((uintptr_t)cmap & 15) || (((uintptr_t)g | (uintptr_t)v) & 7))
    return fail(KW_EINVAL, "FNO lift: the map must be 16-byte and the bf16 arrays 8-byte aligned");
  hipError_t e = launch_fno_lift(B, C, N, cmap, cmid, chalf, lc, static_cast<const uint16_t*>(g), static_cast<uint16_t*>(v),
                                 static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(KW_EHIP, std::string("FNO lift launch: ") + hipGetErrorString(e));
  return KW_OK;
}

int kw_fno_dft_bf16(int dir, int64_t S, const void* in, const void* m1, const void* m2, void* out, void* stream) {
  if (!in || !m1 || !m2 || !out || S < 1) return fail(KW_EINVAL, "FNO DFT: NULL pointer or empty batch");
  if (dir != 1 && dir != -1) return fail(KW_EINVAL, "FNO DFT: direction must be 1 (forward) or -1 (inverse)");
  if (S > 0x7fffffff) return fail(KW_EINVAL, "FNO DFT: too many slabs");
  if (((uintptr_t)in | (uintptr_t)m1 | (uintptr_t)m2 | (uintptr_t)out) & 15) return fail(KW_EINVAL, "FNO DFT: arrays must be 16-byte aligned");
  hipError_t e;
  if (dir == 1)
    e = launch_dft_fwd(S, static_cast<const uint16_t*>(in), static_cast<const uint16_t*>(m1), static_cast<const uint16_t*>(m2),
                       static_cast<float*>(out), static_cast<hipStream_t>(stream));
  else
    e = launch_dft_inv(S, static_cast<const float*>(in), static_cast<const uint16_t*>(m1), static_cast<const uint16_t*>(m2),
                       static_cast<uint16_t*>(out), static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(KW_EHIP, std::string("FNO DFT launch: ") + hipGetErrorString(e));
  return KW_OK;
}

int kw_fno_pointwise_bf16(int B, int C, int64_t P, const void* v, const void* s, const float* pw, const float* pb, void* out, int gelu,
                          void* stream) {
  if (!v || !s || !pw || !pb || !out || B < 1 || P < 4) return fail(KW_EINVAL, "FNO pointwise: NULL pointer or empty shape");
  if (C != 16 && C != 32) return fail(KW_EUNSUPPORTED, "FNO pointwise: channel count must be 16 or 32");
  if (P % 4) return fail(KW_EINVAL, "FNO pointwise: positions per channel must be a multiple of 4");
  if (B > 65535) return fail(KW_EINVAL, "FNO pointwise: at most 65535 samples per call");
  if (((uintptr_t)v | (uintptr_t)s | (uintptr_t)out) & 7) return fail(KW_EINVAL, "FNO pointwise: bf16 arrays must be 8-byte aligned");
  hipError_t e = launch_fno_pointwise(B, C, P, static_cast<const uint16_t*>(v), static_cast<const uint16_t*>(s), pw, pb,
                                      static_cast<uint16_t*>(out), gelu ? 1 : 0, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(KW_EHIP, std::string("FNO pointwise launch: ") + hipGetErrorString(e));
  return KW_OK;
}

int kw_mode_contract_bf16(int B, int C, int md, int NX, int NY, const void* X, const int64_t xstride[4], const void* Bt, void* Y,
                          const int64_t ystride[4], void* stream) {
  if (!X || !Bt || !Y || !xstride || !ystride || B < 1 || md < 1) return fail(KW_EINVAL, "mode contraction: NULL pointer or empty shape");
  if ((2 * C) % 32) return fail(KW_EINVAL, "mode contraction: 2 C must be a multiple of 32");
  if (md > NY || 2 * md > NX) return fail(KW_EINVAL, "mode contraction: retained modes exceed the spectrum");
  if ((uintptr_t)Bt & 7) return fail(KW_EINVAL, "mode contraction: bf16 weights must be 8-byte aligned");
  if (2 * (int64_t)md * md > 65535) return fail(KW_EINVAL, "mode contraction: too many modes");
  for (int i = 0; i < 4; ++i)
    if (xstride[i] < 1 || ystride[i] < 1) return fail(KW_EINVAL, "mode contraction: strides must be positive");
  hipError_t e = launch_mode_contract(B, 2 * C, md, NX, static_cast<const float2*>(X), xstride, static_cast<const uint16_t*>(Bt),
                                      static_cast<float2*>(Y), ystride, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(KW_EHIP, std::string("mode contraction launch: ") + hipGetErrorString(e));
  return KW_OK;
}

int kw_plan_destroy(kw_plan* P) {
  if (!P) return KW_OK;
  (void)hipSetDevice(P->dev);
  if (P->st) (void)hipStreamSynchronize(P->st);
  delete P;
  return KW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// several plans as one captured step graph (kw_batch_*)
// ---------------------------------------------------------------------------
struct kw_batch {
  std::vector<kw_plan*> plans;
  int dev = 0;
  hipStream_t st = nullptr;  // the stream graphs are launched on
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  std::map<std::string, hipGraphExec_t> graphs;  // key: block length and each plan's form
  int graph_k = 8;
  ~kw_batch() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Steps each plan still has to run: its instances' largest Nt minus the common step index.
int64_t plan_remaining(const kw_plan* P) {
  int nt = 0;
  for (const Inst& I : P->h_inst) nt = std::max(nt, I.nt);
  return std::max<int64_t>(0, nt - P->t_host);
}

// One graph for k steps of every plan marked in `form` ('-': idle, 'f': full step, 'l': lean
// step): each plan's steps are a branch forked from and joined into the batch stream.
int batch_graph(kw_batch* B, int k, const std::string& form, hipGraphExec_t* out) {
  const std::string key = std::to_string(k) + ":" + form;
  auto it = B->graphs.find(key);
  if (it != B->graphs.end()) { *out = it->second; return KW_OK; }
  KW_HIP(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
  int rc = KW_OK;
  hipError_t e = hipEventRecord(B->fork, B->st);
  for (size_t i = 0; i < B->plans.size() && rc == KW_OK && e == hipSuccess; ++i) {
    if (form[i] == '-') continue;
    kw_plan* P = B->plans[i];
    e = hipStreamWaitEvent(P->st, B->fork, 0);
    for (int s = 0; s < k && rc == KW_OK && e == hipSuccess; ++s) rc = enqueue_step(P, form[i] == 'l');
    if (rc == KW_OK && e == hipSuccess) e = hipEventRecord(B->join[i], P->st);
    if (rc == KW_OK && e == hipSuccess) e = hipStreamWaitEvent(B->st, B->join[i], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(B->st, &graph);
  if (rc != KW_OK || e != hipSuccess || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != KW_OK) return rc;
    return fail(KW_EHIP, std::string("batch graph capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
  }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(KW_EHIP, std::string("batch graph instantiate: ") + hipGetErrorString(e));
  B->graphs[key] = ex;
  *out = ex;
  return KW_OK;
}

}  // namespace

extern "C" {

int kw_batch_create(kw_plan* const* plans, int n, kw_batch** out) {
  if (!plans || n < 1 || !out) return fail(KW_EINVAL, "batch needs at least one plan");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!plans[i]) return fail(KW_EINVAL, "NULL plan in batch");
    if (plans[i]->dev != plans[0]->dev) return fail(KW_EINVAL, "every plan of a batch must be on one device");
    for (int j = 0; j < i; ++j)
      if (plans[j] == plans[i]) return fail(KW_EINVAL, "a plan appears twice in the batch");
  }
  KW_HIP(hipSetDevice(plans[0]->dev));
  auto* B = new kw_batch();
  B->dev = plans[0]->dev;
  B->plans.assign(plans, plans + n);
  B->join.assign(n, nullptr);
  auto cleanup = [&](int code) { delete B; return code; };
  if (hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) != hipSuccess) return cleanup(fail(KW_EHIP, "batch stream creation failed"));
  if (hipEventCreateWithFlags(&B->fork, hipEventDisableTiming) != hipSuccess) return cleanup(fail(KW_EHIP, "batch event creation failed"));
  for (int i = 0; i < n; ++i)
    if (hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming) != hipSuccess) return cleanup(fail(KW_EHIP, "batch event creation failed"));
  *out = B;
  return KW_OK;
}

int kw_batch_run(kw_batch* B) {
  if (!B) return fail(KW_EINVAL, "NULL batch");
  KW_HIP(hipSetDevice(B->dev));
  const size_t n = B->plans.size();
  std::vector<int64_t> rem(n), qend(n);
  int64_t total = 0;
  int rc;
  for (size_t i = 0; i < n; ++i) {
    kw_plan* P = B->plans[i];
    if (P->t_host < 0) return fail(KW_EINVAL, "every plan of a batch must be initialised (kw_plan_init) first");
    if ((rc = sync_inst(P))) return rc;
    KW_HIP(hipStreamSynchronize(P->st));  // work enqueued on the plan outside the batch comes first
    rem[i] = plan_remaining(P);
    qend[i] = P->lean ? source_end(P) : -1;
    total = std::max(total, rem[i]);
  }
  std::vector<int64_t> done(n, 0);  // steps enqueued per plan
  for (int64_t s = 0; s < total;) {
    // block length: graph_k steps unless a plan switches to its lean form inside the block
    int k = (int)std::min<int64_t>(B->graph_k, total - s);
    for (size_t i = 0; i < n; ++i) {
      const kw_plan* P = B->plans[i];
      if (s >= rem[i] || qend[i] < 0) continue;
      const int64_t t = P->t_host + s;
      if (t < qend[i] && t + k > qend[i]) k = (int)(qend[i] - t);
    }
    std::string form(n, '-');
    for (size_t i = 0; i < n; ++i)
      if (s < rem[i]) form[i] = (qend[i] >= 0 && B->plans[i]->t_host + s >= qend[i]) ? 'l' : 'f';
    hipGraphExec_t g = nullptr;
    if ((rc = batch_graph(B, k, form, &g))) return rc;
    KW_HIP(hipGraphLaunch(g, B->st));
    for (size_t i = 0; i < n; ++i)
      if (form[i] != '-') done[i] += k;
    s += k;
  }
  KW_HIP(hipStreamSynchronize(B->st));
  for (size_t i = 0; i < n; ++i) B->plans[i]->t_host += done[i];
  return KW_OK;
}

int kw_batch_destroy(kw_batch* B) {
  if (!B) return KW_OK;
  (void)hipSetDevice(B->dev);
  if (B->st) (void)hipStreamSynchronize(B->st);
  delete B;
  return KW_OK;
}

int kw_spectral_apply(int ndim, const int n[3], const double spacing[3], int use_sg, double kap, int kind, int axis, double power,
                      const float* f, float* out, int device) {
  if (!n || !spacing || !f || !out) return fail(KW_EINVAL, "NULL argument");
  kw_plan_desc d{};
  d.ndim = ndim;
  for (int i = 0; i < 3; ++i) { d.n[i] = n[i]; d.spacing[i] = spacing[i]; }
  d.batch = 1;
  d.alpha_power = power + 2.0;  // frac1 = power
  d.use_sg = use_sg;
  d.device = device;
  if (kind < 0 || kind > 6 || kind == 4 || axis < 0 || axis >= ndim) return fail(KW_EINVAL, "bad operator kind or axis");
  kw_plan* P = nullptr;
  int rc = kw_plan_create(&d, &P);
  if (rc) return rc;
  const Geom& g = P->g;
  std::vector<float> ones(std::max(n[0], std::max(n[1], ndim == 3 ? n[2] : 1)), 1.f);
  kw_inst_desc id{};
  id.nt = 1; id.dt = 1.0; id.c_ref = 1.0; id.use_kspace = 1;
  for (int i = 0; i < ndim; ++i) { id.pml[i] = ones.data(); id.pml_sg[i] = ones.data(); }
  if ((rc = kw_plan_set_instance(P, 0, &id))) { kw_plan_destroy(P); return rc; }
  P->h_inst[0].kap = (float)kap;
  P->inst_dirty = true;
  auto run = [&]() -> int {
    int r;
    if ((r = sync_inst(P))) return r;
    float* fin;
    if ((r = P->mem.upload(&fin, f, (size_t)g.ntot))) return r;
    KW_LAUNCH(launch_setint(P->fl.tdev, 1, 0, P->st));
    const float2* twz = P->ops.tw[g.nd - 1];
    KW_LAUNCH(launch_zfwd(g, 1, fin, 0, P->fl.spec[0], 0, twz, P->st));
    if ((r = ypass(P, 0, {0}))) return r;
    static const int kinds[7] = {XK_COPY, XK_GRAD, XK_DIV, XK_FRAC1, XK_FRAC2, XK_SRCKAPPA, XK_UNSTAGGER};
    if ((r = xpass(P, {{0, 0, kinds[kind], axis}}, KK_XGRAD))) return r;
    if ((r = ypass(P, 1, {0}))) return r;
    KW_LAUNCH(launch_zinv(g, 1, P->fl.spec[0], 0, fin, 0, twz, 1.f, P->st));
    KW_HIP(hipStreamSynchronize(P->st));
    KW_HIP(hipMemcpy(out, fin, sizeof(float) * g.ntot, hipMemcpyDeviceToHost));
    return KW_OK;
  };
  rc = run();
  kw_plan_destroy(P);
  return rc;
}

}  // extern "C"
