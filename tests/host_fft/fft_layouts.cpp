// CPU check of the register-resident FFT (k-wave-python_amd/csrc/kw_fft.hpp) with its
// stage-ordered twiddle table and its LDS index maps: for every compiled line length, radix plan and tile shape the
// column passes (ColIdx, lines fastest) and the row passes (RowIdx, one line per
// group of threads) use, one POSIX thread per lane runs fft_reg on its owned
// elements against a shared array of the map's declared size, and the result is
// compared with a direct DFT in double precision.  Accesses outside the declared
// LDS size are caught by a guard band around the array.  Prints one line per case
// and "FAIL" for any case off by more than 1e-5 (relative to the peak).
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kw_fft.hpp"

pthread_barrier_t kw_host_barrier;
using namespace kw;
using cd = std::complex<double>;

static int g_fail = 0;

template <int N, int L, int PV, bool ROW>
struct Case {
  using F = FftShape<N, 1, PV>;
  using Idx = typename std::conditional<ROW, RowIdx<N>, ColIdx<L>>::type;
  static constexpr int NT = F::TPL * L;
  static constexpr int SIZE = ROW ? RowIdx<N>::size(L) : ColIdx<L>::size(N);
  static constexpr int GUARD = 4096;
  std::vector<float2> lds_mem = std::vector<float2>(SIZE + 2 * GUARD);
  float2 tw[stage_tw_size<N, PV>()];
  std::vector<cd> in, out;
  int dir = -1;
  struct Arg { Case* self; int tid; };
  static void* lane(void* p) {
    Arg* a = static_cast<Arg*>(p);
    Case& s = *a->self;
    const int tid = a->tid;
    const int c = ROW ? tid / F::TPL : tid % L;
    const int j0 = ROW ? tid % F::TPL : tid / L;
    float2 v[F::E];
    for (int m = 0; m < F::E; ++m) {
      const cd z = s.in[c * N + j0 + m * F::TPL];
      v[m] = make_float2((float)z.real(), (float)z.imag());
    }
    const Idx idx{};
    float2* lds = s.lds_mem.data() + GUARD;
    if (s.dir < 0) fft_reg<N, -1, Idx, 1, PV>(v, lds, s.tw, c, j0, idx);
    else fft_reg<N, +1, Idx, 1, PV>(v, lds, s.tw, c, j0, idx);
    for (int m = 0; m < F::E; ++m) s.out[c * N + j0 + m * F::TPL] = cd(v[m].x, v[m].y);
    return nullptr;
  }
  void run(int d) {
    if (NT > 1024) return;
    dir = d;
    for (int e = 0; e < stage_tw_size<N, PV>(); ++e) {
      const int m = stage_tw_power<N, PV>(e);
      tw[e] = make_float2((float)std::cos(-2 * M_PI * m / N), (float)std::sin(-2 * M_PI * m / N));
    }
    in.assign(N * L, cd());
    out.assign(N * L, cd());
    unsigned seed = 12345u + N * 31u + L;
    for (auto& z : in) {
      seed = seed * 1103515245u + 12345u;
      const double a = (seed >> 8) / double(1u << 24) - 0.5;
      seed = seed * 1103515245u + 12345u;
      const double b = (seed >> 8) / double(1u << 24) - 0.5;
      z = cd(a, b);
    }
    const float2 poison = make_float2(1e30f, 1e30f);
    for (auto& x : lds_mem) x = poison;
    pthread_barrier_init(&kw_host_barrier, nullptr, NT);
    std::vector<pthread_t> th(NT);
    std::vector<Arg> args(NT);
    for (int t = 0; t < NT; ++t) {
      args[t] = {this, t};
      pthread_create(&th[t], nullptr, &Case::lane, &args[t]);
    }
    for (auto& t : th) pthread_join(t, nullptr);
    pthread_barrier_destroy(&kw_host_barrier);
    bool guard_ok = true;
    for (int i = 0; i < GUARD; ++i)
      guard_ok &= lds_mem[i].x == poison.x && lds_mem[GUARD + SIZE + i].x == poison.x;
    double err = 0, peak = 0;
    for (int c = 0; c < L; ++c)
      for (int k = 0; k < N; ++k) {
        cd s = 0;
        for (int n = 0; n < N; ++n) s += in[c * N + n] * std::polar(1.0, dir * 2 * M_PI * double((int64_t)k * n % N) / N);
        err = std::max(err, std::abs(s - out[c * N + k]));
        peak = std::max(peak, std::abs(s));
      }
    const bool ok = guard_ok && err / peak < 1e-5;
    g_fail += !ok;
    std::printf("%s %s N=%d L=%d pv=%d dir=%+d E=%d rel_err=%.2e%s\n", ok ? "ok  " : "FAIL", ROW ? "row" : "col", N, L, PV,
                dir, F::E, err / peak, guard_ok ? "" : " (write outside the declared LDS size)");
  }
};

template <int N, int PV>
void lengths_pv() {
  Case<N, 1, PV, false>().run(-1);
  Case<N, 2, PV, false>().run(-1);
  Case<N, 4, PV, false>().run(+1);
  Case<N, 8, PV, false>().run(-1);
  Case<N, 16, PV, false>().run(+1);
  Case<N, 32, PV, false>().run(-1);
  Case<N, 1, PV, true>().run(-1);
  Case<N, 2, PV, true>().run(+1);
  Case<N, 4, PV, true>().run(-1);
}
template <int N>
void lengths() {
  lengths_pv<N, 0>();
  if constexpr (has_plan_variant(N)) lengths_pv<N, 1>();
  if constexpr ((N & (N - 1)) == 0) lengths_pv<N, 2>();  // radix-4 plan of the row passes
}

int main() {
#define KW_HOST_CASE(S) lengths<S>();
  KW_HOST_CASE(16) KW_HOST_CASE(32) KW_HOST_CASE(64) KW_HOST_CASE(128) KW_HOST_CASE(256) KW_HOST_CASE(512) KW_HOST_CASE(1024)
  KW_HOST_CASE(24) KW_HOST_CASE(48) KW_HOST_CASE(96) KW_HOST_CASE(192) KW_HOST_CASE(384) KW_HOST_CASE(768)
  KW_HOST_CASE(40) KW_HOST_CASE(80) KW_HOST_CASE(160) KW_HOST_CASE(320) KW_HOST_CASE(640)
  KW_HOST_CASE(56) KW_HOST_CASE(112) KW_HOST_CASE(224) KW_HOST_CASE(448)
  std::printf("%d failed\n", g_fail);
  return g_fail ? 1 : 0;
}
