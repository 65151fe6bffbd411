// Host-side construction shared by the plan API and the allocation-free entry
// point: FP64-built 1-D operator tables (stored as FP32), plan geometry and the
// argument block of a column pass.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "kw_types.hpp"

namespace kw {

// Record msg as the thread's kw_last_error() text and return code.
int set_error(int code, const std::string& msg);

// 1-D operator tables for one axis (FP64 construction, FP32 storage)
struct AxisTables {
  std::vector<float2> grad, div, unst, tw;
  std::vector<float> k;
};

// Writes the n entries of each table into the caller's arrays.
inline void build_axis_into(int n, double d, bool use_sg, float2* grad, float2* div, float2* unst, float2* tw, float* kv) {
  const double two_pi = 2.0 * M_PI;
  for (int i = 0; i < n; ++i) {
    // numpy.fft.fftfreq ordering: 0..ceil(n/2)-1, then -floor(n/2)..-1
    const int f = (i < (n + 1) / 2) ? i : i - n;
    const double k = two_pi * f / (n * d);
    std::complex<double> ik(0.0, k);
    std::complex<double> sp = use_sg ? std::exp(std::complex<double>(0.0, k * d / 2)) : 1.0;
    std::complex<double> sm = use_sg ? std::exp(std::complex<double>(0.0, -k * d / 2)) : 1.0;
    std::complex<double> g = ik * sp, dv = ik * sm, un = std::exp(std::complex<double>(0.0, -k * d / 2));
    if (n % 2 == 0 && i == n / 2) {  // self-conjugate bin: keep the Hermitian part
      g = g.real(); dv = dv.real(); un = un.real();
    }
    grad[i] = make_float2((float)g.real(), (float)g.imag());
    div[i] = make_float2((float)dv.real(), (float)dv.imag());
    unst[i] = make_float2((float)un.real(), (float)un.imag());
    kv[i] = (float)k;
    const double a = -two_pi * i / n;
    tw[i] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
}

inline AxisTables build_axis(int n, double d, bool use_sg) {
  AxisTables t;
  t.grad.resize(n); t.div.resize(n); t.unst.resize(n); t.tw.resize(n); t.k.resize(n);
  build_axis_into(n, d, use_sg, t.grad.data(), t.div.data(), t.unst.data(), t.tw.data(), t.k.data());
  return t;
}

inline Geom make_geom(int nd, const int* n) {
  Geom g{};
  g.nd = nd;
  g.n[0] = n[0]; g.n[1] = n[1]; g.n[2] = nd == 3 ? n[2] : 1;
  g.nz = n[nd - 1];
  g.nh = g.nz / 2 + 1;
  g.nhp = (g.nh + 15) & ~15;
  g.ntot = (int64_t)g.n[0] * g.n[1] * g.n[2];
  g.rows = (int)(g.ntot / g.nz);
  g.nspec = (int64_t)g.rows * g.nhp;
  return g;
}

inline ColArgs col_args(const Geom& g, int axis, int mode) {
  ColArgs a{};
  a.mode = mode;
  a.axis = axis;
  a.ncols = g.nh;
  if (g.nd == 2) {
    a.estride = g.nhp; a.ostride = 0; a.nouter = 1;
  } else if (axis == 1) {
    a.estride = g.nhp; a.ostride = (int64_t)g.n[1] * g.nhp; a.nouter = g.n[0];
  } else {
    a.estride = (int64_t)g.n[1] * g.nhp; a.ostride = g.nhp; a.nouter = g.n[1];
  }
  return a;
}

}  // namespace kw
