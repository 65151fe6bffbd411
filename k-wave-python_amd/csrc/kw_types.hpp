// Device-visible parameter blocks shared by the kernels and the host plan code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace kw {

constexpr int kMaxJobs = 8;
constexpr int kMaxBufs = 12;

// Geometry of one plan (every instance in a plan has the same grid shape).
// Real fields are C-ordered with the last axis contiguous; spectra hold the
// half spectrum along that axis:  spec[row * nhp + k],  k < nh = nz/2 + 1.
// The pitch nhp > nh keeps the stride of the outermost-axis pass away from a
// power of two (measured: 3.3 -> 4.2 TB/s for 512^3 column copies).
struct Geom {
  int nd;            // 2 or 3
  int n[3];          // axis lengths (n[2] = 1 in 2-D)
  int nz;            // contiguous-axis length n[nd-1]
  int nh;            // nz / 2 + 1 (spectral columns)
  int nhp;           // row pitch of a spectrum (nh rounded up to 16 complex)
  int rows;          // ntot / nz
  int64_t ntot;      // real elements per instance
  int64_t nspec;     // rows * nhp complex elements per instance (storage)
};

// One time-varying source term (p, or one velocity component).
struct Src {
  int mode;            // SRC_* below
  int len;             // signal length (injection stops once t >= len)
  int nsigrow;         // rows in sig
  const float* sig;    // [nsigrow][len]
  const int* row_ptr;  // [rows + 1] CSR over grid rows (modes DIRICHLET, NOCORR, CHAIN)
  const int* col;      // position along the contiguous axis
  const float* scale;  // per-entry scale
  const int* sigrow;   // per-entry signal row
  const float* pattern;  // mode PATTERN: F^-1[kappa_s F[scale*mask]] (dense, ntot)
  int chain_buf;       // mode CHAIN: spectrum buffer index holding the filtered source
  float fac[3];        // per-axis multiplier applied on injection
};
enum { SRC_NONE = 0, SRC_DIRICHLET = 1, SRC_PATTERN = 2, SRC_CHAIN = 3, SRC_NOCORR = 4 };

// Per-instance scalars, operator tables and medium arrays (stride 0 = scalar).
struct Inst {
  int nt;
  float dt;
  float kap;            // c_ref * dt / 2 (kappa = sinc(kap |k|)); use_kspace=False -> 0
  const float* pml[3];  // 1-D multipliers along each axis (pressure grid)
  const float* pmlsg[3];  // staggered grid
  const float* c0sq; int c0sq_s;
  const float* rho0; int rho0_s;
  const float* dtr0[3]; int dtr0_s;  // dt / rho0_staggered_i
  const float* bona; int bona_s;
  const float* tau; int tau_s;       // absorption coefficient (grid), or null:
  const float* eta; int eta_s;       // dispersion coefficient (grid), or null:
  // with a scalar alpha the coefficients are derived from c0^2 in registers:
  // tau = tau_c (c0^2)^{(y-1)/2},  eta = eta_c (c0^2)^{y/2}
  float tau_c, eta_c, ym1h, yh;
  float p0div[3];       // split density i at t = 0 is p0 / (c0^2 p0div[i]): nd each by default
  Src srcp;
  Src srcu[3];
  // sensors
  int sens_kind;        // 0 none, 1 gather, 2 gather + linear weights
  int nsens;
  int ncorner;          // 1 (binary mask) or 2^nd (linear interpolation)
  const int64_t* sidx;  // [nsens * ncorner]
  const float* sw;      // [nsens * ncorner] (kind 2)
  int rec_start;        // first recorded step (0-based)
  float* rec_p;         // [nrec][nsens] or null
  float* rec_u[3];      // staggered velocity, [nrec][nsens] or null
  float* rec_uc[3];     // velocity shifted to the pressure grid, [nrec][nsens] or null
  float* agg_p;         // [3][nsens]: running max, min, sum of squares (or null)
};

// Spectral operator tables shared by all instances of a plan.
struct Ops {
  const float2* grad[3];   // i k e^{+ik d/2}   (use_sg) or i k; real part only at Nyquist
  const float2* div[3];    // i k e^{-ik d/2}   (use_sg) or i k
  const float2* unst[3];   // e^{-ik d/2}, real part at Nyquist
  const float* k[3];       // wavenumbers 2 pi fftfreq(n, d)
  float inv_ntot;
  float frac1, frac2;      // exponents y-2 and y-1 of the fractional Laplacians
  const float2* tw[3];     // twiddle tables exp(-2 pi i m / n_axis)
};

enum { FLAG_NONLINEAR = 1, FLAG_STOKES = 2, FLAG_POWERLAW_ABS = 4, FLAG_POWERLAW_DISP = 8 };

// Strided column pass (axes other than the contiguous one).
enum { XK_COPY = 0, XK_GRAD = 1, XK_DIV = 2, XK_FRAC1 = 3, XK_FRAC2 = 4, XK_SRCKAPPA = 5, XK_UNSTAGGER = 6 };
struct XJob {
  int in, out, kind, axis;
};
struct ColArgs {
  int mode;        // 0 forward, 1 inverse, 2 forward + jobs + inverse
  int axis;        // axis transformed (0 or 1)
  int nbuf;        // buffers transformed (modes 0/1)
  int bufs[kMaxBufs];
  int njob;
  XJob job[kMaxJobs];
  int64_t estride;  // element stride along the line
  int64_t ostride;  // stride between consecutive "outer" indices
  int nouter;       // outer index count (3-D y-pass: nx; x-pass: ny; 2-D: 1)
  int ncols;        // columns (= nh)
  int tile;         // 0: default tile width, 1: the alternative (chosen per plan by timing)
  int wide;         // 1: keep the chosen width on small grids too (parity tests of the large-grid tiles)
  int persist;      // passes whose workgroups may loop over tiles: bit 0 with jobs, bit 1 FFT-only
  int grid;         // > 0: workgroups per instance when looping (tests of the loop on small grids)
  int pv;           // radix-plan variant of the transform (lengths with a factor 5 or 7; chosen per plan by timing)
};

struct Fields {
  float* p;
  float* u[3];
  float* rho[3];
  float* rhot;
  float2* spec[kMaxBufs];
  float* uc[3];     // velocity on the pressure grid (only when it is recorded)
  int* tdev;        // per-instance step counter
  // generic-length family only (kw_generic.hip): real temporaries of the unfused
  // contiguous-axis updates, and the stash of a column pass's forward transform
  float* tmp[4];
  float2* gstash;
  // adjoint replay only (null otherwise): where the equation of state of the step
  // writes its absorption operand (rho0 div u for Stokes, L1[rho0 div u] for the
  // power law) and its dispersion operand L2[sum rho], [nb][ntot]
  float* adjA;
  float* adjB;
  float* adjD;  // per-axis Dop_i u_i of the step, [nd][nb][ntot] (nonlinear adjoint replay)
};

// State of the reverse-time sweep of kw_plan_adjoint_c (all [nb][ntot] unless noted).
struct AdjFields {
  float* pb;         // adjoint of p
  float* ub[3];      // adjoint of u_i
  float* rb[3];      // adjoint of rho_i
  float* cb;         // accumulated dJ/dc0
  float* R;          // sum_i rb_i * fac_i on source-active steps
  float* U[3];       // ub_i after the step's mass-update adjoint (dJ/du_i^t), axes with an additive velocity source
  float* gl[2];      // generic-length family, power law: L1[tau w] and L2[eta w] back on the grid
  const float* pt;   // forward p of the step being reversed
  const float* adj;  // dJ/d(recorded p), [nb][nrec][nsens]
  int nrec, nsens;
  // absorbing media: 0 none, 1 Stokes, 2 power law; disp: the power law has dispersion
  int absorb, disp;
  float ym1, y;      // alpha_power - 1 and alpha_power
  const float* At;   // forward absorption operand of the step being reversed (Fields::adjA)
  const float* Bt;   // forward dispersion operand (Fields::adjB)
  // B/A: sum of the densities after the step (St) and after the step before (Sp), and
  // the step's per-axis Dop_i u_i ([nd][nb][ntot], Fields::adjD)
  int nonlin;
  const float* St;
  const float* Sp;
  const float* Dt;
};

}  // namespace kw
