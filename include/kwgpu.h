/*
 * kwgpu — C ABI of the MI355X k-space pseudospectral solver (libkwgpu.so).
 *
 * This is the drop-in boundary below k-wave-python's native solver path.  In the
 * reference, kspaceFirstOrder2D/3D (kwave/kspaceFirstOrder2D.py:141-218,
 * kwave/kspaceFirstOrder3D.py:133-211) dispatch to run_python_backend
 * (kwave/solvers/native.py:6-30), which constructs
 * kwave.solvers.kspace_solver.Simulation (kspace_solver.py:122-203) and calls
 * setup()/step()/run() (kspace_solver.py:205, 691, 749).  The Python mirror
 * (kwave_amd/solvers/kspace_solver.py) performs the same FP64 host set-up and
 * hands the time loop to the entry points below through ctypes.
 *
 * Conventions: every function returns 0 on success or a negative KW_E* code;
 * kw_last_error() holds a thread-local message.  The library owns all device
 * memory.  Host pointers are read (or written) only during the call.  A plan is
 * bound to one device and one stream, is not thread-safe, and its calls are
 * synchronous as seen by the host.  A plan holds `batch` instances of the same
 * grid shape; each instance has its own medium, sources, sensors, dt and Nt.
 */
#ifndef KWGPU_H
#define KWGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KW_ABI_VERSION 1

#define KW_OK 0
#define KW_EINVAL (-1)      /* bad argument  -> ValueError   */
#define KW_ENOMEM (-2)      /* out of memory -> MemoryError  */
#define KW_EHIP (-3)        /* HIP failure   -> RuntimeError */
#define KW_EUNSUPPORTED (-4) /* unsupported configuration -> ValueError */

/* physics flags (kw_plan_desc.flags) */
#define KW_NONLINEAR 1          /* B/A term, kspace_solver.py:526-550 */
#define KW_STOKES 2             /* y = 2 absorption, kspace_solver.py:491-494 */
#define KW_POWERLAW_ABS 4       /* tau |k|^{y-2} term, kspace_solver.py:495-498 */
#define KW_POWERLAW_DISP 8      /* eta |k|^{y-1} term, kspace_solver.py:521-524 */

/* field identifiers for kw_plan_set_field / kw_plan_read */
#define KW_F_C0SQ 0     /* c0^2 (kspace_solver.py:662) */
#define KW_F_RHO0 1     /* rho0 */
#define KW_F_DTR0_X 2   /* dt / rho0_sg_x (kspace_solver.py:663); _Y = 3, _Z = 4 */
#define KW_F_BONA 5     /* B/A */
#define KW_F_TAU 6      /* absorption coefficient tau */
#define KW_F_ETA 7      /* dispersion coefficient eta */
#define KW_F_P0 8       /* initial pressure (already smoothed), kspace_solver.py:679-689 */
#define KW_F_P 9        /* read: current pressure */
#define KW_F_UX 10      /* read: current staggered velocity x (_Y 11, _Z 12) */
#define KW_F_RHOX 13    /* read: split density x (_Y 14, _Z 15) */
#define KW_F_REC_P 16   /* read: recorded p, [nrec][nsens] */
#define KW_F_REC_UX 17  /* read: recorded staggered ux (uy 18, uz 19) */
#define KW_F_AGG_P 20   /* read: running [max, min, sum of squares] of recorded p, [3][nsens] */
#define KW_F_STEP 21    /* read: step counter (as float) */
#define KW_F_REC_UCX 23 /* read: recorded velocity on the pressure grid (unstaggered, kspace_solver.py:741-743), uy 24, uz 25 */
#define KW_F_ALPHA_NP 22 /* scalar absorption coefficient [Np/((rad/s)^y m)]: tau, eta derived from c0 on the device
                           (kspace_solver.py:495-498, 522) instead of grid-sized KW_F_TAU / KW_F_ETA */

/* source modes (kspace_solver.py:60-114) */
#define KW_SRC_DIRICHLET 1
#define KW_SRC_ADDITIVE 2          /* k-space corrected: F^-1[kappa_s F[s]] */
#define KW_SRC_ADDITIVE_NOCORR 4

/* sensor record flags */
#define KW_REC_P 1
#define KW_REC_P_AGG 2
#define KW_REC_UX 4  /* staggered velocity components: UY 8, UZ 16 */
#define KW_REC_UCX 32 /* velocity shifted to the pressure grid: UCY 64, UCZ 128 */

typedef struct kw_plan kw_plan;

typedef struct {
  int ndim;          /* 2 or 3 */
  int n[3];          /* grid shape, C order (last axis contiguous) */
  double spacing[3]; /* grid spacing per axis [m] */
  int batch;         /* instances in this plan (same shape) */
  int flags;         /* KW_NONLINEAR | KW_STOKES | KW_POWERLAW_* */
  double alpha_power;/* y, used for the fractional Laplacian exponents */
  int use_sg;        /* staggered grid shifts, kspace_solver.py:442-447 */
  int device;        /* HIP device ordinal */
} kw_plan_desc;

typedef struct {
  int nt;            /* number of time steps (kgrid.Nt) */
  double dt;         /* time step (kgrid.dt) */
  double c_ref;      /* max(c0), kspace_solver.py:231 */
  int use_kspace;    /* kappa on/off, kspace_solver.py:429-434 */
  const float* pml[3];    /* 1-D PML multipliers per axis (n[axis] values), kspace_solver.py:363-406 */
  const float* pml_sg[3]; /* staggered-grid PML multipliers */
} kw_inst_desc;

/* Version of this ABI (KW_ABI_VERSION). */
int kw_abi_version(void);
/* Thread-local message describing the last failure. */
const char* kw_last_error(void);
/* Number of visible HIP devices. */
int kw_device_count(int* count);
/* 1 if an axis of this length has a transform, else 0: the lengths with fused
 * kernels (2^k 16..1024, 3*2^k 24..768, 5*2^k 40..640, 7*2^k 56..448) and every
 * other length from 2 to 4096 through the generic-length kernels (the
 * reference's numpy.fft transforms any length, kspace_solver.py:702-705). */
int kw_size_supported(int n);

/* Replaces Simulation.__init__ + the grid part of setup(), kspace_solver.py:135-241. */
int kw_plan_create(const kw_plan_desc* desc, kw_plan** out);
/* Per-instance time step, c_ref, PML profiles; kspace_solver.py:225-231, 363-406. */
int kw_plan_set_instance(kw_plan* plan, int b, const kw_inst_desc* d);
/* Share of the initial density p0 / c0^2 that each split density receives at
 * t = 0 (share[i] >= 0, summing to 1; 1/ndim each unless set).  The reference
 * divides p0 evenly over its ndim split densities (kspace_solver.py:729-733); a
 * 1-D problem runs here as a 2-D grid that is uniform along axis 0, with the
 * whole initial density in the split density of the propagation axis (shares
 * {0, 1}), so that the sum evolves as the 1-D density does, PML included.  No
 * counterpart in the reference. */
int kw_plan_set_density_split(kw_plan* plan, int b, const double share[3]);
/* Medium and initial-condition arrays, n = 1 (scalar) or the grid size;
 * kspace_solver.py:228-230, 449-550, 647-689. */
int kw_plan_set_field(kw_plan* plan, int b, int which, const float* data, int64_t n);
/* Time-varying source term; which = -1 for p, 0..2 for ux/uy/uz.
 * idx: C-order flat indices of the mask points (ascending), scale: per-point
 * scale, sigrow: per-point row in sig ([nsigrow][len]), fac: per-axis
 * multiplier (pressure sources are injected into each rho_i).
 * Replaces _setup_source_operators + _build_source_op, kspace_solver.py:552-645, 60-114. */
int kw_plan_set_source(kw_plan* plan, int b, int which, int mode, int64_t npts, const int64_t* idx, const float* scale,
                       const int32_t* sigrow, const float* sig, int nsigrow, int len, const double fac[3]);
/* Sensor: nsens points with ncorner (index, weight) pairs each (binary masks:
 * ncorner = 1, weights NULL).  rec_start is 0-based.  Replaces
 * _setup_sensor_mask / _setup_cartesian_extract, kspace_solver.py:243-361. */
int kw_plan_set_sensor(kw_plan* plan, int b, int nsens, int ncorner, const int64_t* idx, const float* w, int rec_start,
                       unsigned rec_flags);
/* Zero the state and apply the t = 0 initial-value override where p0 is set
 * (kspace_solver.py:651-654, 729-733); records step 0 for those instances. */
int kw_plan_init(kw_plan* plan);
/* Advance every active instance by up to nsteps steps (Simulation.step,
 * kspace_solver.py:691-747); instances stop at their own Nt. */
int kw_plan_run(kw_plan* plan, int64_t nsteps);
/* Enqueue nsteps steps on the plan's stream and return without waiting, so
 * that several plans (e.g. the grid-size groups of an uneven batch) run
 * concurrently; kw_plan_sync waits for them.  The batched counterpart of
 * calling Simulation.run() per instance (kspace_solver.py:749-761). */
int kw_plan_launch(kw_plan* plan, int64_t nsteps);
int kw_plan_sync(kw_plan* plan);

/* Several initialised plans on one device -- typically the grid-shape groups of
 * an uneven sweep, each a batch of same-shape instances -- advanced together:
 * every block of steps is one captured graph in which each plan's step sequence
 * is a parallel branch, launched with a single graph launch.  kw_batch_run runs
 * every plan until each of its instances has reached its own Nt (finished plans
 * leave the graph; a plan switches to its lean step form once its filtered
 * sources have ended).  Results equal those of running the plans one by one.
 * The plans stay owned by the caller and must outlive the batch; no other call
 * may use them during kw_batch_run.  The batched counterpart of calling
 * Simulation.run() once per instance (kspace_solver.py:749-778). */
typedef struct kw_batch kw_batch;
int kw_batch_create(kw_plan* const* plans, int n, kw_batch** out);
int kw_batch_run(kw_batch* batch);
int kw_batch_destroy(kw_batch* batch);
/* Like kw_plan_run, without graph capture, bracketing every kernel launch with
 * HIP events on the plan stream.  Per kernel kind k (< kw_kernel_kind_count())
 * it returns the summed device time [ms], the algorithmic bytes moved and the
 * launch count.  No counterpart in the reference (its loop has no per-kernel
 * timing); used by the benchmark to place each kernel against the HBM roofline. */
int kw_plan_run_profiled(kw_plan* plan, int64_t nsteps, double* kernel_ms, double* kernel_bytes, int64_t* launches);
int kw_kernel_kind_count(void);
const char* kw_kernel_kind_name(int kind);
/* Tile-width variants the plan picked by timing at its first kw_plan_init (0:
 * the default width, 1: half of it) for the FFT-only column passes and for the
 * passes with spectral jobs; KW_AUTOTUNE=0 in the environment keeps 0.  No
 * counterpart in the reference (its FFTs are library calls). */
int kw_plan_tiles(kw_plan* plan, int* fft_tile, int* job_tile);
/* Radix-plan variants of the same two pass kinds (0: radix 2 next to a factor 5
 * or 7, 1: radix 4 next to it, fewer stages for more registers; 0 for lengths
 * without such a factor).  0 unless KW_COL_PLANS in the environment forces them
 * ("ab") or adds them to the timing ("tune"): the plans agree to FP32 rounding,
 * not bit for bit.  No counterpart in the reference. */
int kw_plan_radix_plans(kw_plan* plan, int* fft_plan, int* job_plan);
/* Copy a device field of instance b to host memory (n floats). */
int kw_plan_read(kw_plan* plan, int b, int which, float* out, int64_t n);
/* Number of floats kw_plan_read returns for `which`. */
int kw_plan_field_size(kw_plan* plan, int b, int which, int64_t* n);
/* Release every device resource of the plan. */
int kw_plan_destroy(kw_plan* plan);

/* Gradient of a functional J of the recorded pressure traces with respect to
 * the sound speed c0 of every instance, by the discrete adjoint of the step
 * (reverse time, forward state checkpointed every `ckpt` steps and replayed
 * one segment at a time).  adj = dJ/d(recorded p), [batch][nrec][nsens] (the
 * device record layout); grad_c: [batch][grid size] out.  The gradient holds
 * c_ref, dt, the PML and kappa fixed and includes the dependence of the
 * additive source scale 2 dt / (D c0 d) on c0 at the source points and, with an
 * initial pressure p0, of the initial split densities p0 / (c0^2 p0div_i) (the
 * state after step 0 is set from p0, kspace_solver.py:729-733, and only those
 * densities depend on c0), and in absorbing media the dependence of tau and eta
 * on c0 (kspace_solver.py:471-516), for a dirichlet pressure source the
 * written value sig / (D c0^2) at its points (:613-617), and for an additive
 * velocity source its scale 2 c0 dt / d_i (:636-642).  Plans with additive,
 * uncorrected additive or dirichlet pressure and velocity sources and/or p0,
 * lossless or with Stokes or power-law absorption (with or without dispersion),
 * with or without B/A nonlinearity (kspace_solver.py:542-550); every
 * instance records p with a common Nt, sensor count and record start.  The plan's fields
 * and records afterwards are those of the forward run's first segment replay:
 * call kw_plan_init before stepping it again.  The reference has no adjoint:
 * this differentiates its Simulation.step (kspace_solver.py:691-747) and source
 * scaling (:619-624). */
int kw_plan_adjoint_c(kw_plan* plan, const float* adj, int ckpt, float* grad_c);

/* Mode-batched GEMM of the Fourier-neural-operator surrogate, BF16 inputs with
 * FP32 accumulation on the matrix cores:  C[m] = A[m] * Bt[m]^T for m < nmodes,
 * A: [nmodes][R][K] bf16, Bt: [nmodes][N][K] bf16, C: [nmodes][R][N] fp32, all
 * device pointers (row-major, contiguous, bf16 operands 8-byte aligned),
 * R and N multiples of 32, K a multiple of 8; enqueued on `stream` (a
 * hipStream_t, NULL for the default stream) without waiting.  The spectral
 * layers' forward and both gradients are such GEMMs over real-stacked complex
 * modes.  The reference has no learned surrogate; no counterpart. */
int kw_mode_gemm_bf16(int nmodes, int R, int N, int K, const void* A, const void* Bt, float* C, void* stream);

/* The surrogate's spectral contraction with the mode gather and scatter fused:
 * for every retained mode (kx rows 0..md-1 and NX-md..NX-1, ky < md) and batch
 * row b < B, Y[b][:, kx, ky] = X[b][:, kx, ky] W[mode] with complex channels
 * interleaved as (re, im) pairs, i.e. the real GEMM [B x 2C] x [2C x 2C] whose
 * bf16 transposed weights are Bt [2 md md][2C][2C] (row n = output index).  X
 * and Y are complex64 device arrays [B][C][NX][NY] given by their element
 * strides {batch, channel, kx, ky}; Y is written only at the retained modes.
 * 2C must be a multiple of 32.  Enqueued on `stream`.  No counterpart in the
 * reference (it has no surrogate). */
int kw_mode_contract_bf16(int B, int C, int md, int NX, int NY, const void* X, const int64_t xstride[4], const void* Bt, void* Y,
                          const int64_t ystride[4], void* stream);

/* Truncated 2-D DFT of the surrogate's 64 x 64 channel slabs for inference,
 * both axes in one pass (bf16 operands, FP32 accumulation, the intermediate
 * rounded to bf16).  S slabs; all arrays device pointers, 16-byte aligned.
 * dir = 1:  in = v bf16 [S][64][64], m1 = Mx bf16 [64][64] (rows interleave
 *           re/im of 32 retained kx), m2 = My^T bf16 [32][128];
 *           out = X fp32 [S][32][32] = (Mx v viewed 32 x 128) My, i.e. complex
 *           [S][32 kx][16 ky] interleaved (re, im).
 * dir = -1: in = Y fp32 [S][32][32] (same layout), m1 = Ny^T bf16 [128][32],
 *           m2 = Nx bf16 [64][64]; out = bf16 [S][64][64] = Nx (Y Ny viewed 64 x 64).
 * The matrices are those of the network's truncated transforms.  Enqueued on
 * `stream`.  No counterpart in the reference (it has no surrogate). */
int kw_fno_dft_bf16(int dir, int64_t S, const void* in, const void* m1, const void* m2, void* out, void* stream);

/* Input lift of the surrogate for inference: v[b][c][i][j] = bf16(x[b][i][j] lc[c] +
 * g[c][i][j]) with x the 2 x 2 average of the normalised fine map
 * (cmap - cmid) / chalf; cmap fp32 [B][2N][2N] (16-byte aligned), lc fp32 [C],
 * g bf16 [C][N][N] (the lift of the coordinate grid), v bf16 [B][C][N][N]; device
 * pointers, C 16 or 32, N a multiple of 4.  Enqueued on `stream`.  No counterpart
 * in the reference (it has no surrogate). */
int kw_fno_lift_bf16(int64_t B, int C, int N, const float* cmap, float cmid, float chalf, const float* lc, const void* g, void* v,
                     void* stream);

/* Sensor-head gather of the surrogate for inference: sel[b][k][c][x] =
 * v[b][c][x][cols[k]] for b < B, k < K, c < C, x < 64, with v bf16 [B][C][64][64],
 * cols int32 [K] (device, every value in [0, 64): the caller checks), sel bf16
 * [B][K][C][64]; device pointers.  Enqueued on `stream`.  No counterpart in the
 * reference (it has no surrogate). */
int kw_fno_head_gather_bf16(int64_t B, int C, int K, const void* v, const int* cols, void* sel, void* stream);

/* Pointwise part of a surrogate layer for inference, fused into one pass:
 * out[b][o][p] = act(s[b][o][p] + pb[o] + sum_c pw[o][c] v[b][c][p]), with v (the
 * layer input), s (its spectral-conv output) and out bf16 device arrays
 * [B][C][P] (contiguous, 8-byte aligned), pw [C][C] and pb [C] fp32 device
 * arrays, act the exact GELU when gelu != 0 else the identity.  C is 16 or 32,
 * P a multiple of 4; FP32 accumulation, one bf16 rounding of the result.
 * Enqueued on `stream`.  No counterpart in the reference (it has no surrogate). */
int kw_fno_pointwise_bf16(int B, int C, int64_t P, const void* v, const void* s, const float* pw, const float* pb, void* out, int gelu,
                          void* stream);

/* Spectral operator f -> real(ifftn(op * fftn(f))) on the device, the
 * reference's Simulation._diff (kspace_solver.py:781-796), for a single field.
 * kind: 0 identity, 1 gradient (op_grad), 2 divergence (op_div),
 * 3 |k|^{power}, 5 source kappa, 6 unstagger;  kap = c_ref*dt/2 (0: kappa = 1).
 * Used by the parity tests of the transform kernels. */
int kw_spectral_apply(int ndim, const int n[3], const double spacing[3], int use_sg, double kap, int kind, int axis,
                      double power, const float* f, float* out, int device);

/* Allocation-free variant of kw_spectral_apply: the same operator
 * real(ifftn(op * fftn(f))) with the same kernels (bit-identical results), for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  f and out are device arrays of the grid size on the calling
 * thread's current device; work is enqueued on `stream` (a hipStream_t, NULL for
 * the default stream) and the call returns once it is done.  The half spectrum
 * and the operator tables live in a fixed 64 MiB device array per GPU, held
 * until the process ends: grids whose spectrum and tables exceed it are refused
 * with KW_EINVAL before any device work (e.g. 3-D 192^3 fits, 256^3 does not).
 * Calls that overlap on one GPU share that array and race.  No hipMalloc,
 * hipFree or other allocating routine is called. */
int kw_spectral_apply_static(int ndim, const int n[3], const double spacing[3], int use_sg, double kap, int kind, int axis,
                             double power, const float* f, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* KWGPU_H */
