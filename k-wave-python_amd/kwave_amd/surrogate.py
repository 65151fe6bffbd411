"""Fourier-neural-operator surrogate: sound-speed map -> sensor traces, trained on solver data.

The problem is ``workloads.surrogate_2d``: a 128^2 c map in [1400, 1650] m/s, a
32-cell additive line source with a coincident 32-cell receive line, 600 steps.
Training data come from the GPU solver (batched plans); the solver also measures
the surrogate's error and is its fallback.

Network (runs at 64^2, the c map averaged over 2 x 2 cells):
  lift     (c, x, y) -> 32 channels, pointwise
  4 x      v <- GELU(SpectralConv(v) + W v + b)   (16 x 16 retained modes, both kx signs)
  head     for every sensor, the feature columns at its 64^2 column and the
           neighbour on its side (2 x 32 channels x 64 depths) and the full-resolution
           c column under the sensor (128 depths; the injected amplitude scales with
           the local c) -> 600 samples by one shared linear map, plus a learned
           per-sensor mean trace.

The truncated 2-D transforms of every layer run in the library's BF16 MFMA kernel
(kw_fno_dft_bf16), in training as well: each transform is linear, so its gradient
is the opposite-direction kernel with transposed matrices.  The spectral
contraction - forward, input gradient and weight gradient - runs in
kw_mode_contract_bf16 / kw_mode_gemm_bf16 (FP32 accumulation).  Only 2 md x md of
the 64 x 33 Fourier modes are kept, so the truncated transforms are dense matrix
products rather than full FFTs.  At inference the pointwise part of a layer
(channel mix, bias, residual sum, GELU) is one kernel (kw_fno_pointwise_bf16);
in training it and the optimiser are PyTorch.

Per-map fallback: an error head (a small MLP on pooled last-layer features and a
few scalars of the map and of the prediction) estimates each map's relative L2
error; it is fitted on solver-checked calibration maps, and the routing threshold
is chosen on a second calibration split so that few accepted maps exceed the
error bound.  Maps above the threshold, and maps outside the training range, go
to the solver.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _abi, workloads

N_FINE = 128
N_COARSE = 64
MODES = 16
WIDTH = 32
LAYERS = 4
TRACE_SCALE = 1e5  # Pa; traces are learned in units of the source amplitude
C_MID, C_HALF = 1525.0, 125.0


# ---------------------------------------------------------------------------
# the BF16 MFMA contraction
# ---------------------------------------------------------------------------
def mode_gemm(a, bt):
    """C[m] = A[m] @ Bt[m]^T on the matrix cores; a: (M, R, K), bt: (M, N, K) bfloat16 CUDA tensors."""
    import torch

    if a.dtype != torch.bfloat16 or bt.dtype != torch.bfloat16 or not a.is_cuda:
        raise ValueError("mode_gemm takes bfloat16 device tensors")
    a, bt = a.contiguous(), bt.contiguous()
    m, r, k = a.shape
    n = bt.shape[1]
    if bt.shape[0] != m or bt.shape[2] != k:
        raise ValueError(f"mode_gemm shapes {tuple(a.shape)} x {tuple(bt.shape)}")
    out = torch.empty((m, r, n), dtype=torch.float32, device=a.device)
    stream = torch.cuda.current_stream(a.device).cuda_stream
    _abi.check(_abi.lib().kw_mode_gemm_bf16(m, r, n, k, C.c_void_p(a.data_ptr()), C.c_void_p(bt.data_ptr()),
                                            C.c_void_p(out.data_ptr()), C.c_void_p(stream)))
    return out


def _contract_fn():
    import torch

    class SpectralContract(torch.autograd.Function):
        """Y2[m] = X2[m] W2[m] with real-stacked complex modes (see oracle/fno_ref.py for the algebra)."""

        @staticmethod
        def forward(ctx, x2, w2):
            xb = x2.to(torch.bfloat16).contiguous()
            wb = w2.to(torch.bfloat16).contiguous()
            ctx.save_for_backward(xb, wb)
            return mode_gemm(xb, wb.transpose(1, 2))

        @staticmethod
        def backward(ctx, g2):
            xb, wb = ctx.saved_tensors
            gb = g2.to(torch.bfloat16).contiguous()
            dx = mode_gemm(gb, wb)                                              # G2 W2^T
            dw = mode_gemm(xb.transpose(1, 2), gb.transpose(1, 2))              # X2^T G2
            return dx, dw

    return SpectralContract


def mode_contract(x, bt, md):
    """Spectrum (complex64, [B, C, NX, NY]) -> zeros except the retained modes, contracted with bt.

    bt: bf16 [2 md md, 2C, 2C], row n = output index (interleaved re/im), i.e. the
    transposed real weight matrix of the forward pass or the plain one of the
    input gradient.  Runs kw_mode_contract_bf16 (gather + MFMA + scatter fused).
    """
    import torch

    b, c, nx, ny = x.shape
    if x.dtype != torch.complex64 or bt.dtype != torch.bfloat16 or not bt.is_contiguous():
        raise ValueError("mode_contract takes a complex64 spectrum and contiguous bf16 weights")
    if tuple(bt.shape) != (2 * md * md, 2 * c, 2 * c):
        raise ValueError(f"weights {tuple(bt.shape)} do not match {2 * md * md} modes of {c} channels")
    # the kernels write every retained mode; only a spectrum with more than those needs zeros elsewhere
    alloc = torch.empty if (nx == 2 * md and ny == md) else torch.zeros
    y = torch.view_as_complex(alloc((b, c, nx, ny, 2), dtype=torch.float32, device=x.device))
    xs = (C.c_int64 * 4)(*x.stride())
    ys = (C.c_int64 * 4)(*y.stride())
    stream = torch.cuda.current_stream(x.device).cuda_stream
    _abi.check(_abi.lib().kw_mode_contract_bf16(b, c, md, nx, ny, C.c_void_p(x.data_ptr()), xs, C.c_void_p(bt.data_ptr()),
                                                C.c_void_p(y.data_ptr()), ys, C.c_void_p(stream)))
    return y


def fno_lift(cmap, lc, g):
    """First-layer channels of the inference forward in one pass (kw_fno_lift_bf16): bf16
    (B, C, n, n) = (2 x 2 average of (cmap - C_MID) / C_HALF) * lc[c] + g[c], from the fine
    maps cmap (B, 2n, 2n) float32, lc (C,) float32 and the grid lift g (C, n, n) bf16."""
    import torch

    cmap = cmap.float().contiguous()
    b, n2 = cmap.shape[0], cmap.shape[-1]
    n = n2 // 2
    lc = lc.detach().float().reshape(-1).contiguous()
    g = g.detach().to(torch.bfloat16).reshape(-1, n, n).contiguous()
    c = lc.numel()
    v = torch.empty((b, c, n, n), dtype=torch.bfloat16, device=cmap.device)
    stream = torch.cuda.current_stream(cmap.device).cuda_stream
    _abi.check(_abi.lib().kw_fno_lift_bf16(b, c, n, C.c_void_p(cmap.data_ptr()), C_MID, C_HALF, C.c_void_p(lc.data_ptr()),
                                           C.c_void_p(g.data_ptr()), C.c_void_p(v.data_ptr()), C.c_void_p(stream)))
    return v


def fno_pointwise(v, s, pw, pb, gelu):
    """act(s + pb + pw v) per sample and position in one pass (inference; kw_fno_pointwise_bf16).

    v, s: bfloat16 CUDA tensors (B, C, ...) of the same shape; pw (C, C), pb (C,) or
    (C, 1) weights; returns bfloat16 of v's shape.
    """
    import torch

    if v.dtype != torch.bfloat16 or s.dtype != torch.bfloat16 or not v.is_cuda or v.shape != s.shape:
        raise ValueError("fno_pointwise takes two bfloat16 device tensors of one shape")
    v, s = v.contiguous(), s.contiguous()
    b, c = v.shape[:2]
    p = v[0, 0].numel()
    pw = pw.detach().float().contiguous()
    pb = pb.detach().float().reshape(-1).contiguous()
    out = torch.empty_like(v)
    stream = torch.cuda.current_stream(v.device).cuda_stream
    _abi.check(_abi.lib().kw_fno_pointwise_bf16(b, c, p, C.c_void_p(v.data_ptr()), C.c_void_p(s.data_ptr()),
                                                C.c_void_p(pw.data_ptr()), C.c_void_p(pb.data_ptr()),
                                                C.c_void_p(out.data_ptr()), int(bool(gelu)), C.c_void_p(stream)))
    return out


def fno_dft(direction, x, m1, m2):
    """Both truncated transforms of 64 x 64 slabs in one kernel (inference; kw_fno_dft_bf16).

    direction 1: x bf16 (..., 64, 64), m1 = Mx, m2 = My^T (bf16) -> float32 (..., 32, 32),
    the retained modes with (re, im) interleaved along the last axis.
    direction -1: x float32 (..., 32, 32) in that layout, m1 = Ny^T, m2 = Nx -> bf16 (..., 64, 64).
    """
    import torch

    fwd = direction == 1
    shape_in, shape_out = ((64, 64), (32, 32)) if fwd else ((32, 32), (64, 64))
    dt_in, dt_out = (torch.bfloat16, torch.float32) if fwd else (torch.float32, torch.bfloat16)
    if not x.is_cuda or x.dtype != dt_in or tuple(x.shape[-2:]) != shape_in:
        raise ValueError(f"fno_dft({direction}) takes {dt_in} device slabs of {shape_in}")
    x = x.contiguous()
    lead = x.shape[:-2]
    out = torch.empty(*lead, *shape_out, dtype=dt_out, device=x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    _abi.check(_abi.lib().kw_fno_dft_bf16(int(direction), x.numel() // (shape_in[0] * shape_in[1]), C.c_void_p(x.data_ptr()),
                                          C.c_void_p(m1.data_ptr()), C.c_void_p(m2.data_ptr()), C.c_void_p(out.data_ptr()),
                                          C.c_void_p(stream)))
    return out


def _dft_fns():
    """Truncated 2-D DFT of 64 x 64 slabs (16 modes) on the project's kernels, with gradients.

    Both transforms are linear, X = view(Mx v) My and v' = Nx view(Y Ny) (dft_mats), so each
    backward is the opposite-direction kernel with transposed matrices:
        dL/dv = Mx^T view(G My^T)   -> kw_fno_dft_bf16(-1) with (My, Mx^T) in the (Ny^T, Nx) slots
        dL/dY = view(Nx^T G) Ny^T   -> kw_fno_dft_bf16(+1) with (Nx^T, Ny) in the (Mx, My^T) slots
    Activations on the v side are bf16, spectra fp32 (FP32 accumulation throughout).
    """
    import torch

    class DftFwd(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v, conv):
            ctx.conv = conv
            return fno_dft(1, v.to(torch.bfloat16), conv.mx_b, conv.myt_b)

        @staticmethod
        def backward(ctx, g):
            c = ctx.conv
            return fno_dft(-1, g.float().contiguous(), c.my_b, c.mxt_b), None

    class DftInv(torch.autograd.Function):
        @staticmethod
        def forward(ctx, y, conv):
            ctx.conv = conv
            return fno_dft(-1, y.float().contiguous(), conv.nyt_b, conv.nx_b)

        @staticmethod
        def backward(ctx, g):
            c = ctx.conv
            return fno_dft(1, g.to(torch.bfloat16).contiguous(), c.nxt_b, c.ny_b), None

    return DftFwd, DftInv


_DFT = None


def fno_dft_train(direction, x, conv):
    """fno_dft with gradients (training): direction 1 maps bf16 (..., 64, 64) slabs to fp32 (..., 32, 32)
    retained modes, -1 maps them back; the matrices come from the SpectralConv2d module ``conv``."""
    global _DFT
    if _DFT is None:
        _DFT = _dft_fns()
    return (_DFT[0] if direction == 1 else _DFT[1]).apply(x, conv)


def head_gather(v, cols):
    """sel[b, k] = v[b, :, :, cols[k]] as (B, K, C, 64) bf16 (inference; kw_fno_head_gather_bf16).

    v: bf16 (B, C, 64, 64) device tensor; cols: int32 device tensor of column indices in [0, 64).
    """
    import torch

    if v.dtype != torch.bfloat16 or not v.is_cuda or tuple(v.shape[-2:]) != (64, 64) or cols.dtype != torch.int32:
        raise ValueError("head_gather takes bf16 (B, C, 64, 64) slabs and int32 column indices")
    v, cols = v.contiguous(), cols.contiguous()
    b, c = v.shape[:2]
    sel = torch.empty(b, cols.numel(), c, 64, dtype=torch.bfloat16, device=v.device)
    stream = torch.cuda.current_stream(v.device).cuda_stream
    _abi.check(_abi.lib().kw_fno_head_gather_bf16(b, c, cols.numel(), C.c_void_p(v.data_ptr()), C.c_void_p(cols.data_ptr()),
                                                  C.c_void_p(sel.data_ptr()), C.c_void_p(stream)))
    return sel


def gather_modes(x, md):
    """Retained modes of a spectrum as interleaved real pairs, [2 md md, B, 2C]."""
    import torch

    b, c, nx, _ = x.shape
    kx = torch.cat([torch.arange(md), torch.arange(nx - md, nx)]).to(x.device)
    sel = x.index_select(2, kx)[..., :md]
    return torch.view_as_real(sel.permute(2, 3, 0, 1)).reshape(2 * md * md, b, 2 * c)


def _modes_fn():
    import torch
    import torch.nn.functional as F

    class SpectralModes(torch.autograd.Function):
        """out = spectrum whose retained modes are X W (per mode), zeros elsewhere."""

        @staticmethod
        def forward(ctx, ft, w2, md):
            wb = w2.to(torch.bfloat16).contiguous()
            ctx.save_for_backward(ft, wb)
            ctx.md = md
            return mode_contract(ft, wb.transpose(1, 2).contiguous(), md)

        @staticmethod
        def backward(ctx, g):
            ft, wb = ctx.saved_tensors
            md = ctx.md
            dft = mode_contract(g, wb, md)                                       # G2 W2^T at the modes
            b = ft.shape[0]
            bp = (b + 7) // 8 * 8
            xt = gather_modes(ft, md).to(torch.bfloat16).transpose(1, 2)
            gt = gather_modes(g, md).to(torch.bfloat16).transpose(1, 2)
            if bp != b:
                xt, gt = F.pad(xt, (0, bp - b)), F.pad(gt, (0, bp - b))
            dw = mode_gemm(xt.contiguous(), gt.contiguous())                     # X2^T G2
            return dft, dw, None

    return SpectralModes


_MODES = None


def spectral_modes(ft, w2, md):
    global _MODES
    if _MODES is None:
        _MODES = _modes_fn()
    return _MODES.apply(ft, w2, md)


_CONTRACT = None


def spectral_contract(x2, w2):
    global _CONTRACT
    if _CONTRACT is None:
        _CONTRACT = _contract_fn()
    return _CONTRACT.apply(x2, w2)


# ---------------------------------------------------------------------------
# truncated 2-D DFT as matrix products
# ---------------------------------------------------------------------------
def dft_mats(n, md):
    """Real matrices of the truncated forward and inverse 2-D DFT of an n x n field.

    Retained modes: kx in [0, md) and [n - md, n) (row index j of the compact
    spectrum, 2 md rows), ky in [0, md).  Complex values are interleaved
    (re, im) along the axis they sit on.

      forward   B = Mx v       (4 md x n: rows 2j, 2j+1 = Re, Im of sum_x v e^{-i kx_j x})
                X = B' My      (B' = B viewed 2 md x 2n, row j = [Re B_j | Im B_j];
                                X: 2 md x 2 md, columns 2k, 2k+1 = Re, Im at ky = k)
      inverse   C = Y Ny       (2 md x 2n, row j = [Re C_j | Im C_j],
                                C_j(y) = sum_k w_k Y_jk e^{+i k y}, w_0 = 1, w_k = 2)
                v = Nx C'      (C' = C viewed 4 md x n; v = Re sum_j C_j e^{+i kx_j x} / n^2)

    X equals torch.fft.rfft2(v) at the retained modes and the inverse equals
    torch.fft.irfft2 of the spectrum that is zero elsewhere (the Nyquist column
    ky = n/2 is not retained).
    """
    kx = np.concatenate([np.arange(md), np.arange(n - md, n)])
    x = np.arange(n)
    ph = 2 * np.pi * np.outer(kx, x) / n                                    # (2 md, n)
    mx = np.zeros((4 * md, n))
    mx[0::2], mx[1::2] = np.cos(ph), -np.sin(ph)
    th = 2 * np.pi * np.outer(x, np.arange(md)) / n                         # (n, md)
    my = np.zeros((2 * n, 2 * md))
    my[:n, 0::2], my[n:, 0::2] = np.cos(th), np.sin(th)
    my[:n, 1::2], my[n:, 1::2] = -np.sin(th), np.cos(th)
    w = np.where(np.arange(md) == 0, 1.0, 2.0)[:, None]
    ny_ = np.zeros((2 * md, 2 * n))
    ny_[0::2, :n], ny_[1::2, :n] = w * np.cos(th.T), -w * np.sin(th.T)
    ny_[0::2, n:], ny_[1::2, n:] = w * np.sin(th.T), w * np.cos(th.T)
    nx_ = np.zeros((n, 4 * md))
    nx_[:, 0::2], nx_[:, 1::2] = np.cos(ph.T) / n ** 2, -np.sin(ph.T) / n ** 2
    return [a.astype(np.float32) for a in (mx, my, ny_, nx_)]


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------
def build_model(device="cuda", seed=0, width=WIDTH, layers=LAYERS, modes=MODES):
    import torch
    import torch.nn as nn
    import torch.nn.functional as F

    torch.manual_seed(seed)

    class SpectralConv2d(nn.Module):
        """Retained modes kx in [0, m) and [n - m, n), ky in [0, m), contracted per mode.

        Real and imaginary parts are interleaved per channel (x2[..., 2i + r]), so the
        FFT output reaches the contraction through one gather and a view; the weight
        matrix has the matching interleaved block pattern
        W2[2i + a, 2o + b] = [[Re W, Im W], [-Im W, Re W]][a][b].
        """

        def __init__(self, c, modes, n):
            super().__init__()
            self.c, self.modes = c, modes
            scale = 1.0 / (c * c)
            self.wr = nn.Parameter(scale * torch.randn(2 * modes * modes, c, c))
            self.wi = nn.Parameter(scale * torch.randn(2 * modes * modes, c, c))
            mats = dft_mats(n, modes)
            for name, a in zip(("mx", "my", "ny", "nx"), mats):
                self.register_buffer(name, torch.as_tensor(a))
            # bf16 operands of the fused transforms (kw_fno_dft_bf16: 64 x 64 slabs, 16 modes), and the
            # transposed pairs their gradients use
            self.fused_dft = n == 64 and modes == 16
            mx, my, ny_, nx_ = mats
            for name, a in (("mx_b", mx), ("myt_b", my.T), ("nyt_b", ny_.T), ("nx_b", nx_),
                            ("my_b", my), ("mxt_b", mx.T), ("nxt_b", nx_.T), ("ny_b", ny_)):
                self.register_buffer(name, torch.as_tensor(np.ascontiguousarray(a)).to(torch.bfloat16), persistent=False)

        def weights(self):
            c = self.c
            w2 = torch.stack([torch.stack([self.wr, self.wi], -1), torch.stack([-self.wi, self.wr], -1)], 2)
            return w2.reshape(-1, 2 * c, 2 * c)

        def forward(self, v):
            b, c, n, _ = v.shape
            md = self.modes
            if self.fused_dft and v.is_cuda:
                # both truncated transforms on the project's kernels, in inference and in training
                xm = fno_dft_train(1, v, self).reshape(b, c, 2 * md, md, 2)
                ym = spectral_modes(torch.view_as_complex(xm.contiguous()), self.weights(), md)
                return fno_dft_train(-1, torch.view_as_real(ym).reshape(b, c, 2 * md, 2 * md), self)
            bx = torch.matmul(self.mx, v).reshape(b, c, 2 * md, 2 * n)
            xm = torch.matmul(bx, self.my).float().reshape(b, c, 2 * md, md, 2)  # retained modes
            ym = spectral_modes(torch.view_as_complex(xm), self.weights(), md)
            cy = torch.matmul(torch.view_as_real(ym).reshape(b, c, 2 * md, 2 * md), self.ny)
            return torch.matmul(self.nx, cy.reshape(b, c, 4 * md, n))

    class FNO(nn.Module):
        def __init__(self, n_sens, nt, cols_a, cols_b, cols_f):
            super().__init__()
            self.width, self.layers = width, layers
            # default-scale initialisation (uniform, bound 1/sqrt(fan_in)); the extra head terms start at zero
            self.lift_c = nn.Parameter(torch.empty(width).uniform_(-3 ** -0.5, 3 ** -0.5))
            self.lift_g = nn.Conv2d(2, width, 1)
            self.spec = nn.ModuleList([SpectralConv2d(width, modes, N_COARSE) for _ in range(layers)])
            self.pw = nn.ParameterList([nn.Parameter(torch.empty(width, width).uniform_(-width ** -0.5, width ** -0.5))
                                        for _ in range(layers)])
            self.pb = nn.ParameterList([nn.Parameter(torch.zeros(width, 1)) for _ in range(layers)])
            self.head_a = nn.Linear(width * N_COARSE, nt)
            self.head_b = nn.Linear(width * N_COARSE, nt, bias=False)
            self.head_f = nn.Linear(N_FINE, nt, bias=False)
            nn.init.zeros_(self.head_b.weight)
            nn.init.zeros_(self.head_f.weight)
            self.mean = nn.Parameter(torch.zeros(n_sens, nt))
            ab = np.stack([cols_a, cols_b], 1).reshape(-1)
            if ab.min() < 0 or ab.max() >= N_COARSE:
                raise ValueError("sensor columns outside the coarse grid")
            self.register_buffer("cols_ab", torch.as_tensor(ab, dtype=torch.long))
            self.register_buffer("cols_ab32", torch.as_tensor(ab, dtype=torch.int32), persistent=False)
            self.register_buffer("cols_f", torch.as_tensor(cols_f, dtype=torch.long))
            g = torch.linspace(0, 1, N_COARSE)
            self.register_buffer("grid", torch.stack(torch.meshgrid(g, g, indexing="ij"))[None])
            self.n_sens = n_sens

        def forward(self, cmap, features=False):
            """Traces (b, n_sens, nt) in units of TRACE_SCALE; with ``features`` also the
            error head's inputs (b, 2 width + 3): per-channel mean and standard deviation of
            the last layer over the grid, the standard deviation and largest magnitude of the
            normalised map, and the prediction's distance from the mean trace set."""
            b = cmap.shape[0]
            n = N_COARSE
            wd = self.width
            # inference under bf16 autocast: the lift and each layer's pointwise part are one
            # fused kernel each
            fused = not torch.is_grad_enabled() and cmap.is_cuda and torch.is_autocast_enabled()
            if fused:
                v = fno_lift(cmap, self.lift_c, self.lift_g(self.grid))
            else:
                cn = (cmap - C_MID) / C_HALF
                x = F.avg_pool2d(cn[:, None], 2)                                     # (b, 1, n, n)
                v = x * self.lift_c[None, :, None, None] + self.lift_g(self.grid)    # lift of (c, x, y)
            for i in range(self.layers):
                if fused:
                    v = fno_pointwise(v, self.spec[i](v), self.pw[i], self.pb[i], i < self.layers - 1)
                    continue
                w = torch.matmul(self.pw[i], v.reshape(b, wd, n * n)) + self.pb[i]
                v = self.spec[i](v) + w.reshape(b, wd, n, n)
                if i < self.layers - 1:
                    v = F.gelu(v)
            # per sensor s the feature rows [column a_s | column b_s] (2 x wd x n), one product for both heads
            if fused:
                sel = head_gather(v, self.cols_ab32).reshape(b * self.n_sens, 2 * wd * n)
            else:
                sel = v.permute(0, 3, 1, 2).index_select(1, self.cols_ab).reshape(b * self.n_sens, 2 * wd * n)
            wab = torch.cat([self.head_a.weight, self.head_b.weight], 1)
            ha = F.linear(sel, wab, self.head_a.bias).reshape(b, self.n_sens, -1)
            ff = ((cmap.index_select(2, self.cols_f) - C_MID) / C_HALF).permute(0, 2, 1)  # (b, n_sens, 128)
            out = self.mean[None] + ha + self.head_f(ff)
            if not features:
                return out
            vf = v.float()
            dev = (out.float() - self.mean[None]).flatten(1).norm(dim=1) / self.mean.norm()
            flat = ((cmap - C_MID) / C_HALF).flatten(1)
            feats = torch.cat([vf.mean((2, 3)), vf.std((2, 3)), flat.std(1, keepdim=True),
                               flat.abs().amax(1, keepdim=True), dev[:, None]], 1)
            return out, feats

    kg, med, src, sen, pml, alpha = workloads.surrogate_2d(0)
    ys = np.where(np.asarray(sen.mask).any(axis=0))[0]
    cols_a = ys // 2
    cols_b = np.clip(cols_a + np.where(ys % 2 == 1, 1, -1), 0, N_COARSE - 1)
    return FNO(len(ys), kg.Nt, cols_a, cols_b, ys).to(device)


# ---------------------------------------------------------------------------
# data from the solver
# ---------------------------------------------------------------------------
def solve_batch(cmaps, plan_batch=256):
    """Recorded traces (n, 32, 600) float32 of the surrogate problem for each c map, batched on the GPU."""
    _torch_first()
    from .solvers.plan import DevicePlan
    from .solvers.setup import build_instance

    out = []
    for s in range(0, len(cmaps), plan_batch):
        sts = []
        for c in cmaps[s:s + plan_batch]:
            kg, med, src, sen, pml, alpha = workloads.surrogate_2d(0, c=c)
            sts.append(build_instance(kg, med, src, sen, pml_size=pml, pml_alpha=alpha))
        plan = DevicePlan(sts)
        try:
            plan.run(sts[0].Nt - plan.step_index(0))
            out.extend(plan.recorded(b, _abi.F_REC_P).copy() for b in range(len(sts)))
        finally:
            plan.close()
    return np.stack(out).astype(np.float32)


def make_dataset(seeds, plan_batch=256):
    cmaps = np.stack([workloads.surrogate_c_map(int(s)) for s in seeds]).astype(np.float32)
    return cmaps, solve_batch(cmaps.astype(np.float64), plan_batch)


def rel_l2(pred, true):
    """Per-sample relative L2 error of trace sets (n, 32, 600)."""
    d = (pred - true).reshape(len(true), -1)
    return np.linalg.norm(d, axis=1) / np.linalg.norm(true.reshape(len(true), -1), axis=1)


# ---------------------------------------------------------------------------
# training and inference
# ---------------------------------------------------------------------------
def train(model, cmaps, traces, epochs=40, batch=64, lr=2e-3, log=None):
    """Minimise the mean per-sample relative L2 of the traces with Adam (cosine schedule)."""
    import torch

    dev = next(model.parameters()).device
    x = torch.as_tensor(cmaps, device=dev)
    y = torch.as_tensor(traces / TRACE_SCALE, device=dev)
    with torch.no_grad():
        model.mean.copy_(y.mean(0))
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    steps = epochs * ((len(x) + batch - 1) // batch)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=steps, pct_start=0.1)
    g = torch.Generator(device="cpu").manual_seed(0)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(len(x), generator=g).to(dev)
        tot = 0.0
        for i in range(0, len(x), batch):
            idx = perm[i:i + batch]
            pred = model(x[idx])
            t = y[idx]
            loss = ((pred - t).flatten(1).norm(dim=1) / t.flatten(1).norm(dim=1)).mean()
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            sched.step()
            tot += float(loss.detach()) * len(idx)
        hist.append(tot / len(x))
        if log:
            log(f"epoch {ep + 1}/{epochs}: train rel-L2 {hist[-1]:.4f}")
    return hist


def predict(model, cmaps, batch=512, features=False, tta=False, ranges=False):
    """Surrogate traces (n, 32, 600) in Pa (pointwise layers and head in bf16 autocast).

    ``tta``: also run the map mirrored across the source line (y -> n - 1 - y; the
    acquisition is symmetric, so the solver's traces of the mirrored map are the
    traces in reverse sensor order, to ~5e-8) and average the two predictions.
    ``features``: also return the error head's inputs (n, F) float32 - the forward's
    features, plus with ``tta`` log10 of the relative disagreement of the two
    predictions.  ``ranges``: also return each map's (min, max) (n, 2), taken on the
    device from the uploaded map.

    Host traffic: each batch of maps is copied into one of two page-locked staging
    buffers and uploaded asynchronously, so that the copy of the next batch overlaps
    the forward of the current one; the traces come back asynchronously into a
    page-locked array that is returned as is (no further host copy)."""
    import torch

    dev = next(model.parameters()).device
    cm = np.ascontiguousarray(cmaps, dtype=np.float32)
    n = len(cm)
    src = torch.from_numpy(cm)
    stream = torch.cuda.current_stream(dev)
    stage = [torch.empty((min(batch, n),) + cm.shape[1:], dtype=torch.float32, pin_memory=True) for _ in range(2)]
    uploaded = [None, None]
    out, fts, rng = None, [], []
    with torch.no_grad(), torch.autocast("cuda", dtype=torch.bfloat16):
        for k, i in enumerate(range(0, n, batch)):
            nb = min(batch, n - i)
            buf = stage[k % 2]
            if uploaded[k % 2] is not None:
                uploaded[k % 2].synchronize()  # the upload from this buffer two batches ago has finished
            buf[:nb].copy_(src[i:i + nb])
            x = buf[:nb].to(dev, non_blocking=True)
            uploaded[k % 2] = torch.cuda.Event()
            uploaded[k % 2].record(stream)
            if ranges:
                flat = x.reshape(nb, -1)
                rng.append(torch.stack([flat.amin(1), flat.amax(1)], 1))
            r = model(torch.cat([x, x.flip(-1)]) if tta else x, features=features)
            if features:
                r, f = r
                f = f.float()
            r = r.float()
            if tta:
                a, m = r[:nb], r[nb:].flip(1)
                r = 0.5 * (a + m)
                if features:
                    dis = (a - m).flatten(1).norm(dim=1) / r.flatten(1).norm(dim=1).clamp_min(1e-30)
                    f = torch.cat([f[:nb], torch.log10(dis.clamp_min(1e-12))[:, None]], 1)
            r = r * TRACE_SCALE
            if out is None:
                out = torch.empty((n,) + tuple(r.shape[1:]), dtype=torch.float32, pin_memory=True)
            out[i:i + nb].copy_(r, non_blocking=True)
            if features:
                fts.append(f)
    stream.synchronize()
    res = [out.numpy()]
    if features:
        res.append(torch.cat(fts).cpu().numpy())
    if ranges:
        res.append(torch.cat(rng).cpu().numpy())
    return res[0] if len(res) == 1 else tuple(res)


def build_error_head(n_in, device="cuda", seed=1, hidden=64):
    """MLP estimating log10 of a map's relative L2 error from the forward's features."""
    import torch
    import torch.nn as nn

    torch.manual_seed(seed)

    class ErrorHead(nn.Module):
        def __init__(self):
            super().__init__()
            self.register_buffer("mu", torch.zeros(n_in))
            self.register_buffer("sd", torch.ones(n_in))
            self.net = nn.Sequential(nn.Linear(n_in, hidden), nn.GELU(), nn.Linear(hidden, hidden), nn.GELU(),
                                     nn.Linear(hidden, 1))

        def forward(self, f):
            return self.net((f - self.mu) / self.sd)[:, 0]

    return ErrorHead().to(device)


def fit_error_head(feats, err, steps=3000, lr=3e-3, weight_decay=1e-4, log=None):
    """Fit the error head to log10(err) of solver-checked maps (full-batch Adam, cosine decay)."""
    import torch

    head = build_error_head(feats.shape[1])
    dev = head.mu.device
    f = torch.as_tensor(feats, dtype=torch.float32, device=dev)
    t = torch.as_tensor(np.log10(np.maximum(err, 1e-6)), dtype=torch.float32, device=dev)
    with torch.no_grad():
        head.mu.copy_(f.mean(0))
        head.sd.copy_(f.std(0).clamp_min(1e-6))
    opt = torch.optim.AdamW(head.parameters(), lr=lr, weight_decay=weight_decay)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    for _ in range(steps):
        loss = ((head(f) - t) ** 2).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
    if log:
        log(f"error head: fit rms of log10 error {float(loss.detach()) ** 0.5:.3f}")
    return head.eval()


def predicted_error(head, feats):
    """The error head's estimate of each map's relative L2 error."""
    import torch

    with torch.no_grad():
        z = head(torch.as_tensor(feats, dtype=torch.float32, device=head.mu.device))
    return np.power(10.0, z.double().cpu().numpy())


def routing_threshold(pred_err, err, max_err, miss_rate):
    """Largest threshold on the predicted error such that, among these maps, the ones
    accepted (predicted <= threshold) with a true error above ``max_err`` are at most
    ``miss_rate`` of all maps."""
    order = np.argsort(pred_err, kind="stable")
    bad = np.cumsum(err[order] > max_err)
    ok = np.nonzero(bad <= miss_rate * len(err))[0]
    if len(ok) == 0:
        return 0.0
    k = ok[-1]
    # accept everything up to and including order[k]; stop below the next predicted value
    return float(pred_err[order[k]]) if k + 1 == len(order) else float(
        0.5 * (pred_err[order[k]] + pred_err[order[k + 1]]))


class Surrogate:
    """Trained network with the solver as per-map fallback.

    ``__call__`` returns traces for c maps and the mask of maps the solver ran.
    A map goes to the solver when it lies outside the training range
    [1400, 1650] m/s, or, once ``calibrate`` has fitted the error head, when the
    head's estimate of its error exceeds the calibrated threshold.  ``validate``
    measures the error on solver-checked maps; if the measured mean error exceeds
    ``max_err`` (2 %), the surrogate disables itself and every later call goes to
    the solver.
    """

    def __init__(self, model, max_err=0.02, tta=True):
        self.model, self.max_err, self.enabled, self.tta = model, max_err, True, tta
        self.head, self.threshold = None, None

    def validate(self, cmaps, traces=None):
        traces = solve_batch(np.asarray(cmaps, dtype=np.float64)) if traces is None else traces
        err = rel_l2(predict(self.model, np.asarray(cmaps, dtype=np.float32), tta=self.tta), traces)
        self.enabled = bool(err.mean() <= self.max_err)
        return err

    def calibrate(self, cmaps, traces=None, miss_rate=0.005, fit_fraction=0.6, policy="predicted", log=None):
        """Fit the error head on the first ``fit_fraction`` of these solver-checked maps and set
        the routing threshold: ``policy="predicted"`` routes maps whose estimated error exceeds
        ``max_err``; ``policy="budget"`` takes the largest threshold that leaves at most
        ``miss_rate`` of the remaining maps accepted while above ``max_err``.  Returns the
        calibration summary (on the maps not used for the fit)."""
        if policy not in ("predicted", "budget"):
            raise ValueError(f"policy must be 'predicted' or 'budget', got {policy!r}")
        cmaps = np.asarray(cmaps, dtype=np.float32)
        traces = solve_batch(cmaps.astype(np.float64)) if traces is None else traces
        pred, feats = predict(self.model, cmaps, features=True, tta=self.tta)
        err = rel_l2(pred, traces)
        nf = int(round(fit_fraction * len(cmaps)))
        if nf < 16 or len(cmaps) - nf < 16:
            raise ValueError("calibrate needs at least 16 maps in each split")
        self.head = fit_error_head(feats[:nf], err[:nf], log=log)
        pe = predicted_error(self.head, feats[nf:])
        self.threshold = (self.max_err if policy == "predicted"
                          else routing_threshold(pe, err[nf:], self.max_err, miss_rate))
        routed = pe > self.threshold
        return {"policy": policy, "n_fit": nf, "n_threshold": len(cmaps) - nf, "threshold": self.threshold,
                "routed_frac": float(routed.mean()),
                "miss_frac": float(((err[nf:] > self.max_err) & ~routed).mean()),
                "log10_err_corr": float(np.corrcoef(np.log10(pe), np.log10(np.maximum(err[nf:], 1e-6)))[0, 1])}

    def route(self, cmaps):
        """(traces of the surrogate for every map, mask of maps to send to the solver).

        A map goes to the solver when its sound speed leaves the training range or,
        with an error head, when the head's estimate of its error exceeds the threshold."""
        cmaps = np.asarray(cmaps, dtype=np.float32)
        lo, hi = workloads.SURROGATE_C_RANGE
        if not self.enabled:
            return None, np.ones(len(cmaps), dtype=bool)
        if self.head is None:
            pred, rng = predict(self.model, cmaps, tta=self.tta, ranges=True)
        else:
            pred, feats, rng = predict(self.model, cmaps, features=True, tta=self.tta, ranges=True)
        to_solver = (rng[:, 0] < lo - 1e-3) | (rng[:, 1] > hi + 1e-3)
        if self.head is not None:
            to_solver |= predicted_error(self.head, feats) > self.threshold
        return pred, to_solver

    def __call__(self, cmaps):
        """(traces of every map, mask of the maps the solver ran): the surrogate's traces
        with the routed maps' rows replaced by the solver's."""
        cmaps = np.asarray(cmaps, dtype=np.float32)
        pred, to_solver = self.route(cmaps)
        if pred is None:
            return solve_batch(cmaps.astype(np.float64)).astype(np.float32), to_solver
        if to_solver.any():
            pred[to_solver] = solve_batch(cmaps[to_solver].astype(np.float64))
        return pred, to_solver


def time_surrogate(model, cmaps, batch=1024, reps=5, head=None, tta=False):
    """Device time per sample of the surrogate forward (inputs already on the device); with
    ``tta`` the forward of the maps and their mirror images and the average; with an error
    ``head``, also the head's features and the head itself."""
    import torch

    x = torch.as_tensor(cmaps, device=next(model.parameters()).device)
    xb = x.repeat((batch + len(x) - 1) // len(x), 1, 1)[:batch].contiguous()

    def once():
        xx = torch.cat([xb, xb.flip(-1)]) if tta else xb
        if head is None and not tta:
            return model(xx)
        r, f = model(xx, features=True)
        f = f.float()
        if tta:
            a, m = r[:batch].float(), r[batch:].float().flip(1)
            r = 0.5 * (a + m)
            dis = (a - m).flatten(1).norm(dim=1) / r.flatten(1).norm(dim=1).clamp_min(1e-30)
            f = torch.cat([f[:batch], torch.log10(dis.clamp_min(1e-12))[:, None]], 1)
        return head(f) if head is not None else r

    with torch.no_grad(), torch.autocast("cuda", dtype=torch.bfloat16):
        once()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            once()
        torch.cuda.synchronize()
    return (time.perf_counter() - t0) / (reps * batch)


def time_solver(cmaps):
    """Device time per sample of the batched solver's time loop (plan uploaded, graph captured)."""
    from .solvers.plan import DevicePlan
    from .solvers.setup import build_instance

    sts = []
    for c in cmaps:
        kg, med, src, sen, pml, alpha = workloads.surrogate_2d(0, c=c)
        sts.append(build_instance(kg, med, src, sen, pml_size=pml, pml_alpha=alpha))
    plan = DevicePlan(sts)
    try:
        plan.run(sts[0].Nt)
        plan.reset()
        plan.run(0)
        t0 = time.perf_counter()
        plan.run(sts[0].Nt)
        return (time.perf_counter() - t0) / len(sts)
    finally:
        plan.close()


def _torch_first():
    """Initialise PyTorch's GPU context before the solver library touches the device.

    torch and the solver library each load a HIP runtime; initialising torch
    first is the order that has always worked (once, torch failed to find the
    GPU when the solver had run first in the same process).
    """
    import torch

    torch.cuda.init()


def benchmark(n_train=20000, n_test=1000, n_cal=3000, epochs=36, max_err=0.02, miss_rate=0.001, log=print):
    """Train on solver data, calibrate the per-map fallback, then report held-out error before
    and after routing and the surrogate's speed against the batched solver, with the fallback's
    solver time included.

    Seeds: training 5,000,000 + j, held-out 6,000,000 + j, calibration 7,000,000 + j.  The
    routed surrogate averages each map's prediction with that of its mirror image (tta) and
    feeds the disagreement of the two to the error head.
    """
    _torch_first()
    t0 = time.perf_counter()
    xtr, ytr = make_dataset(5_000_000 + np.arange(n_train))
    t_gen = time.perf_counter() - t0
    xte, yte = make_dataset(6_000_000 + np.arange(n_test))
    xca, yca = make_dataset(7_000_000 + np.arange(n_cal))
    log(f"surrogate: {n_train} training samples generated in {t_gen:.1f} s")
    model = build_model()
    t0 = time.perf_counter()
    train(model, xtr, ytr, epochs=epochs, log=log)
    t_train = time.perf_counter() - t0
    err = rel_l2(predict(model, xte), yte)
    err_tta = rel_l2(predict(model, xte, tta=True), yte)
    base = rel_l2(np.broadcast_to(ytr.mean(0), yte.shape), yte)
    sur = Surrogate(model, max_err=max_err, tta=True)
    cal_b = sur.calibrate(xca, yca, miss_rate=miss_rate, policy="budget", log=log)
    _, routed_b = sur.route(xte)
    after_b = np.where(routed_b, 0.0, err_tta)
    # the routing the surrogate keeps: maps whose estimated error exceeds max_err go to the solver
    cal = sur.calibrate(xca, yca, policy="predicted", log=log)
    _, routed = sur.route(xte)
    after = np.where(routed, 0.0, err_tta)  # routed maps carry the solver's own traces
    # the same routing without the mirror image: one forward per map, the head without the
    # disagreement feature
    sur_n = Surrogate(model, max_err=max_err, tta=False)
    cal_n = sur_n.calibrate(xca, yca, policy="predicted", log=log)
    _, routed_n = sur_n.route(xte)
    after_n = np.where(routed_n, 0.0, err)
    # throughput, inputs resident on the device: surrogate forward (+ mirror, error head) vs the
    # batched solver's time loop; with the fallback, the routed fraction also pays the solver
    t_sur = time_surrogate(model, xte)
    t_sur_r = time_surrogate(model, xte, head=sur.head, tta=True)
    t_sol = time_solver(xte.astype(np.float64))
    t_eff = t_sur_r + routed.mean() * t_sol
    t_eff_b = t_sur_r + routed_b.mean() * t_sol
    t_sur_n = time_surrogate(model, xte, head=sur_n.head, tta=False)
    t_eff_n = t_sur_n + routed_n.mean() * t_sol
    # host to host: c maps in host memory -> traces in host memory (the solver incl. its host set-up)
    t0 = time.perf_counter()
    predict(model, xte)
    t_sur_hh = (time.perf_counter() - t0) / n_test
    t0 = time.perf_counter()
    out, used = sur(xte)
    t_fb_hh = (time.perf_counter() - t0) / n_test
    t0 = time.perf_counter()
    solve_batch(xte.astype(np.float64), plan_batch=n_test)
    t_sol_hh = (time.perf_counter() - t0) / n_test
    err_out = rel_l2(out, yte)
    res = {"n_train": n_train, "n_test": n_test, "n_cal": n_cal, "epochs": epochs, "train_s": t_train,
           "datagen_s": t_gen, "width": model.width, "layers": model.layers, "modes": MODES, "grid": N_COARSE,
           "test_rel_l2_mean": float(err.mean()), "test_rel_l2_p95": float(np.percentile(err, 95)),
           "test_frac_above_2pct": float((err > max_err).mean()), "mean_trace_baseline_rel_l2": float(base.mean()),
           "tta_rel_l2_mean": float(err_tta.mean()), "tta_rel_l2_p95": float(np.percentile(err_tta, 95)),
           "tta_frac_above_2pct": float((err_tta > max_err).mean()),
           "calibration": cal, "routed_frac": float(routed.mean()),
           "budget_policy": {"calibration": cal_b, "routed_frac": float(routed_b.mean()),
                             "after_routing_frac_above_2pct": float((after_b > max_err).mean()),
                             "speedup_with_fallback": None},
           "no_mirror_policy": {"calibration": cal_n, "routed_frac": float(routed_n.mean()),
                                "after_routing_rel_l2_mean": float(after_n.mean()),
                                "after_routing_frac_above_2pct": float((after_n > max_err).mean()),
                                "routed_surrogate_ms_per_sample": 1e3 * t_sur_n, "speedup_with_fallback": t_sol / t_eff_n},
           "after_routing_rel_l2_mean": float(after.mean()), "after_routing_rel_l2_p95": float(np.percentile(after, 95)),
           "after_routing_frac_above_2pct": float((after > max_err).mean()),
           "after_routing_output_rel_l2_max": float(err_out.max()),
           "surrogate_ms_per_sample": 1e3 * t_sur, "routed_surrogate_ms_per_sample": 1e3 * t_sur_r,
           "solver_ms_per_sample_batched": 1e3 * t_sol, "speedup_vs_batched_solver": t_sol / t_sur,
           "speedup_with_fallback": t_sol / t_eff,
           "host_to_host_surrogate_ms_per_sample": 1e3 * t_sur_hh,
           "host_to_host_with_fallback_ms_per_sample": 1e3 * t_fb_hh,
           "host_to_host_solver_ms_per_sample": 1e3 * t_sol_hh,
           "host_to_host_speedup": t_sol_hh / t_sur_hh, "host_to_host_speedup_with_fallback": t_sol_hh / t_fb_hh}
    res["budget_policy"]["speedup_with_fallback"] = t_sol / t_eff_b
    log(f"surrogate: held-out rel-L2 {err.mean():.4f} ({err_tta.mean():.4f} with the mirror average; mean-trace "
        f"baseline {base.mean():.4f}), {100 * (err_tta > max_err).mean():.1f} % above {100 * max_err:.0f} %; routed "
        f"{100 * routed.mean():.1f} % -> {100 * (after > max_err).mean():.2f} % above; {1e3 * t_sur:.4f} ms/sample vs "
        f"solver {1e3 * t_sol:.3f} -> {t_sol / t_sur:.0f}x, {t_sol / t_eff:.1f}x routed with the fallback")
    return res, model


def _main():
    """`python -m kwave_amd.surrogate --forward [--reps R]`: time the inference forward
    of an untrained network on 1024 maps (the kernels are the trained network's; used
    for profiling), or with no flag run the full benchmark (training, calibration and
    held-out set sizes and epochs selectable) and print its results as one JSON line."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--forward", action="store_true", help="time the forward only (no data, no training)")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=1000)
    ap.add_argument("--cal", type=int, default=3000)
    ap.add_argument("--epochs", type=int, default=36)
    args = ap.parse_args()
    _torch_first()
    if not args.forward:
        res, _ = benchmark(n_train=args.train, n_test=args.test, n_cal=args.cal, epochs=args.epochs)
        print(json.dumps(res), flush=True)
        return
    model = build_model().eval()
    cmaps = np.stack([workloads.surrogate_c_map(6_000_000 + i) for i in range(64)]).astype(np.float32)
    t = time_surrogate(model, cmaps, reps=args.reps)
    print(f"surrogate forward: {1e3 * t:.4f} ms per map (batch 1024, {args.reps} repetitions)", flush=True)


if __name__ == "__main__":
    _main()
