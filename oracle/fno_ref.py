"""FP64 CPU reference of the surrogate's spectral-mode contraction (test infrastructure only).

Only ``tests/`` import this module: it is the checker of the BF16 MFMA kernel
behind ``kw_mode_gemm_bf16`` and of the spectral layer's gradients.

For retained Fourier mode m the layer computes the complex product
Y[m] = X[m] W[m] ((batch x Cin) (Cin x Cout)).  In real-stacked form
X2 = [Re X, Im X], W2 = [[Re W, Im W], [-Im W, Re W]], Y2 = X2 W2 = [Re Y, Im Y],
and for a real loss L with G2 = dL/dY2:  dL/dX2 = G2 W2^T,  dL/dW2 = X2^T G2,
dL/dRe W = dW2[:C, :C] + dW2[C:, C:],  dL/dIm W = dW2[:C, C:] - dW2[C:, :C].
The kernel rounds its operands to bfloat16 (round to nearest even) and
accumulates in FP32; ``bf16`` reproduces the operand rounding so that the
comparison isolates the accumulation.
Pins: tests/test_fno_ref.py (complex identity, finite differences).
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def stack_weights(wr, wi):
    """W2 = [[Re W, Im W], [-Im W, Re W]] for mode-major weights (M, Cin, Cout)."""
    top = np.concatenate([wr, wi], axis=2)
    bot = np.concatenate([-wi, wr], axis=2)
    return np.concatenate([top, bot], axis=1)


def contract(x2, w2):
    """Y2[m] = X2[m] W2[m]."""
    return np.einsum("mbi,mio->mbo", x2, w2)


def contract_grads(x2, w2, g2):
    """(dL/dX2, dL/dW2) for dL/dY2 = g2."""
    return np.einsum("mbo,mio->mbi", g2, w2), np.einsum("mbi,mbo->mio", x2, g2)


def weight_grads(dw2, c):
    """dL/dRe W and dL/dIm W from dL/dW2 (C = Cin = Cout)."""
    return dw2[:, :c, :c] + dw2[:, c:, c:], dw2[:, :c, c:] - dw2[:, c:, :c]
