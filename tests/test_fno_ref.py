"""Pin the FP64 reference of the spectral-mode contraction (oracle/fno_ref.py)."""
import numpy as np


def test_real_stacking_equals_complex_product():
    from oracle.fno_ref import contract, stack_weights

    rng = np.random.default_rng(0)
    m, b, c = 5, 7, 6
    x = rng.standard_normal((m, b, c)) + 1j * rng.standard_normal((m, b, c))
    w = rng.standard_normal((m, c, c)) + 1j * rng.standard_normal((m, c, c))
    y = np.einsum("mbi,mio->mbo", x, w)
    y2 = contract(np.concatenate([x.real, x.imag], axis=2), stack_weights(w.real, w.imag))
    assert np.allclose(y2[..., :c], y.real, atol=1e-12) and np.allclose(y2[..., c:], y.imag, atol=1e-12)


def test_gradients_match_finite_differences():
    from oracle.fno_ref import contract, contract_grads, stack_weights, weight_grads

    rng = np.random.default_rng(1)
    m, b, c = 3, 4, 5
    x2 = rng.standard_normal((m, b, 2 * c))
    wr, wi = rng.standard_normal((m, c, c)), rng.standard_normal((m, c, c))
    t = rng.standard_normal((m, b, 2 * c))

    def loss(x2_, wr_, wi_):
        return 0.5 * float(np.sum((contract(x2_, stack_weights(wr_, wi_)) - t) ** 2))

    g2 = contract(x2, stack_weights(wr, wi)) - t
    dx2, dw2 = contract_grads(x2, stack_weights(wr, wi), g2)
    dwr, dwi = weight_grads(dw2, c)
    h = 1e-6
    for arr, grad, which in ((x2, dx2, 0), (wr, dwr, 1), (wi, dwi, 2)):
        for idx in [(0, 0, 0), (1, 2, 3), (2, 3, 1)]:
            a = [x2.copy(), wr.copy(), wi.copy()]
            a[which][idx] += h
            lp = loss(*a)
            a[which][idx] -= 2 * h
            lm = loss(*a)
            assert abs((lp - lm) / (2 * h) - grad[idx]) < 1e-6 * max(1.0, abs(grad[idx]))


def test_bf16_rounding():
    from oracle.fno_ref import bf16

    # exactly representable values stay; ties round to even
    assert bf16(np.array([1.0, -2.5, 3.0]))[1] == -2.5
    one_ulp = 2.0 ** -7
    assert bf16(np.array([1.0 + one_ulp / 2]))[0] == 1.0            # tie -> even (1.0)
    assert bf16(np.array([1.0 + 1.5 * one_ulp]))[0] == 1.0 + 2 * one_ulp  # tie -> even
    assert bf16(np.array([1.0 + 0.6 * one_ulp]))[0] == 1.0 + one_ulp
