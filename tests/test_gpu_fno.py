"""The surrogate's BF16 MFMA spectral contraction and training, on the GPU.

* kw_mode_gemm_bf16 is exact on integer-valued bf16 operands (this also fixes the
  MFMA lane layout the kernel assumes);
* forward, input gradient and weight gradients of the spectral contraction match
  the FP64 CPU reference (oracle/fno_ref.py) on the same bf16-rounded operands;
* the truncated DFTs of training (kw_fno_dft_bf16 with its gradients) against FP64
  products of the same matrices, and the network's parameter gradients against the
  FP32 matrix-product form of the transforms;
* a short training run on solver data beats the method without the network
  (predicting the training-mean traces) on held-out maps;
* the surrogate hands out-of-range maps to the solver, routes maps whose estimated
  error is above the calibrated threshold to it (those carry the solver's traces),
  and disables itself when validation shows more than the allowed error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_mode_gemm_exact_on_integers():
    import torch

    from kwave_amd.surrogate import mode_gemm

    g = torch.Generator().manual_seed(0)
    m, r, n, k = 3, 64, 96, 40
    a = torch.randint(-8, 9, (m, r, k), generator=g).float()
    bt = torch.randint(-8, 9, (m, n, k), generator=g).float()
    c = mode_gemm(a.to("cuda", torch.bfloat16), bt.to("cuda", torch.bfloat16)).cpu()
    want = torch.bmm(a.double(), bt.double().transpose(1, 2))
    assert torch.equal(c.double(), want)


def test_spectral_contract_matches_fp64_reference():
    import torch

    from kwave_amd.surrogate import spectral_contract
    from oracle.fno_ref import bf16, contract, contract_grads, stack_weights, weight_grads

    rng = np.random.default_rng(1)
    m, b, c = 12, 64, 32
    x2 = rng.standard_normal((m, b, 2 * c)).astype(np.float32)
    wr = (0.1 * rng.standard_normal((m, c, c))).astype(np.float32)
    wi = (0.1 * rng.standard_normal((m, c, c))).astype(np.float32)
    g2 = rng.standard_normal((m, b, 2 * c)).astype(np.float32)
    xt = torch.tensor(x2, device="cuda", requires_grad=True)
    wrt = torch.tensor(wr, device="cuda", requires_grad=True)
    wit = torch.tensor(wi, device="cuda", requires_grad=True)
    w2t = torch.cat([torch.cat([wrt, wit], 2), torch.cat([-wit, wrt], 2)], 1)
    y2 = spectral_contract(xt, w2t)
    y2.backward(torch.tensor(g2, device="cuda"))
    xb, wb, gb = bf16(x2), bf16(stack_weights(wr, wi)), bf16(g2)
    y_ref = contract(xb, wb)
    dx_ref, dw2_ref = contract_grads(xb, wb, gb)
    dwr_ref, dwi_ref = weight_grads(dw2_ref, c)
    errs = [_rel(y2.detach().cpu().numpy(), y_ref), _rel(xt.grad.cpu().numpy(), dx_ref),
            _rel(wrt.grad.cpu().numpy(), dwr_ref), _rel(wit.grad.cpu().numpy(), dwi_ref)]
    print("forward, dX, dRe W, dIm W rel-L2 vs FP64:", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-5, errs


def test_surrogate_training_beats_baseline_and_falls_back():
    from kwave_amd import surrogate, workloads

    xtr, ytr = surrogate.make_dataset(5_000_000 + np.arange(1024))
    xte, yte = surrogate.make_dataset(6_000_000 + np.arange(128))
    model = surrogate.build_model()
    hist = surrogate.train(model, xtr, ytr, epochs=30, log=print)
    assert hist[-1] < hist[0]
    err = surrogate.rel_l2(surrogate.predict(model, xte), yte)
    base = surrogate.rel_l2(np.broadcast_to(ytr.mean(0), yte.shape), yte)
    print(f"held-out rel-L2: surrogate {err.mean():.4f}, mean-trace baseline {base.mean():.4f}")
    assert err.mean() < 0.7 * base.mean(), (err.mean(), base.mean())

    xca, yca = surrogate.make_dataset(7_000_000 + np.arange(256))
    sur = surrogate.Surrogate(model)
    cal = sur.calibrate(xca, yca, miss_rate=0.01, policy="budget")
    print("calibration:", cal)
    assert cal["miss_frac"] <= 0.01 + 1e-12 and sur.threshold is not None
    cal = sur.calibrate(xca, yca)                       # route maps estimated above 2 %
    assert cal["policy"] == "predicted" and sur.threshold == sur.max_err
    out, used_solver = sur(xte)
    err = surrogate.rel_l2(out, yte)
    assert np.all(err[used_solver] == 0.0)            # routed maps carry the solver's traces
    print(f"routed {used_solver.mean():.3f}; after routing {np.mean(err > 0.02):.3f} of maps above 2 %")
    odd = xte[:2].copy()
    odd[1, 40:50, 60:70] = workloads.SURROGATE_C_RANGE[1] + 50.0  # outside the training range
    out, used_solver = surrogate.Surrogate(model)(odd)                # no error head: range check only
    assert used_solver.tolist() == [False, True]
    assert np.array_equal(out[1], surrogate.solve_batch(odd[1:].astype(np.float64))[0])
    strict = surrogate.Surrogate(model, max_err=1e-9)
    strict.validate(xte[:4], yte[:4])
    assert not strict.enabled
    out, used_solver = strict(xte[:2])
    assert used_solver.all()


def test_training_dft_gradients_match_fp64():
    """The kernel-backed truncated DFTs of training: forward values and the gradients (the
    opposite-direction kernel with transposed matrices) against FP64 products."""
    import torch

    from kwave_amd import surrogate

    conv = surrogate.build_model().spec[0]
    mx, my, ny, nx = [np.asarray(a, dtype=np.float64) for a in surrogate.dft_mats(64, 16)]
    g = torch.Generator().manual_seed(6)
    v = torch.randn(2, 32, 64, 64, generator=g).to("cuda", torch.bfloat16).float().requires_grad_()
    x = surrogate.fno_dft_train(1, v, conv)
    gx = torch.randn(tuple(x.shape), generator=g)
    x.backward(gx.cuda())
    vv = v.detach().cpu().double().numpy()
    assert _rel(x.detach().cpu().numpy(), ((mx @ vv).reshape(2, 32, 32, 128) @ my)) < 1e-2
    want = mx.T @ (gx.double().numpy() @ my.T).reshape(2, 32, 64, 64)       # Mx^T view(G My^T)
    assert _rel(v.grad.cpu().numpy(), want) < 1e-2

    y = torch.randn(2, 32, 32, 32, generator=g).cuda().requires_grad_()
    out = surrogate.fno_dft_train(-1, y, conv)
    go = torch.randn(tuple(out.shape), generator=g).to(torch.bfloat16)
    out.backward(go.cuda())
    yy = y.detach().cpu().double().numpy()
    assert _rel(out.detach().float().cpu().numpy(), nx @ (yy @ ny).reshape(2, 32, 64, 64)) < 1e-2
    want_y = (nx.T @ go.double().numpy()).reshape(2, 32, 32, 128) @ ny.T     # view(Nx^T G) Ny^T
    assert _rel(y.grad.cpu().numpy(), want_y) < 1e-2


def test_network_gradients_with_kernel_dfts_match_fp32_products():
    """One training loss through the network: the parameter gradients with the kernel-backed
    transforms agree with those of the FP32 matrix-product transforms (bf16 rounding)."""
    import torch

    from kwave_amd import surrogate, workloads

    x = torch.as_tensor(np.stack([workloads.surrogate_c_map(5_000_000 + i) for i in range(8)]).astype(np.float32)).cuda()
    grads = []
    for fused in (True, False):
        model = surrogate.build_model(seed=0)
        for conv in model.spec:
            conv.fused_dft = fused
        with torch.no_grad():
            model.mean.normal_(generator=torch.Generator(device="cuda").manual_seed(1))
        loss = model(x).square().mean()
        loss.backward()
        grads.append({n: p.grad.detach().cpu().numpy() for n, p in model.named_parameters() if p.grad is not None})
    worst = max(_rel(grads[0][k], grads[1][k]) for k in grads[1] if np.linalg.norm(grads[1][k]) > 0)
    print(f"worst parameter-gradient rel-L2, kernel vs FP32 transforms: {worst:.2e}")
    assert worst < 3e-2


@pytest.mark.parametrize("b,c", [(40, 16), (70, 32)])
def test_fused_mode_contraction_matches_gather_path(b, c):
    """kw_mode_contract_bf16 (gather + MFMA + scatter fused) == explicit gather + kw_mode_gemm_bf16 + scatter,
    forward and gradients, bit for bit (same bf16 operands, same accumulation order).  16 retained
    modes run the tiled kernel (32 batch rows per workgroup: 40 and 70 leave ragged last tiles)."""
    import torch

    from kwave_amd.surrogate import gather_modes, spectral_contract, spectral_modes

    torch.manual_seed(3)
    n, md = 64, 16
    v = torch.randn(b, c, n, n, device="cuda")
    wr = (0.05 * torch.randn(2 * md * md, c, c, device="cuda")).requires_grad_()
    wi = (0.05 * torch.randn(2 * md * md, c, c, device="cuda")).requires_grad_()

    def w2():
        return torch.stack([torch.stack([wr, wi], -1), torch.stack([-wi, wr], -1)], 2).reshape(-1, 2 * c, 2 * c)

    ft = torch.fft.rfft2(v).requires_grad_()
    g = torch.randn(b, c, n, n // 2 + 1, dtype=torch.complex64, device="cuda")
    out = spectral_modes(ft, w2(), md)
    (torch.view_as_real(out) * torch.view_as_real(g)).sum().backward()
    fused = [out.detach().clone(), ft.grad.clone(), wr.grad.clone(), wi.grad.clone()]
    ft.grad = wr.grad = wi.grad = None

    bp = (b + 31) // 32 * 32
    x2 = torch.nn.functional.pad(gather_modes(ft, md), (0, 0, 0, bp - b))
    y2 = spectral_contract(x2, w2())[:, :b]
    y = torch.view_as_complex(y2.reshape(2 * md, md, b, c, 2)).permute(2, 3, 0, 1)
    kx = torch.cat([torch.arange(md), torch.arange(n - md, n)]).cuda()
    out_ref = torch.zeros(b, c, n, n // 2 + 1, dtype=torch.complex64, device="cuda")
    out_ref[:, :, kx, :md] = y
    (torch.view_as_real(out_ref) * torch.view_as_real(g)).sum().backward()
    assert torch.equal(fused[0], out_ref.detach())
    assert torch.equal(fused[1], ft.grad)
    for a, bb in zip(fused[2:], (wr.grad, wi.grad)):
        assert torch.allclose(a, bb, rtol=0, atol=1e-6 * float(bb.abs().max())), float((a - bb).abs().max())


@pytest.mark.parametrize("c,p", [(16, 4096), (32, 12), (16, 4 * 257)])
@pytest.mark.parametrize("gelu", [True, False])
def test_fused_pointwise_matches_fp32_reference(c, p, gelu):
    """kw_fno_pointwise_bf16 against the plain PyTorch FP32 form of the same op on the
    same bf16 inputs; the only difference allowed is the final bf16 rounding."""
    import torch
    import torch.nn.functional as F

    from kwave_amd.surrogate import fno_pointwise

    g = torch.Generator().manual_seed(c + p)
    b = 5
    v = torch.randn(b, c, p, generator=g).to("cuda", torch.bfloat16)
    s = torch.randn(b, c, p, generator=g).to("cuda", torch.bfloat16)
    pw = (torch.randn(c, c, generator=g) / c ** 0.5).cuda()
    pb = torch.randn(c, 1, generator=g).cuda()
    out = fno_pointwise(v, s, pw, pb, gelu).float()
    ref = s.float() + pb + torch.matmul(pw, v.float())
    if gelu:
        ref = F.gelu(ref)
    err = (out - ref).abs()
    # one bf16 rounding: relative 2^-8 of the value
    assert float((err - 2.0 ** -8 * ref.abs()).max()) <= 1e-6, float(err.max())
    # rounding to nearest errs by a quarter unit in the last place on average
    assert float(err.mean()) < 2.0 ** -8 * float(ref.abs().mean())


@pytest.mark.parametrize("c", [16, 32])
def test_fused_lift_matches_fp32_reference(c):
    """kw_fno_lift_bf16 == bf16 of (2 x 2 mean of the normalised map) * lc + g in FP32, to one
    bf16 rounding (the kernel fuses the multiply-add)."""
    import torch
    import torch.nn.functional as F

    from kwave_amd import surrogate

    torch.manual_seed(5)
    b, n = 37, 64
    cmap = surrogate.C_MID + 120.0 * torch.rand(b, 2 * n, 2 * n, device="cuda") - 60.0
    lc = torch.randn(c, device="cuda")
    g = torch.randn(c, n, n, device="cuda").to(torch.bfloat16)
    got = surrogate.fno_lift(cmap, lc, g).float()
    x = F.avg_pool2d(((cmap.double() - surrogate.C_MID) / surrogate.C_HALF)[:, None], 2)
    ref = x * lc.double()[None, :, None, None] + g.double()[None]
    err = (got.double() - ref).abs() / ref.abs().clamp_min(1e-3)
    assert float(err.max()) <= 1.01 * 2.0 ** -8, float(err.max())


def test_fused_inference_matches_unfused_forward():
    """The fused inference forward of the surrogate network agrees with its unfused
    (autograd) form under the same bf16 autocast to within bf16 rounding."""
    import torch

    from kwave_amd import surrogate, workloads

    model = surrogate.build_model().eval()
    x = torch.as_tensor(np.stack([workloads.surrogate_c_map(5_000_000 + i) for i in range(8)]).astype(np.float32)).cuda()
    with torch.autocast("cuda", dtype=torch.bfloat16):
        with torch.no_grad():
            fused = model(x).float()
        unfused = model(x).float().detach()
    assert _rel(fused.cpu().numpy(), unfused.cpu().numpy()) < 1e-2


def _dft_operands():
    import torch

    from kwave_amd.surrogate import dft_mats

    mx, my, ny, nx = [torch.as_tensor(np.ascontiguousarray(a)) for a in dft_mats(64, 16)]
    b16 = lambda a: a.contiguous().to("cuda", torch.bfloat16)  # noqa: E731
    return mx, my, ny, nx, b16(mx), b16(my.T), b16(ny.T), b16(nx)


def test_fused_dft_forward_matches_rfft2_and_bf16_emulation():
    """kw_fno_dft_bf16 forward: against rfft2 (FP64) at the retained modes of the same
    bf16 slabs, and against a PyTorch emulation with the same bf16 roundings."""
    import torch

    from kwave_amd.surrogate import fno_dft

    mx, my, _, _, mxb, mytb, _, _ = _dft_operands()
    g = torch.Generator().manual_seed(3)
    v = torch.randn(3, 16, 64, 64, generator=g).to("cuda", torch.bfloat16)
    x = fno_dft(1, v, mxb, mytb).cpu().double()                             # (3, 16, 32, 32)
    ft = torch.fft.rfft2(v.cpu().double())
    kx = torch.cat([torch.arange(16), torch.arange(48, 64)])
    ref = torch.view_as_real(ft.index_select(2, kx)[..., :16]).reshape(3, 16, 32, 32)
    assert _rel(x.numpy(), ref.numpy()) < 1e-2
    t = (mxb.float() @ v.float()).to(torch.bfloat16).float().reshape(3, 16, 32, 128)
    emu = (t @ mytb.float().T).cpu().double()
    assert _rel(x.numpy(), emu.numpy()) < 2e-3


def test_fused_dft_inverse_matches_irfft2_and_bf16_emulation():
    """kw_fno_dft_bf16 inverse: against irfft2 (FP64) of the zero-padded spectrum and
    against a PyTorch emulation with the same bf16 roundings."""
    import torch

    from kwave_amd.surrogate import fno_dft

    _, _, ny, nx, _, _, nytb, nxb = _dft_operands()
    g = torch.Generator().manual_seed(4)
    y = torch.randn(3, 16, 32, 32, generator=g).cuda()
    out = fno_dft(-1, y, nytb, nxb).cpu().double()                            # (3, 16, 64, 64)
    yc = torch.view_as_complex(y.cpu().double().reshape(3, 16, 32, 16, 2))
    full = torch.zeros(3, 16, 64, 33, dtype=torch.complex128)
    full[:, :, :16, :16], full[:, :, 48:, :16] = yc[:, :, :16], yc[:, :, 16:]
    ref = torch.fft.irfft2(full, s=(64, 64))
    assert _rel(out.numpy(), ref.numpy()) < 1e-2
    c = (y.to(torch.bfloat16).float() @ nytb.float().T).to(torch.bfloat16).float().reshape(3, 16, 64, 64)
    emu = (nxb.float() @ c).cpu().double()
    assert _rel(out.numpy(), emu.numpy()) < 2e-3


def test_head_gather_matches_index_select():
    """kw_fno_head_gather_bf16 moves bf16 values: bit-identical to PyTorch's gather."""
    import torch

    from kwave_amd.surrogate import head_gather

    g = torch.Generator().manual_seed(5)
    v = torch.randn(3, 16, 64, 64, generator=g).to("cuda", torch.bfloat16)
    cols = torch.tensor([0, 63, 5, 5, 17, 40, 62, 1], dtype=torch.int32)
    got = head_gather(v, cols.cuda())
    want = v.permute(0, 3, 1, 2).index_select(1, cols.long().cuda())
    assert torch.equal(got, want)
