"""The denoiser's hand-written training kernels (include/pyr3nn.h) against
PyTorch FP32 references computed on the CPU: the BF16-MFMA GEMM in every
operand layout and epilogue, im2col / col2im (padding, stride 2, upsample,
concat), the features, one full forward/backward of the 4-level U-Net
against autograd, and Adam."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bf(x):
    """float32 tensor rounded to bf16 values (as float32)."""
    return x.to(torch.bfloat16).to(torch.float32)


@pytest.fixture(scope="module")
def lib():
    from pyr3_amd import _lib
    return _lib.load()


def _stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("ak,bk", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_layouts_and_split_k(lib, ak, bk):
    from pyr3_amd.denoise.net import _Gemm, EPI_STORE, EPI_ATOMIC
    g = torch.Generator().manual_seed(7 + 2 * ak + bk)
    M, N, K = 77, 45, 168
    A = _bf(torch.randn(M, K, generator=g))
    B = _bf(torch.randn(N, K, generator=g))
    ref = A.double() @ B.double().T

    def stored(X, kmajor):
        """X (rows x K) as the kernel reads it: [rows][K] or [K][rows], the
        leading dimension padded to a multiple of 8 (rows 77, 45)."""
        Y = X.T if kmajor else X
        ld = (Y.shape[1] + 7) // 8 * 8
        Z = torch.zeros(Y.shape[0], ld)
        Z[:, :Y.shape[1]] = Y
        return Z.to(torch.bfloat16).cuda(), ld
    Ad, lda = stored(A, ak)
    Bd, ldb = stored(B, bk)
    for epi, split in ((EPI_STORE, 1), (EPI_ATOMIC, 3)):
        Cd = torch.zeros(M, N, dtype=torch.float32, device="cuda")
        gm = _Gemm(Ad.data_ptr(), Bd.data_ptr(), Cd.data_ptr(), None, M, N, K, lda, ldb, N)
        assert lib.p3nn_gemm(C.byref(gm), ak, bk, epi, split, _stream()) == 0
        torch.cuda.synchronize()
        got = Cd.cpu().double()
        err = (got - ref).abs().max().item() / ref.abs().max().item()
        print(f"layout A{ak} B{bk} epi {epi} split {split}: max rel err {err:.2e}", flush=True)
        assert err < 1e-5


def test_gemm_bias_relu_epilogue(lib):
    from pyr3_amd.denoise.net import _Gemm, EPI_BIAS_RELU_BF16
    g = torch.Generator().manual_seed(3)
    M, N, K = 130, 64, 96
    A = _bf(torch.randn(M, K, generator=g))
    B = _bf(torch.randn(N, K, generator=g))
    bias = torch.randn(N, generator=g)
    ref = torch.relu(A.double() @ B.double().T + bias.double()).float()
    Cd = torch.zeros(M, N, dtype=torch.bfloat16, device="cuda")
    Ad, Bd, bd = A.to(torch.bfloat16).cuda(), B.to(torch.bfloat16).cuda(), bias.cuda()
    gm = _Gemm(Ad.data_ptr(), Bd.data_ptr(), Cd.data_ptr(), bd.data_ptr(), M, N, K, K, K, N)
    assert lib.p3nn_gemm(C.byref(gm), 0, 0, EPI_BIAS_RELU_BF16, 1, _stream()) == 0
    torch.cuda.synchronize()
    got = Cd.float().cpu()
    assert torch.allclose(got, _bf(ref), rtol=8e-3, atol=1e-6)


def _im2col_ref(a, b, stride, up, out_h, out_w):
    """(n, H, W, C) inputs -> (n*out_h*out_w, 9*Cin) by explicit taps."""
    n = a.shape[0]
    if up:
        a = a.repeat_interleave(2, 1).repeat_interleave(2, 2)[:, :out_h, :out_w]
    x = torch.cat([a, b], 3) if b is not None else a
    H, W, Cin = x.shape[1:]
    xp = torch.zeros(n, H + 2, W + 2, Cin)
    xp[:, 1:H + 1, 1:W + 1] = x
    cols = []
    s = 1 if up else stride
    for t in range(9):
        dy, dx = t // 3, t % 3
        cols.append(xp[:, dy:dy + s * (out_h - 1) + 1:s, dx:dx + s * (out_w - 1) + 1:s])
    return torch.cat(cols, 3).reshape(n * out_h * out_w, 9 * Cin)


@pytest.mark.parametrize("case", ["plain", "stride2", "up", "concat", "feat8"])
def test_im2col_and_col2im(lib, case):
    from pyr3_amd.denoise.net import _Conv
    g = torch.Generator().manual_seed(11)
    n = 2
    if case == "plain":
        ih, iw, oh, ow, ca, cb, st, up = 13, 9, 13, 9, 16, 0, 1, 0
    elif case == "stride2":
        ih, iw, oh, ow, ca, cb, st, up = 13, 9, 7, 5, 32, 0, 2, 0
    elif case == "up":
        ih, iw, oh, ow, ca, cb, st, up = 7, 5, 13, 9, 64, 0, 1, 1
    elif case == "concat":
        ih, iw, oh, ow, ca, cb, st, up = 11, 6, 11, 6, 16, 16, 1, 0
    else:
        ih, iw, oh, ow, ca, cb, st, up = 10, 12, 10, 12, 8, 0, 1, 0
    a = _bf(torch.randn(n, ih, iw, ca, generator=g)).requires_grad_(True)
    b = _bf(torch.randn(n, oh, ow, cb, generator=g)).requires_grad_(True) if cb else None
    ref = _im2col_ref(a, b, st, up, oh, ow)
    kp = (9 * (ca + cb) + 31) // 32 * 32
    cv = _Conv(n, iw, ih, ow, oh, ca, cb, st, up)
    ad = a.detach().to(torch.bfloat16).cuda()
    bd = b.detach().to(torch.bfloat16).cuda() if cb else None
    col = torch.empty(n * oh * ow, kp, dtype=torch.bfloat16, device="cuda")
    assert lib.p3nn_im2col(C.byref(cv), ad.data_ptr(), bd.data_ptr() if cb else None,
                           col.data_ptr(), kp, _stream()) == 0
    torch.cuda.synchronize()
    got = col.float().cpu()
    assert torch.equal(got[:, :ref.shape[1]], ref.detach())
    assert not got[:, ref.shape[1]:].any()
    # col2im = the gradient of im2col
    dcol = torch.randn(n * oh * ow, kp, generator=g)
    (ref * dcol[:, :ref.shape[1]]).sum().backward()
    da = torch.zeros(n, ih, iw, ca, device="cuda")
    db = torch.zeros(n, oh, ow, cb, device="cuda") if cb else None
    dcd = dcol.cuda()
    assert lib.p3nn_col2im(C.byref(cv), dcd.data_ptr(), kp, da.data_ptr(),
                           db.data_ptr() if cb else None, _stream()) == 0
    torch.cuda.synchronize()
    assert torch.allclose(da.cpu(), a.grad, rtol=1e-5, atol=1e-5)
    if cb:
        assert torch.allclose(db.cpu(), b.grad, rtol=1e-5, atol=1e-5)


def test_features_kernel(lib):
    g = torch.Generator().manual_seed(5)
    H, W = 9, 14
    hist = torch.randint(0, 3000, (H, W, 4), generator=g, dtype=torch.int64)
    hist[::3, ::2] = 0
    img = torch.rand(H, W, 4, generator=g)
    k2 = 3.7e-4
    hd = hist.to(torch.int32).cuda()
    out = torch.empty(1, H, W, 8, dtype=torch.bfloat16, device="cuda")
    assert lib.p3nn_features(hd.data_ptr(), img.cuda().data_ptr(), k2, H * W, out.data_ptr(),
                             _stream()) == 0
    torch.cuda.synchronize()
    h = hist.double()
    cnt = h[..., 3]
    inv = torch.where(cnt > 0, 1.0 / cnt.clamp(min=1), torch.zeros_like(cnt))
    ref = torch.stack([img[..., 0].double(), img[..., 1].double(), img[..., 2].double(),
                       torch.log1p(16 * cnt * k2) / 4, h[..., 0] * inv, h[..., 1] * inv,
                       h[..., 2] * inv, (cnt > 0).double()], -1)
    got = out[0].float().cpu().double()
    assert torch.allclose(got, ref, rtol=8e-3, atol=1e-6)


class _Bf16(torch.autograd.Function):
    """Rounds to bf16 in the forward and the gradient to bf16 in the backward:
    the HIP path stores activations in bf16 and feeds its GEMMs bf16 output
    gradients (p3nn_relu_bwd)."""

    @staticmethod
    def forward(ctx, x):
        return _bf(x)

    @staticmethod
    def backward(ctx, g):
        return _bf(g)


def _torch_loss_and_grads(params, feat, base, target):
    """Autograd of the FP32 restatement with the HIP path's bf16 roundings
    (weights and inputs bf16-exact; each layer's output and its gradient
    rounded to bf16; relu(round(z)) = round(relu(z)))."""
    from pyr3_amd.denoise import net
    import torch.nn.functional as F
    tp = {k: torch.from_numpy(v.copy()).requires_grad_(True) for k, v in params.items()}
    acts = {"feat": feat}
    for L in net.LAYERS:
        a = acts[L.a]
        if L.up:
            ref = acts[net.SKIP_OF_LEVEL[L.level]]
            a = F.interpolate(a, scale_factor=2, mode="nearest")[..., :ref.shape[-2],
                                                                 :ref.shape[-1]]
        x = torch.cat([a, acts[L.b]], 1) if L.b else a
        w = tp[L.name + ".w"][:, :9 * L.cin].reshape(L.cout, 3, 3, L.cin).permute(0, 3, 1, 2)
        z = F.conv2d(x, w, tp[L.name + ".b"], stride=L.stride, padding=1)
        acts[L.name] = F.relu(_Bf16.apply(z))
    hw = tp["head"]
    out = base + F.conv2d(acts["c1"], hw[:48].reshape(3, 16, 1, 1), hw[48:51])
    loss = ((out - target) ** 2).mean()
    loss.backward()
    return loss.item(), {k: v.grad for k, v in tp.items()}


def test_network_forward_backward_matches_autograd():
    from pyr3_amd.denoise import net
    g = torch.Generator().manual_seed(2)
    params = net.init_params(4)
    for k, v in params.items():   # a non-trivial head and biases, bf16-exact weights
        if k == "head" or k.endswith(".b"):
            params[k] = (0.05 * np.random.default_rng(len(k)).standard_normal(v.shape)).astype(
                np.float32)
        else:
            params[k] = _bf(torch.from_numpy(v)).numpy()
    n, H, W = 2, 37, 29
    feat = _bf(torch.rand(n, H, W, 8, generator=g))
    base = torch.zeros(n, H, W, 4)
    base[..., :3] = torch.rand(n, H, W, 3, generator=g)
    target = torch.rand(n, H, W, 3, generator=g)
    hn = net.HipNet(params)
    out = hn.forward(feat.to(torch.bfloat16).cuda(), base.cuda())
    loss = hn.loss_backward(target.cuda())
    torch.cuda.synchronize()
    ref_out = net.torch_forward(params, feat.permute(0, 3, 1, 2), base[..., :3].permute(0, 3, 1, 2),
                                round_bf16=True).permute(0, 2, 3, 1)
    err = (out.cpu() - ref_out).abs().max().item()
    print(f"forward max abs err {err:.3e}", flush=True)
    assert err < 2e-2
    ref_loss, ref_g = _torch_loss_and_grads(params, feat.permute(0, 3, 1, 2),
                                            base[..., :3].permute(0, 3, 1, 2),
                                            target.permute(0, 3, 1, 2))
    assert abs(loss.item() - ref_loss) < 2e-2 * ref_loss
    for k, gr in ref_g.items():
        got = hn.g[k].cpu()
        rel = (got - gr).norm().item() / max(gr.norm().item(), 1e-12)
        print(f"grad {k}: rel L2 err {rel:.3e}", flush=True)
        assert rel < 1e-2, k


def test_adam_step_matches_the_formula(lib):
    g = torch.Generator().manual_seed(9)
    n = 1000
    w = torch.randn(n, generator=g)
    gr = torch.randn(n, generator=g)
    m = torch.randn(n, generator=g) * 0.1
    v = torch.rand(n, generator=g) * 0.01
    wd, gd, md, vd = w.cuda(), gr.cuda(), m.cuda(), v.cuda()
    wbf = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    lr, b1, b2, eps, step = 1e-3, 0.9, 0.999, 1e-8, 5
    assert lib.p3nn_adam(wd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), n, lr, b1, b2,
                         eps, step, wbf.data_ptr(), _stream()) == 0
    torch.cuda.synchronize()
    m2 = b1 * m + (1 - b1) * gr
    v2 = b2 * v + (1 - b2) * gr * gr
    w2 = w - lr * (m2 / (1 - b1 ** step)) / ((v2 / (1 - b2 ** step)).sqrt() + eps)
    assert torch.allclose(wd.cpu(), w2, rtol=1e-5, atol=1e-6)
    assert torch.equal(wbf.cpu(), w2.to(torch.bfloat16))


def test_training_reduces_the_loss():
    """A few Adam steps on one batch lower the loss (the kernels compose into
    a working optimiser)."""
    from pyr3_amd.denoise import net
    g = torch.Generator().manual_seed(8)
    n, H, W = 2, 32, 32
    feat = torch.rand(n, H, W, 8, generator=g)
    base = torch.zeros(n, H, W, 4)
    base[..., :3] = feat[..., 4:7]
    target = (feat[..., 4:7] * 0.5 + 0.25).contiguous()
    hn = net.HipNet(net.init_params(1))
    fd, bd, td = feat.to(torch.bfloat16).cuda(), base.cuda(), target.cuda()
    losses = []
    # split-K weight gradients add with atomics, so the trajectory varies a
    # little from run to run; 60 steps leave a wide margin under the bar
    for _ in range(60):
        hn.forward(fd, bd)
        losses.append(hn.loss_backward(td).item())
        hn.adam_step(3e-3)
    print("losses", [round(x, 5) for x in losses[::10]], flush=True)
    assert losses[-1] < 0.6 * losses[0]
