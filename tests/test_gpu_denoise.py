"""BF16-MFMA denoiser inference on the GPU vs the network in PyTorch FP32.

The renderer presents a low-quality frame without DE into its f32 image;
the HIP kernels compute the network input from the histogram and that image,
run the fourteen 3x3 layers with BF16 activations and weights (FP32
accumulation) and the 1x1 head, and overwrite the RGBA8 image. The reference
is the same network (pyr3_amd/denoise/net.py) in FP32 on the CPU from the
features the training kernels compute (p3nn_features) of the same frame.
Tolerance (BF16 rounding of every activation): RGB within 3 LSB everywhere,
mean absolute difference below 0.35 LSB; alpha identical.
"""
import numpy as np
import pytest
import torch

from helpers import genome, sierpinski, xform

pytestmark = pytest.mark.gpu


def _params(seed):
    from pyr3_amd.denoise import net
    p = net.init_params(seed)
    rng = np.random.default_rng(seed)
    p["head"] = (rng.standard_normal(net.HEAD) * 0.05).astype(np.float32)
    for L in net.LAYERS:
        p[L.name + ".b"] = (rng.standard_normal(L.cout) * 0.01).astype(np.float32)
    return p


def _frame(W, H):
    from pyr3_amd.renderer import Renderer
    g = genome([xform(0.6, (0.6, 0.3, 0.1, -0.2, 0.7, 0.05), vars=((1, 0.5), (2, 0.5)), color=0.1),
                xform(0.4, (-0.5, 0.2, -0.3, 0.4, 0.5, 0.2), vars=((0, 0.7), (1, 0.3)), color=0.9)],
               scale=30.0)
    r = Renderer(W, H, 1)
    r.set_image_f32(True)
    r.reset(g)
    s = r.iterate(g, 5, 2048, 40, fuse=20)
    r.present(g, s, force_de_off=True)
    return r, g, s


def reference_output(params, r, g, s):
    """FP32 network output (H, W, 3) of the renderer's last DE-off present."""
    from pyr3_amd import pack
    from pyr3_amd.denoise import net
    hist = r.read_histogram()
    img = r.read_image_f32()
    _, k2 = pack.calibration(g.scale, s, g.tonemap.brightness, 1)
    hn = net.HipNet(params)
    feat = hn.features(torch.from_numpy(hist.view(np.int32)).cuda(),
                       torch.from_numpy(img).cuda(), k2)
    torch.cuda.synchronize()
    f = feat.float().cpu().permute(0, 3, 1, 2)
    base = torch.from_numpy(img[..., :3].transpose(2, 0, 1).copy())[None]
    with torch.no_grad():
        out = net.torch_forward(params, f, base)[0].clamp(0, 1).numpy().transpose(1, 2, 0)
    return out, img


@pytest.mark.parametrize("size", [(96, 80), (131, 67), (200, 37)])
def test_denoiser_matches_fp32_reference(size):
    from pyr3_amd.denoise.weights import flat_params
    from pyr3_amd.renderer import Denoiser
    W, H = size
    r, g, s = _frame(W, H)
    p = _params(3)
    ref, img = reference_output(p, r, g, s)
    d = Denoiser(flat_params(p))
    d.apply(r)
    got = r.read_image()
    want = np.floor(ref * 255.0 + 0.5).astype(np.int32)
    diff = np.abs(got[..., :3].astype(np.int32) - want)
    print(f"{W}x{H}: max diff {diff.max()} LSB, mean {diff.mean():.3f}", flush=True)
    assert diff.max() <= 3 and diff.mean() < 0.35, (diff.max(), diff.mean())
    alpha = np.floor(np.clip(img[..., 3], 0, 1) * 255.0 + 0.5)
    assert np.array_equal(got[..., 3].astype(np.float64), alpha)
    d.destroy()
    r.destroy()


def test_denoise_requires_a_de_off_f32_present():
    from pyr3_amd._lib import P3Error
    from pyr3_amd.denoise.weights import flat_params
    from pyr3_amd.renderer import Denoiser, Renderer
    g = genome(sierpinski(), scale=20.0)
    r = Renderer(32, 32, 1)
    r.reset(g)
    s = r.iterate(g, 1, 256, 10)
    r.present(g, s, force_de_off=True)
    d = Denoiser(flat_params(_params(0)))
    with pytest.raises(P3Error):      # no f32 image
        d.apply(r)
    r.set_image_f32(True)
    r.present(g, s)                   # with DE
    with pytest.raises(P3Error):
        d.apply(r)
    r.present(g, s, force_de_off=True)
    d.apply(r)
    d.destroy()
    r.destroy()
