"""BF16-MFMA denoiser on the GPU vs the network in PyTorch FP32.

The HIP kernels read the renderer's f32 image and histogram count, run the
ten 3x3 layers with BF16 activations and weights (FP32 accumulation) and the
1x1 head, and overwrite the RGBA8 image. The reference is the same network
(pyr3_amd/denoise/model.py) in FP32 on the CPU from the same inputs.
Tolerance (BF16 rounding of every activation): RGB within 3 LSB everywhere,
mean absolute difference below 0.35 LSB; alpha identical.
"""
import numpy as np
import pytest
import torch

from helpers import genome, sierpinski, xform

pytestmark = pytest.mark.gpu


def _net(seed):
    from pyr3_amd.denoise.model import Denoiser
    torch.manual_seed(seed)
    net = Denoiser()
    torch.nn.init.normal_(net.out.weight, std=0.05)
    torch.nn.init.normal_(net.out.bias, std=0.01)
    return net.eval()


@pytest.mark.parametrize("size", [(96, 80), (131, 67)])
def test_denoiser_matches_fp32_reference(size):
    from pyr3_amd.denoise.model import features
    from pyr3_amd.denoise.weights import from_module
    from pyr3_amd.renderer import Denoiser, Renderer
    W, H = size
    g = genome([xform(0.6, (0.6, 0.3, 0.1, -0.2, 0.7, 0.05), vars=((1, 0.5), (2, 0.5)), color=0.1),
                xform(0.4, (-0.5, 0.2, -0.3, 0.4, 0.5, 0.2), vars=((0, 0.7), (1, 0.3)), color=0.9)],
               scale=30.0)
    r = Renderer(W, H, 1)
    r.set_image_f32(True)
    r.reset(g)
    s = r.iterate(g, 5, 2048, 40, fuse=20)
    r.present(g, s)
    img = r.read_image_f32()
    cnt = r.read_histogram()[..., 3].astype(np.float32)
    net = _net(3)
    d = Denoiser(from_module(net))
    d.apply(r)
    got = r.read_image()
    x = features(torch.from_numpy(img[..., :3].transpose(2, 0, 1).copy())[None],
                 torch.from_numpy(cnt)[None, None])
    with torch.no_grad():
        ref = net(x)[0].clamp(0, 1).numpy().transpose(1, 2, 0)
    want = np.floor(ref * 255.0 + 0.5).astype(np.int32)
    diff = np.abs(got[..., :3].astype(np.int32) - want)
    assert diff.max() <= 3 and diff.mean() < 0.35, (diff.max(), diff.mean())
    alpha = np.floor(np.clip(img[..., 3], 0, 1) * 255.0 + 0.5)
    assert np.array_equal(got[..., 3].astype(np.float64), alpha)
    d.destroy()
    r.destroy()


def test_denoise_requires_f32_image():
    from pyr3_amd._lib import P3Error
    from pyr3_amd.denoise.weights import from_module
    from pyr3_amd.renderer import Denoiser, Renderer
    g = genome(sierpinski(), scale=20.0)
    r = Renderer(32, 32, 1)
    r.reset(g)
    s = r.iterate(g, 1, 256, 10)
    r.present(g, s)
    d = Denoiser(from_module(_net(0)))
    with pytest.raises(P3Error):
        d.apply(r)
