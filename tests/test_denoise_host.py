"""Denoiser host side (CPU): the network definition, the parameter layout
shared by the trainer's .npz, the C ABI's flat vector and the PyTorch
restatement."""
import numpy as np
import pytest
import torch

from pyr3_amd.denoise import net
from pyr3_amd.denoise.weights import DEFAULT, flat_params, load


def test_param_count_matches_c_abi():
    from pyr3_amd import _lib
    p = net.init_params(0)
    assert _lib.load().p3_denoiser_param_count() == net.param_count() == flat_params(p).size
    assert 480_000 < net.param_count() < 500_000     # ~0.5 M parameters


def test_flat_order_is_layer_weight_then_bias():
    p = net.init_params(3)
    flat = flat_params(p)
    off = 0
    for L in net.LAYERS:
        w = net.to_torch_weight(p[L.name + ".w"], L)
        assert w.shape == (L.cout, L.cin, 3, 3)
        # tap t = 3 (dy + 1) + (dx + 1) of input channel c is column t * cin + c
        assert w[1, L.cin - 1, 2, 0] == p[L.name + ".w"][1, 6 * L.cin + L.cin - 1]
        assert np.array_equal(flat[off:off + w.numel()], w.numpy().reshape(-1))
        off += w.numel()
        assert np.array_equal(flat[off:off + L.cout], p[L.name + ".b"])
        off += L.cout
    assert off + net.HEAD == flat.size


def test_network_shapes_levels_and_cost():
    p = net.init_params(1)
    for h, w in ((67, 93), (64, 64), (9, 3)):
        out = net.torch_forward(p, torch.zeros(1, net.FEAT_CH, h, w), torch.zeros(1, 3, h, w))
        assert out.shape == (1, 3, h, w)
    assert net.level_sizes(1080, 1920) == [(1080, 1920), (540, 960), (270, 480), (135, 240)]
    assert net.level_sizes(67, 93)[3] == (9, 12)
    assert 40_000 < net.macs_per_pixel() < 42_000


def test_zero_head_is_the_base_image():
    p = net.init_params(2)      # the head starts at zero: output = base RGB
    base = torch.rand(1, 3, 20, 30)
    assert torch.equal(net.torch_forward(p, torch.rand(1, net.FEAT_CH, 20, 30), base), base)


def test_padding_columns_are_zero():
    p = net.init_params(4)
    for L in net.LAYERS:
        assert not p[L.name + ".w"][:, 9 * L.cin:].any()


def test_shipped_weights_load():
    if not DEFAULT.exists():
        pytest.skip("no trained weights in the tree")
    from pyr3_amd import _lib
    p = load()
    flat = flat_params(p)
    assert flat.size == _lib.load().p3_denoiser_param_count()
    assert np.isfinite(flat).all() and np.abs(flat).max() < 100
    for L in net.LAYERS:
        assert not p[L.name + ".w"][:, 9 * L.cin:].any()


def test_hipnet_refuses_a_library_loaded_before_torch(monkeypatch):
    """torch's own HIP runtime cannot open the GPU once the native library has
    bound /opt/rocm's: HipNet says so instead of failing inside torch."""
    from pyr3_amd import _lib
    from pyr3_amd.denoise import net
    monkeypatch.setattr(_lib, "_lib", object())
    monkeypatch.setattr(_lib, "loaded_after_torch", False)
    with pytest.raises(RuntimeError, match="loaded before torch"):
        net.HipNet(net.init_params(0))
