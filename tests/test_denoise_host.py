"""Denoiser host side (CPU): parameter layout shared by the network
definition, the weight flattener and the C ABI."""
import numpy as np
import torch

from pyr3_amd.denoise.model import LAYERS, Denoiser, macs_per_pixel
from pyr3_amd.denoise.weights import flat_params, from_module


def test_param_count_matches_c_abi():
    from pyr3_amd import _lib
    net = Denoiser()
    n = sum(p.numel() for p in net.parameters())
    assert _lib.load().p3_denoiser_param_count() == n == from_module(net).size


def test_flat_order_is_layer_weight_then_bias():
    torch.manual_seed(0)
    net = Denoiser()
    flat = from_module(net)
    off = 0
    for name, ci, co, _ in LAYERS:
        w = net.convs[name].weight.detach().numpy().reshape(-1)
        assert np.array_equal(flat[off:off + w.size], w)
        off += w.size
        b = net.convs[name].bias.detach().numpy()
        assert np.array_equal(flat[off:off + co], b)
        off += co
    assert off + 48 + 3 == flat.size


def test_network_shapes_and_cost():
    net = Denoiser()
    for h, w in ((67, 93), (64, 64), (1, 3)):
        assert net(torch.zeros(1, 4, h, w)).shape == (1, 3, h, w)
    assert 25_000 < macs_per_pixel() < 32_000


def test_zero_head_is_identity():
    net = Denoiser()    # the head starts at zero: output = input RGB
    x = torch.rand(1, 4, 20, 30)
    assert torch.equal(net(x), x[:, :3])


def test_shipped_weights_load():
    from pathlib import Path
    from pyr3_amd import _lib
    path = Path(__file__).resolve().parents[1] / "pyr3_amd" / "denoise" / "weights.npz"
    z = np.load(path, allow_pickle=False)
    flat = flat_params(z)
    assert flat.size == _lib.load().p3_denoiser_param_count()
    assert np.isfinite(flat).all() and np.abs(flat).max() < 100
    net = Denoiser()
    net.load_state_dict({**{f"convs.{n}.{p}": torch.from_numpy(z[f"{n}.{p}"])
                            for n, _, _, _ in LAYERS for p in ("weight", "bias")},
                         "out.weight": torch.from_numpy(z["out.weight"]),
                         "out.bias": torch.from_numpy(z["out.bias"])})
