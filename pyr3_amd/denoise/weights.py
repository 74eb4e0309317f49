"""Flat float32 parameter vector of the denoiser in the C ABI's order."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .model import LAYERS

DEFAULT = Path(__file__).resolve().parent / "weights.npz"


def flat_params(arrays) -> np.ndarray:
    """arrays: mapping name -> array (an .npz from train.py, or a state dict
    converted with .numpy()). Order: each 3x3 layer's weight then bias, then
    the head's weight and bias."""
    parts = []
    for n, ci, co, _ in LAYERS:
        w = np.asarray(arrays[f"{n}.weight"], dtype=np.float32)
        assert w.shape == (co, ci, 3, 3), (n, w.shape)
        parts += [w.reshape(-1), np.asarray(arrays[f"{n}.bias"], dtype=np.float32).reshape(-1)]
    parts += [np.asarray(arrays["out.weight"], dtype=np.float32).reshape(-1),
              np.asarray(arrays["out.bias"], dtype=np.float32).reshape(-1)]
    return np.concatenate(parts)


def from_module(net) -> np.ndarray:
    arrays = {f"{n}.{p}": getattr(m, p).detach().float().cpu().numpy()
              for n, m in net.convs.items() for p in ("weight", "bias")}
    arrays["out.weight"] = net.out.weight.detach().float().cpu().numpy()
    arrays["out.bias"] = net.out.bias.detach().float().cpu().numpy()
    return flat_params(arrays)


def load_module(cls, path=DEFAULT):
    """An eval-mode network of class `cls` holding the weights of `path`."""
    import torch
    arrays = np.load(path, allow_pickle=False)
    net = cls()
    with torch.no_grad():
        for n, m in net.convs.items():
            m.weight.copy_(torch.from_numpy(np.asarray(arrays[f"{n}.weight"], np.float32)))
            m.bias.copy_(torch.from_numpy(np.asarray(arrays[f"{n}.bias"], np.float32)))
        net.out.weight.copy_(torch.from_numpy(np.asarray(arrays["out.weight"], np.float32)))
        net.out.bias.copy_(torch.from_numpy(np.asarray(arrays["out.bias"], np.float32)))
    return net.eval()
