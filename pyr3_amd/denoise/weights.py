"""Denoiser parameters: the .npz the trainer writes and the flat float32
vector of the C ABI (p3_denoiser_create)."""
from __future__ import annotations

from pathlib import Path
from typing import Dict

import numpy as np

from .net import HEAD, LAYERS, to_torch_weight

DEFAULT = Path(__file__).resolve().parent / "weights.npz"


def load(path=DEFAULT) -> Dict[str, np.ndarray]:
    """name -> float32 array (W[o][kp] per layer as `<name>.w`, `<name>.b`,
    `head`), read without unpickling."""
    with np.load(path, allow_pickle=False) as z:
        return {k: np.asarray(z[k], np.float32) for k in z.files if k != "meta"}


def flat_params(params: Dict[str, np.ndarray]) -> np.ndarray:
    """The C ABI's order: per 3x3 layer its weight [c_out][c_in][3][3] then
    its bias, then the head's weight [3][16] and bias."""
    parts = []
    for L in LAYERS:
        w = np.asarray(params[L.name + ".w"], np.float32)
        assert w.shape == (L.cout, L.kp), (L.name, w.shape)
        parts += [to_torch_weight(w, L).numpy().reshape(-1),
                  np.asarray(params[L.name + ".b"], np.float32).reshape(-1)]
    head = np.asarray(params["head"], np.float32).reshape(-1)
    assert head.size == HEAD
    parts.append(head)
    return np.concatenate(parts)
