"""Denoiser network definition (PyTorch; training and the FP32 reference).

Input per output pixel: the tonemapped RGB of the low-quality render (the
pre-quantisation f32 image) and a density feature log1p(count / 255) / 4 of
its histogram count channel. Output: RGB = input RGB + a predicted residual,
clamped to [0, 1] by the caller.

Layers (3x3 convolutions, ReLU after each except the last):
  full  e0: 4 -> 16, e1: 16 -> 16                       (skip s0)
  half  d0: 16 -> 32 stride 2, d1: 32 -> 32              (skip s1)
  qtr   b0: 32 -> 64 stride 2, b1: 64 -> 64
  half  u0: up2(64) -> 32, u1: cat(u0, s1) 64 -> 32
  full  v0: up2(32) -> 16, v1: cat(v0, s0) 32 -> 16
  full  out: 1x1 16 -> 3 (residual)
Odd sizes: stride-2 convolutions use padding 1 (output ceil(n/2)); the
nearest-neighbour upsample is cropped to the skip's size.
"""
from __future__ import annotations

import torch
import torch.nn as nn
import torch.nn.functional as F

LAYERS = [  # name, c_in, c_out, stride
    ("e0", 4, 16, 1), ("e1", 16, 16, 1),
    ("d0", 16, 32, 2), ("d1", 32, 32, 1),
    ("b0", 32, 64, 2), ("b1", 64, 64, 1),
    ("u0", 64, 32, 1), ("u1", 64, 32, 1),
    ("v0", 32, 16, 1), ("v1", 32, 16, 1),
]


class Denoiser(nn.Module):
    def __init__(self):
        super().__init__()
        self.convs = nn.ModuleDict({n: nn.Conv2d(ci, co, 3, stride=s, padding=1)
                                    for n, ci, co, s in LAYERS})
        self.out = nn.Conv2d(16, 3, 1)
        nn.init.zeros_(self.out.weight)
        nn.init.zeros_(self.out.bias)

    @staticmethod
    def _up(x, like):
        y = F.interpolate(x, scale_factor=2, mode="nearest")
        return y[..., :like.shape[-2], :like.shape[-1]]

    def forward(self, x):
        c = self.convs
        r = F.relu
        s0 = r(c["e1"](r(c["e0"](x))))
        s1 = r(c["d1"](r(c["d0"](s0))))
        b = r(c["b1"](r(c["b0"](s1))))
        u = r(c["u0"](self._up(b, s1)))
        u = r(c["u1"](torch.cat([u, s1], 1)))
        v = r(c["v0"](self._up(u, s0)))
        v = r(c["v1"](torch.cat([v, s0], 1)))
        return x[:, :3] + self.out(v)


def features(rgb: torch.Tensor, count: torch.Tensor) -> torch.Tensor:
    """(N, 4, H, W) network input from RGB in [0, 1] (N, 3, H, W) and the
    histogram count channel (N, 1, H, W, units of 1/255 sample)."""
    return torch.cat([rgb, torch.log1p(count / 255.0) / 4.0], 1)


def macs_per_pixel() -> float:
    """Multiply-adds per output pixel (full-resolution pixel count)."""
    scale = {"e": 1.0, "d": 0.25, "b": 1 / 16, "u": 0.25, "v": 1.0}
    total = sum(ci * co * 9 * scale[n[0]] for n, ci, co, _ in LAYERS)
    # d0 / b0 run at the output (lower) resolution of their stride
    total += 16 * 3
    return total
