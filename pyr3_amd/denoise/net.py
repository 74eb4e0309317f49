"""The learned denoiser's network: a 4-level U-Net over histogram features.

Input per pixel (8 bf16 channels, `p3nn_features`): the DE-off tonemapped
RGB of the low-quality frame (the base), a quality-independent log density
log1p(16 count k2) / 4, the mean splat colour and occupancy. Output: RGB = base + a 1x1 head over the last layer, clamped by
the caller. The target is the plain path's high-quality DE + tonemap image,
so the network learns the density estimation and the noise removal.

Layers (3x3, padding 1, ReLU; stride-2 layers output ceil(n / 2); "up"
layers convolve the nearest 2x upsample of their input cropped to the skip's
size; "cat" layers take (previous, skip) concatenated):

  full     e0 8->16, e1 16->16                     skip s0 = e1
  half     d1 16->32 /2, e2 32->32                 skip s1 = e2
  quarter  d2 32->64 /2, e3 64->64                 skip s2 = e3
  eighth   d3 64->128 /2, b0 128->128
  quarter  u3 up(b0) 128->64, c3 cat(u3, s2) 128->64
  half     u2 up(c3) 64->32,  c2 cat(u2, s1) 64->32
  full     u1 up(c2) 32->16,  c1 cat(u1, s0) 32->16, head 1x1 16->3

487,296 3x3 weights + 1,024 biases + 51 head parameters (488,371). Weights are stored W[o][k], k = tap * Cin + c
(tap = 3 (dy + 1) + (dx + 1)), rows zero-padded to kp = 9 Cin rounded up to
32: the layout of the im2col GEMMs (pyr3nn.h) and of the inference kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np

FEAT_CH = 8       # network input channels


@dataclass(frozen=True)
class Layer:
    name: str
    a: str             # input tensor name ("feat" or a layer)
    b: Optional[str]   # concatenated second input (a skip) or None
    ca: int
    cb: int
    cout: int
    stride: int
    up: bool
    level: int         # output resolution level: 0 full, 1 half, 2 quarter, 3 eighth

    @property
    def cin(self) -> int:
        return self.ca + self.cb

    @property
    def kp(self) -> int:
        return (9 * self.cin + 31) // 32 * 32


LAYERS: List[Layer] = [
    Layer("e0", "feat", None, 8, 0, 16, 1, False, 0),
    Layer("e1", "e0", None, 16, 0, 16, 1, False, 0),
    Layer("d1", "e1", None, 16, 0, 32, 2, False, 1),
    Layer("e2", "d1", None, 32, 0, 32, 1, False, 1),
    Layer("d2", "e2", None, 32, 0, 64, 2, False, 2),
    Layer("e3", "d2", None, 64, 0, 64, 1, False, 2),
    Layer("d3", "e3", None, 64, 0, 128, 2, False, 3),
    Layer("b0", "d3", None, 128, 0, 128, 1, False, 3),
    Layer("u3", "b0", None, 128, 0, 64, 1, True, 2),
    Layer("c3", "u3", "e3", 64, 64, 64, 1, False, 2),
    Layer("u2", "c3", None, 64, 0, 32, 1, True, 1),
    Layer("c2", "u2", "e2", 32, 32, 32, 1, False, 1),
    Layer("u1", "c2", None, 32, 0, 16, 1, True, 0),
    Layer("c1", "u1", "e1", 16, 16, 16, 1, False, 0),
]
HEAD = 3 * 16 + 3
SKIP_OF_LEVEL = {0: "e1", 1: "e2", 2: "e3"}


def level_sizes(h: int, w: int) -> List[Tuple[int, int]]:
    out = [(h, w)]
    for _ in range(3):
        ph, pw = out[-1]
        out.append(((ph + 1) // 2, (pw + 1) // 2))
    return out


def param_count() -> int:
    return sum(9 * L.cin * L.cout + L.cout for L in LAYERS) + HEAD


def macs_per_pixel() -> float:
    """Multiply-adds per full-resolution output pixel (3x3 layers at their
    resolution, plus the head)."""
    return sum(9 * L.cin * L.cout / 4.0 ** L.level for L in LAYERS) + 48


def init_params(seed: int = 0) -> Dict[str, np.ndarray]:
    """He-initialised weights in the W[o][kp] layout (padding columns 0),
    zero biases, zero head (the network starts as the identity on the base
    image)."""
    rng = np.random.default_rng(seed)
    p = {}
    for L in LAYERS:
        w = np.zeros((L.cout, L.kp), np.float32)
        w[:, :9 * L.cin] = rng.normal(0, np.sqrt(2.0 / (9 * L.cin)), (L.cout, 9 * L.cin))
        p[L.name + ".w"] = w
        p[L.name + ".b"] = np.zeros(L.cout, np.float32)
    p["head"] = np.zeros(HEAD, np.float32)
    return p


def to_torch_weight(w: np.ndarray, L: Layer):
    """W[o][kp] -> PyTorch conv weight [o][cin][3][3]."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(w[:, :9 * L.cin]))
    return t.reshape(L.cout, 3, 3, L.cin).permute(0, 3, 1, 2).contiguous()


def torch_forward(params: Dict[str, np.ndarray], feat, base, round_bf16: bool = False):
    """FP32 PyTorch restatement of the network (NCHW): feat (N, 16, H, W),
    base (N, 3, H, W); returns (N, 3, H, W). round_bf16 rounds each layer's
    output to bf16 as the HIP path stores it."""
    import torch
    import torch.nn.functional as F
    acts = {"feat": feat}
    for L in LAYERS:
        a = acts[L.a]
        if L.up:   # cropped to the size of the skip at the output level
            ref = acts[SKIP_OF_LEVEL[L.level]]
            a = F.interpolate(a, scale_factor=2, mode="nearest")[..., :ref.shape[-2],
                                                                 :ref.shape[-1]]
        x = torch.cat([a, acts[L.b]], 1) if L.b else a
        wt = to_torch_weight(params[L.name + ".w"], L).to(x.device)
        bt = torch.from_numpy(params[L.name + ".b"]).to(x.device)
        y = F.relu(F.conv2d(x, wt, bt, stride=L.stride, padding=1))
        if round_bf16:
            y = y.to(torch.bfloat16).to(torch.float32)
        acts[L.name] = y
    hw = torch.from_numpy(params["head"]).to(feat.device)
    out = base + F.conv2d(acts["c1"], hw[:48].reshape(3, 16, 1, 1), hw[48:51])
    return out


# ---------------------------------------------------------------- HIP ---

class _Gemm(C.Structure):
    _fields_ = [("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p), ("bias", C.c_void_p),
                ("M", C.c_int), ("N", C.c_int), ("K", C.c_int), ("lda", C.c_int),
                ("ldb", C.c_int), ("ldc", C.c_int)]


class _Conv(C.Structure):
    _fields_ = [("n", C.c_int), ("in_w", C.c_int), ("in_h", C.c_int), ("out_w", C.c_int),
                ("out_h", C.c_int), ("ca", C.c_int), ("cb", C.c_int), ("stride", C.c_int),
                ("up", C.c_int)]


EPI_STORE, EPI_ATOMIC, EPI_BIAS_RELU_BF16, EPI_BIAS_F32 = 0, 1, 2, 3


def _ptr(t) -> int:
    return t.data_ptr()


class HipNet:
    """The network on the GPU through the hand-written training kernels
    (pyr3nn.h): forward, backward and Adam, on torch-allocated device
    tensors (torch is the allocator and the stream, not the arithmetic)."""

    def __init__(self, params: Dict[str, np.ndarray], device: str = "cuda"):
        from .. import _lib
        if _lib._lib is not None and not _lib.loaded_after_torch:
            raise RuntimeError(
                "pyr3_amd: libpyr3hip.so was loaded before torch, so torch's own HIP runtime "
                "cannot open the GPU in this process; import torch before the first use of "
                "pyr3_amd's renderer or kernels")
        import torch
        self.lib = _lib.load()
        self.torch = torch
        self.dev = device
        self.p = {k: torch.from_numpy(v.copy()).to(device) for k, v in params.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.wbf = {L.name: torch.empty(L.cout, L.kp, dtype=torch.bfloat16, device=device)
                    for L in LAYERS}
        self.step_n = 0
        for L in LAYERS:
            self._ok(self.lib.p3nn_cast_bf16(_ptr(self.p[L.name + ".w"]), _ptr(self.wbf[L.name]),
                                             self.p[L.name + ".w"].numel(), self._s()))

    def _s(self):
        return self.torch.cuda.current_stream().cuda_stream

    @staticmethod
    def _ok(rc):
        if rc != 0:
            raise RuntimeError(f"pyr3nn call failed ({rc})")

    def params_numpy(self) -> Dict[str, np.ndarray]:
        return {k: v.detach().cpu().numpy() for k, v in self.p.items()}

    def _conv(self, L: Layer, n: int, sizes) -> _Conv:
        oh, ow = sizes[L.level]
        if L.up:
            ih, iw = sizes[L.level + 1]
        elif L.stride == 2:
            ih, iw = sizes[L.level - 1]
        else:
            ih, iw = oh, ow
        return _Conv(n, iw, ih, ow, oh, L.ca, L.cb, L.stride, 1 if L.up else 0)

    def _gemm(self, a, b, c, bias, M, N, K, lda, ldb, ldc, ak, bk, epi, split=1):
        g = _Gemm(_ptr(a), _ptr(b), _ptr(c), _ptr(bias) if bias is not None else None,
                  M, N, K, lda, ldb, ldc)
        self._ok(self.lib.p3nn_gemm(C.byref(g), ak, bk, epi, split, self._s()))

    def forward(self, feat, base):
        """feat: (n, H, W, 16) bf16; base: (n, H, W, 4) f32. Returns the
        (n, H, W, 3) f32 output; keeps what backward needs."""
        torch = self.torch
        n, H, W, _ = feat.shape
        sizes = level_sizes(H, W)
        self.n, self.sizes = n, sizes
        acts = {"feat": feat}
        cols = {}
        for L in LAYERS:
            cv = self._conv(L, n, sizes)
            P = n * cv.out_h * cv.out_w
            col = torch.empty(P, L.kp, dtype=torch.bfloat16, device=self.dev)
            b = acts[L.b] if L.b else None
            self._ok(self.lib.p3nn_im2col(C.byref(cv), _ptr(acts[L.a]), _ptr(b) if b is not None
                                          else None, _ptr(col), L.kp, self._s()))
            y = torch.empty(n, cv.out_h, cv.out_w, L.cout, dtype=torch.bfloat16, device=self.dev)
            self._gemm(col, self.wbf[L.name], y, self.p[L.name + ".b"], P, L.cout, L.kp, L.kp,
                       L.kp, L.cout, 0, 0, EPI_BIAS_RELU_BF16)
            acts[L.name] = y
            cols[L.name] = col
        self.acts, self.cols = acts, cols
        P = n * H * W
        out = torch.empty(n, H, W, 3, dtype=torch.float32, device=self.dev)
        self._base = base
        self._ok(self.lib.p3nn_head(_ptr(acts["c1"]), _ptr(base), _ptr(self.p["head"]), None, P,
                                    _ptr(out), None, None, None, self._s()))
        return out

    def loss_backward(self, target) -> "torch.Tensor":
        """MSE against target (n, H, W, 3) f32 for the last forward; fills
        the gradients. Returns the loss (a device scalar)."""
        torch = self.torch
        n = self.n
        H, W = self.sizes[0]
        P = n * H * W
        for g in self.g.values():
            g.zero_()
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        out = torch.empty(n, H, W, 3, dtype=torch.float32, device=self.dev)
        dacts = {L.name: torch.zeros(n, *self.sizes[L.level], L.cout, dtype=torch.float32,
                                     device=self.dev) for L in LAYERS}
        self._ok(self.lib.p3nn_head(_ptr(self.acts["c1"]), _ptr(self._base), _ptr(self.p["head"]),
                                    _ptr(target), P, _ptr(out), _ptr(dacts["c1"]),
                                    _ptr(self.g["head"]), _ptr(loss), self._s()))
        for L in reversed(LAYERS):
            cv = self._conv(L, n, self.sizes)
            Pl = n * cv.out_h * cv.out_w
            dy = torch.empty(Pl, L.cout, dtype=torch.bfloat16, device=self.dev)
            self._ok(self.lib.p3nn_relu_bwd(_ptr(dacts[L.name]), _ptr(self.acts[L.name]),
                                            _ptr(dy), Pl * L.cout, self._s()))
            self._ok(self.lib.p3nn_bias_grad(_ptr(dy), Pl, L.cout, _ptr(self.g[L.name + ".b"]),
                                             self._s()))
            split = max(1, min(256, Pl // 4096))
            self._gemm(dy, self.cols[L.name], self.g[L.name + ".w"], None, L.cout, L.kp, Pl,
                       L.cout, L.kp, L.kp, 1, 1, EPI_ATOMIC, split)
            if L.a == "feat":
                continue
            dcol = torch.empty(Pl, L.kp, dtype=torch.float32, device=self.dev)
            self._gemm(dy, self.wbf[L.name], dcol, None, Pl, L.kp, L.cout, L.cout, L.kp, L.kp,
                       0, 1, EPI_STORE)
            db = dacts[L.b] if L.b else None
            self._ok(self.lib.p3nn_col2im(C.byref(cv), _ptr(dcol), L.kp, _ptr(dacts[L.a]),
                                          _ptr(db) if db is not None else None, self._s()))
        return loss

    def adam_step(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.step_n += 1
        for k, w in self.p.items():
            wbf = self.wbf[k[:-2]] if k.endswith(".w") else None
            self._ok(self.lib.p3nn_adam(_ptr(w), _ptr(self.g[k]), _ptr(self.m[k]),
                                        _ptr(self.v[k]), w.numel(), lr, b1, b2, eps, self.step_n,
                                        _ptr(wbf) if wbf is not None else None, self._s()))

    def features(self, hist, img, k2: float):
        """(H, W, 16) bf16 network input of one frame (hist: (H, W, 4) u32
        device tensor viewed as int32; img: (H, W, 4) f32)."""
        torch = self.torch
        H, W = hist.shape[:2]
        out = torch.empty(1, H, W, FEAT_CH, dtype=torch.bfloat16, device=self.dev)
        self._ok(self.lib.p3nn_features(_ptr(hist), _ptr(img), float(k2), H * W, _ptr(out),
                                        self._s()))
        return out
