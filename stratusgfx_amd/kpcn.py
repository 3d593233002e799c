"""The learned denoiser's host side: parameters, the training forward / backward
over the C ABI of include/sgi_kpcn.h, Adam, and weight files.

Every tensor operation runs in libsgi.so's HIP kernels (MFMA convolutions,
pooling, filter application, loss, Adam); torch only allocates device memory
and adds per-image gradients across a batch.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi

C = _abi.KPCN_CHANNELS
LEVELS = _abi.KPCN_LEVELS
LAYERS = _abi.KPCN_LAYERS
LAYER_WEIGHTS = _abi.KPCN_LAYER_WEIGHTS
LAYER_PARAMS = _abi.KPCN_LAYER_PARAMS
PARAMS = _abi.KPCN_PARAMS
# the filter head of level l reads u_0 = e_0 + up(d_1) at level 0, d_l at levels 1..3 and e_4 at level 4
HEADS = tuple(12 + l for l in range(LEVELS))

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "kpcn_city.npy")


def init_params(seed: int) -> np.ndarray:
    """He-normal weights (fan-in 9 x 32), zero biases; the filter heads start at
    1/10 of that scale so the first filters are close to a 5x5 box and every
    blend close to 1/2."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    for l in range(LAYERS):
        std = np.sqrt(2.0 / (9 * C)) * (0.1 if l in HEADS else 1.0)
        p[l * LAYER_PARAMS:l * LAYER_PARAMS + LAYER_WEIGHTS] = rng.normal(0.0, std, LAYER_WEIGHTS)
    return p


def save_params(params: np.ndarray, path: str) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(params, np.float32))


def load_params(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.shape != (PARAMS,) or p.dtype != np.float32:
        raise ValueError(f"{path}: expected {PARAMS} float32 parameters, found {p.shape} {p.dtype}")
    return p


def level_sizes(H: int, W: int) -> list[tuple[int, int]]:
    out = []
    for _ in range(LEVELS):
        out.append((H, W))
        H, W = (H + 1) // 2, (W + 1) // 2
    return out


def _p(t) -> int:
    return 0 if t is None else t.data_ptr()


class Weights:
    """FP32 parameters on the device and their BF16 / FP32 inference copies."""

    def __init__(self, params: np.ndarray, device="cuda", lib=None):
        import torch

        self.lib = lib if lib is not None else _abi.load()
        self.params = torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(device)
        self.w = torch.empty(LAYERS * LAYER_WEIGHTS, dtype=torch.bfloat16, device=device)
        self.wt = torch.empty(LAYERS * LAYER_WEIGHTS, dtype=torch.bfloat16, device=device)
        self.biases = torch.empty(LAYERS * C, dtype=torch.float32, device=device)
        self.pack()

    def pack(self, stream=0) -> None:
        _abi.call(self.lib, "sgi_kpcn_pack_weights", _p(self.params), _p(self.w), _p(self.wt), _p(self.biases), stream)

    def net(self) -> _abi.KpcnNet:
        return _abi.KpcnNet(_p(self.w), _p(self.biases))

    def layer_w(self, l: int):
        return self.w[l * LAYER_WEIGHTS:(l + 1) * LAYER_WEIGHTS]

    def layer_wt(self, l: int):
        return self.wt[l * LAYER_WEIGHTS:(l + 1) * LAYER_WEIGHTS]

    def layer_b(self, l: int):
        return self.biases[l * C:(l + 1) * C]


class Trainer:
    """Training-mode forward / backward at a fixed image size (activations kept
    for the backward pass), gradient accumulation over a batch and Adam."""

    def __init__(self, weights: Weights, height: int, width: int, device="cuda"):
        import torch

        self.torch = torch
        self.wts = weights
        self.lib = weights.lib
        self.H, self.W = height, width
        self.sizes = level_sizes(height, width)
        bf = dict(dtype=torch.bfloat16, device=device)
        f32 = dict(dtype=torch.float32, device=device)

        def act(l):
            h, w = self.sizes[l]
            return torch.empty((h, w, C), **bf)

        self.a = {}
        self.a["e0"], self.a["u0"] = act(0), act(0)
        for l in range(1, LEVELS):
            self.a[f"p{l}"], self.a[f"c{l}"], self.a[f"e{l}"] = act(l), act(l), act(l)
        for l in range(1, LEVELS - 1):
            self.a[f"u{l}"], self.a[f"d{l}"] = act(l), act(l)
        # gradient buffers: two per level (the gradient of a tensor and of its producer's input)
        self.g = [[act(l), act(l)] for l in range(LEVELS)]
        self.gd = [act(l) for l in range(LEVELS)]  # decoder-side gradients kept for the skip sums
        # per level: head logits (FP32), their gradient (BF16), the radiance pyramid, the
        # reconstruction and its gradient, and the gradient of the upsampled coarser output
        self.logits = [torch.empty((h, w, C), **f32) for h, w in self.sizes]
        self.dlogits = [torch.empty((h, w, C), **bf) for h, w in self.sizes]
        self.pyr = torch.empty(int(self.lib.sgi_kpcn_pyramid_floats(height, width)), **f32)
        self.rad = [None] + self._split(self.pyr)
        self.o = [torch.empty((h, w, 3), **f32) for h, w in self.sizes]
        self.do = [torch.empty((h, w, 3), **f32) for h, w in self.sizes]
        self.dup = [torch.empty((h, w, 3), **f32) for h, w in self.sizes]
        self.derr = torch.empty((height, width), **f32)
        self.partials = torch.empty(int(self.lib.sgi_kpcn_wgrad_partials(height, width)), **f32)
        self.layer_grad = torch.empty(LAYER_PARAMS, **f32)
        self.grad = torch.zeros(PARAMS, **f32)
        self.m = torch.zeros(PARAMS, **f32)
        self.v = torch.zeros(PARAMS, **f32)
        self.loss_sums = torch.zeros(2, **f32)
        self.step_count = 0

    def _split(self, pyr):
        out, off = [], 0
        for h, w in self.sizes[1:]:
            out.append(pyr[off:off + h * w * 3].view(h, w, 3))
            off += h * w * 3
        return out

    # -- forward -----------------------------------------------------------
    def _conv(self, l, x, out, level, relu=True, out_f32=False, stream=0):
        h, w = self.sizes[level]
        _abi.call(self.lib, "sgi_kpcn_conv3x3", _p(x), 0, _p(self.wts.layer_w(l)), _p(self.wts.layer_b(l)),
                  1 if relu else 0, h, w, _p(out), 1 if out_f32 else 0, stream)

    def head_input(self, l):
        if l == 0:
            return self.a["u0"]
        return self.a[f"d{l}"] if l < LEVELS - 1 else self.a[f"e{LEVELS - 1}"]

    def forward(self, feat, radiance, stream=0):
        """The network and the multi-scale reconstruction; returns the denoised radiance o_0."""
        a = self.a
        self.feat = feat
        self.rad[0] = radiance
        self._conv(0, feat, a["e0"], 0, stream=stream)
        for l in range(1, LEVELS):
            h, w = self.sizes[l - 1]
            _abi.call(self.lib, "sgi_kpcn_pool", _p(a[f"e{l - 1}"]), h, w, _p(a[f"p{l}"]), stream)
            self._conv(2 * l - 1, a[f"p{l}"], a[f"c{l}"], l, stream=stream)
            self._conv(2 * l, a[f"c{l}"], a[f"e{l}"], l, stream=stream)
        low = a[f"e{LEVELS - 1}"]
        for l, layer in zip(range(LEVELS - 2, 0, -1), range(9, 12)):
            h, w = self.sizes[l]
            _abi.call(self.lib, "sgi_kpcn_upsample_add", _p(a[f"e{l}"]), _p(low), h, w, _p(a[f"u{l}"]), stream)
            self._conv(layer, a[f"u{l}"], a[f"d{l}"], l, stream=stream)
            low = a[f"d{l}"]
        h, w = self.sizes[0]
        _abi.call(self.lib, "sgi_kpcn_upsample_add", _p(a["e0"]), _p(low), h, w, _p(a["u0"]), stream)
        for l in range(LEVELS):
            self._conv(HEADS[l], self.head_input(l), self.logits[l], l, relu=False, out_f32=True, stream=stream)
        _abi.call(self.lib, "sgi_kpcn_radiance_pyramid", _p(radiance), self.H, self.W, _p(self.pyr), stream)
        for l in range(LEVELS - 1, -1, -1):
            h, w = self.sizes[l]
            coarse = self.o[l + 1] if l + 1 < LEVELS else None
            _abi.call(self.lib, "sgi_kpcn_apply", _p(self.logits[l]), _p(self.rad[l]), _p(coarse), h, w,
                      _p(self.o[l]), 0, stream)
        return self.o[0]

    # -- backward ------------------------------------------------------------
    def _layer_back(self, l, x, y, dy, dx_out, level, need_dx=True, stream=0):
        """Adds layer l's parameter gradient into self.grad; writes the input gradient to dx_out."""
        h, w = self.sizes[level]
        _abi.call(self.lib, "sgi_kpcn_conv3x3_wgrad", _p(x), _p(dy), _p(y), h, w, _p(self.partials),
                  _p(self.layer_grad), stream)
        self.grad[l * LAYER_PARAMS:(l + 1) * LAYER_PARAMS] += self.layer_grad
        if need_dx:
            _abi.call(self.lib, "sgi_kpcn_conv3x3", _p(dy), _p(y), _p(self.wts.layer_wt(l)), 0, 0, h, w,
                      _p(dx_out), 0, stream)

    def loss_backward(self, target, err_weight=0.05, scale=1.0, stream=0, alternative=None):
        """Loss of the last forward against target and the gradient of every layer (added to grad).
        alternative: another output of the same input (the a-trous filter's); the error channel then
        learns the log ratio of the two outputs' relative squared errors (sgi_kpcn_loss_ratio)."""
        a, g = self.a, self.g
        if alternative is None:
            _abi.call(self.lib, "sgi_kpcn_loss", _p(self.o[0]), _p(self.logits[0]), _p(target), self.H, self.W,
                      ctypes.c_float(err_weight), ctypes.c_float(scale), _p(self.do[0]), _p(self.derr),
                      _p(self.loss_sums), stream)
        else:
            _abi.call(self.lib, "sgi_kpcn_loss_ratio", _p(self.o[0]), _p(self.logits[0]), _p(target),
                      _p(alternative), self.H, self.W, ctypes.c_float(err_weight), ctypes.c_float(scale),
                      _p(self.do[0]), _p(self.derr), _p(self.loss_sums), stream)
        # the reconstruction, fine to coarse: each level's logits and the coarser output's gradient
        for l in range(LEVELS):
            h, w = self.sizes[l]
            coarse = self.o[l + 1] if l + 1 < LEVELS else None
            _abi.call(self.lib, "sgi_kpcn_apply_backward", _p(self.logits[l]), _p(self.rad[l]), _p(coarse),
                      _p(self.do[l]), _p(self.derr) if l == 0 else 0, h, w, _p(self.dlogits[l]),
                      _p(self.dup[l]) if coarse is not None else 0, stream)
            if coarse is not None:
                _abi.call(self.lib, "sgi_kpcn_upsample_bilinear_backward", _p(self.dup[l]), h, w,
                          _p(self.do[l + 1]), stream)
        # the heads (no ReLU): u0's gradient (also e0's skip gradient) lands in gd[0], d_l's
        # (e4's) from its head in g[l][0]
        self._layer_back(HEADS[0], self.head_input(0), None, self.dlogits[0], self.gd[0], 0, stream=stream)
        for l in range(1, LEVELS):
            self._layer_back(HEADS[l], self.head_input(l), None, self.dlogits[l], g[l][0], l, stream=stream)
        h, w = self.sizes[0]
        _abi.call(self.lib, "sgi_kpcn_upsample_backward", _p(self.gd[0]), h, w, _p(g[1][0]), 1, stream)
        cur = g[1][0]  # gradient of d1
        for l, layer in zip(range(1, LEVELS - 1), range(11, 8, -1)):
            # gradient of u_l, which is also the skip gradient of e_l
            self._layer_back(layer, a[f"u{l}"], a[f"d{l}"], cur, self.gd[l], l, stream=stream)
            h, w = self.sizes[l]
            nxt = g[l + 1][0]  # holds d_{l+1}'s (e4's) head gradient; add the decoder's
            _abi.call(self.lib, "sgi_kpcn_upsample_backward", _p(self.gd[l]), h, w, _p(nxt), 1, stream)
            cur = nxt  # gradient of d_{l+1} (or e4 when l + 1 == 4)
        for l in range(LEVELS - 1, 0, -1):
            # cur = total gradient of e_l
            self._layer_back(2 * l, a[f"c{l}"], a[f"e{l}"], cur, g[l][1], l, stream=stream)
            self._layer_back(2 * l - 1, a[f"p{l}"], a[f"c{l}"], g[l][1], g[l][0], l, stream=stream)
            h, w = self.sizes[l - 1]
            # e_{l-1}'s gradient: the decoder's skip gradient plus the pooled path
            _abi.call(self.lib, "sgi_kpcn_pool_backward", _p(g[l][0]), h, w, _p(self.gd[l - 1]), 1, stream)
            cur = self.gd[l - 1]
        self._layer_back(0, self.feat, a["e0"], cur, None, 0, need_dx=False, stream=stream)

    def zero_grad(self):
        self.grad.zero_()
        self.loss_sums.zero_()

    def step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=0):
        self.step_count += 1
        _abi.call(self.lib, "sgi_kpcn_adam", _p(self.wts.params), _p(self.grad), _p(self.m), _p(self.v),
                  ctypes.c_float(lr), ctypes.c_float(beta1), ctypes.c_float(beta2), ctypes.c_float(eps),
                  self.step_count, _p(self.wts.w), _p(self.wts.wt), _p(self.wts.biases), stream)
