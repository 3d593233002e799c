"""ORACLE — test infrastructure only.

numpy float64 restatement of the learned denoiser (the kernel-predicting
encoder-decoder whose interface include/sgi_kpcn.h declares): forward pass,
loss and the full backward pass.  The reference has no learned denoiser, so
this oracle follows the interface's own definitions; it is pinned by central
finite differences of its loss (tests/test_kpcn_oracle_cpu.py) and by small
closed forms (a one-hot filter copies its tap, pooling of a constant is the
constant, upsampling commutes with pooling's adjoint).

`bf16=True` rounds every stored tensor to BF16 (round to nearest even) at the
points where the GPU stores BF16 (weights, layer outputs, gradients between
layers), so that the GPU path can be compared element-wise; arithmetic stays
float64.  Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

C = 32
LEVELS = 5
LAYERS = 15
TAPS = 25
LAYER_WEIGHTS = 9 * C * C
LAYER_PARAMS = LAYER_WEIGHTS + C
PARAMS = LAYERS * LAYER_PARAMS


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, layer: int):
    """(W [9][out][in], b [32]) of one layer from the flat FP32 parameters."""
    p = params[layer * LAYER_PARAMS:(layer + 1) * LAYER_PARAMS]
    return p[:LAYER_WEIGHTS].reshape(9, C, C).astype(np.float64), p[LAYER_WEIGHTS:].astype(np.float64)


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """s[y][x] = x[y + dy][x + dx] with zeros outside."""
    H, W = x.shape[:2]
    s = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    if ys < ye and xs < xe:
        s[ys:ye, xs:xe] = x[ys + dy:ye + dy, xs + dx:xe + dx]
    return s


def conv3x3(x, Wt, b):
    """y[p][o] = b[o] + sum_t sum_i W[t][o][i] x[p + off(t)][i], off(t) = (t // 3 - 1, t % 3 - 1)."""
    y = np.zeros(x.shape[:2] + (C,))
    for t in range(9):
        y += _shift(x, t // 3 - 1, t % 3 - 1) @ Wt[t].T
    return y + (b if b is not None else 0.0)


def conv3x3_backward(x, dz, Wt):
    """(dW [9][o][i], db [o], dx) of y = conv3x3(x) for the pre-activation gradient dz."""
    dW = np.zeros((9, C, C))
    dx = np.zeros_like(x)
    for t in range(9):
        dy_, dx_ = t // 3 - 1, t % 3 - 1
        xs = _shift(x, dy_, dx_)
        dW[t] = dz.reshape(-1, C).T @ xs.reshape(-1, C)
        dx += _shift(dz @ Wt[t], -dy_, -dx_)
    return dW, dz.reshape(-1, C).sum(0), dx


def pool(x):
    H, W = x.shape[:2]
    h, w = (H + 1) // 2, (W + 1) // 2
    s = np.zeros((h, w) + x.shape[2:])
    n = np.zeros((h, w, 1))
    for dy in (0, 1):
        for dx in (0, 1):
            part = x[dy::2, dx::2]
            s[:part.shape[0], :part.shape[1]] += part
            n[:part.shape[0], :part.shape[1]] += 1
    return s / n


def pool_backward(dout, H, W):
    h, w = dout.shape[:2]
    n = np.zeros((h, w, 1))
    for dy in (0, 1):
        for dx in (0, 1):
            n[:len(range(dy, H, 2)), :len(range(dx, W, 2))] += 1
    g = dout / n
    return np.repeat(np.repeat(g, 2, 0), 2, 1)[:H, :W]


def upsample(low, H, W):
    return np.repeat(np.repeat(low, 2, 0), 2, 1)[:H, :W]


def upsample_backward(dout):
    H, W = dout.shape[:2]
    h, w = (H + 1) // 2, (W + 1) // 2
    s = np.zeros((h, w) + dout.shape[2:])
    for dy in (0, 1):
        for dx in (0, 1):
            part = dout[dy::2, dx::2]
            s[:part.shape[0], :part.shape[1]] += part
    return s


def _taps(radiance):
    """[25][H][W][3]: radiance at p + (j % 5 - 2, j // 5 - 2), clamped to the image."""
    H, W = radiance.shape[:2]
    out = []
    for j in range(TAPS):
        ox, oy = j % 5 - 2, j // 5 - 2
        ys = np.clip(np.arange(H) + oy, 0, H - 1)
        xs = np.clip(np.arange(W) + ox, 0, W - 1)
        out.append(radiance[ys][:, xs])
    return np.stack(out)


def softmax25(logits):
    lg = logits[..., :TAPS]
    e = np.exp(lg - lg.max(-1, keepdims=True))
    return e / e.sum(-1, keepdims=True)


def apply(logits, radiance):
    """(out [H][W][3], err [H][W]) as sgi_kpcn_apply."""
    w = softmax25(logits)
    out = np.einsum("hwj,jhwc->hwc", w, _taps(radiance))
    return out, logits[..., TAPS]


def loss(logits, radiance, target, err_weight):
    """(loss, dlogits [H][W][32]) as sgi_kpcn_loss with scale 1."""
    H, W = radiance.shape[:2]
    taps = _taps(radiance)
    w = softmax25(logits)
    out = np.einsum("hwj,jhwc->hwc", w, taps)
    den = target * target + 0.01
    e = out - target
    r = (e * e / den)
    rp = r.mean(-1)
    err_t = np.log(rp + 1e-4)
    de = logits[..., TAPS] - err_t
    N, P = 3.0 * H * W, float(H * W)
    L = r.sum() / N + err_weight * (de * de).sum() / P
    d_out = 2.0 * e / den / N
    dw = np.einsum("hwc,jhwc->hwj", d_out, taps)
    dl = np.zeros(logits.shape)
    dl[..., :TAPS] = w * (dw - (w * dw).sum(-1, keepdims=True))
    dl[..., TAPS] = 2.0 * err_weight * de / P
    return L, dl


class Net:
    """Forward / backward of the whole network on one image."""

    def __init__(self, params: np.ndarray, bf16: bool = False):
        self.bf16 = bf16
        self.W, self.b = [], []
        for l in range(LAYERS):
            Wt, b = unpack(params, l)
            self.W.append(to_bf16(Wt) if bf16 else Wt)
            self.b.append(b)

    def _r(self, x):
        return to_bf16(x) if self.bf16 else x

    def _conv(self, l, x, relu=True):
        y = conv3x3(x, self.W[l], self.b[l])
        return self._r(np.maximum(y, 0.0) if relu else y)

    def forward(self, feat):
        a = {"feat": feat}
        a["c0"] = self._conv(0, feat)
        a["e0"] = self._conv(1, a["c0"])
        for l in range(1, LEVELS):
            a[f"p{l}"] = self._r(pool(a[f"e{l - 1}"]))
            a[f"c{l}"] = self._conv(2 * l, a[f"p{l}"])
            a[f"e{l}"] = self._conv(2 * l + 1, a[f"c{l}"])
        low = a[f"e{LEVELS - 1}"]
        for l, layer in zip(range(LEVELS - 2, -1, -1), range(10, 14)):
            e = a[f"e{l}"]
            a[f"u{l}"] = self._r(e + upsample(low, *e.shape[:2]))
            a[f"d{l}"] = self._conv(layer, a[f"u{l}"])
            low = a[f"d{l}"]
        a["logits"] = conv3x3(a["d0"], self.W[14], self.b[14])
        if self.bf16:
            a["logits"] = a["logits"].astype(np.float32).astype(np.float64)
        self.a = a
        return a["logits"]

    def backward(self, dlogits):
        """Gradient of every layer's parameters [15][9248] for the logits' gradient."""
        a, r = self.a, self._r
        grads = np.zeros((LAYERS, LAYER_PARAMS))

        def layer_back(l, x, y, dy, relu=True, need_dx=True):
            dz = dy * (y > 0) if relu else dy
            dW, db, dx = conv3x3_backward(x, dz, self.W[l])
            grads[l, :LAYER_WEIGHTS] = dW.reshape(-1)
            grads[l, LAYER_WEIGHTS:] = db
            return r(dx) if need_dx else None

        g = layer_back(14, a["d0"], None, r(dlogits), relu=False)
        de = {}
        for l, layer in zip(range(0, LEVELS - 1), range(13, 9, -1)):
            du = layer_back(layer, a[f"u{l}"], a[f"d{l}"], g)
            de[l] = du
            g = r(upsample_backward(du))
        # g is now the gradient of e4 from the decoder
        for l in range(LEVELS - 1, 0, -1):
            gc = layer_back(2 * l + 1, a[f"c{l}"], a[f"e{l}"], g)
            gp = layer_back(2 * l, a[f"p{l}"], a[f"c{l}"], gc)
            H, W = a[f"e{l - 1}"].shape[:2]
            g = r(de[l - 1] + pool_backward(gp, H, W))
        gc = layer_back(1, a["c0"], a["e0"], g)
        layer_back(0, a["feat"], a["c0"], gc, need_dx=False)
        return grads

