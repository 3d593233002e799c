"""ORACLE — test infrastructure only.

numpy float64 restatement of the learned denoiser (the kernel-predicting
encoder-decoder whose interface include/sgi_kpcn.h declares): forward pass,
multi-scale filter reconstruction, loss and the full backward pass.  The reference has no learned denoiser, so
this oracle follows the interface's own definitions; it is pinned by central
finite differences of its loss (tests/test_kpcn_oracle_cpu.py), by adjoint
identities of every linear piece (bilinear upsampling against its transpose
included) and by small closed forms (a one-hot filter copies its tap, pooling
of a constant is the constant, bilinear upsampling reproduces linear ramps).

`bf16=True` rounds every stored tensor to BF16 (round to nearest even) at the
points where the GPU stores BF16 (weights, layer outputs, gradients between
layers), so that the GPU path can be compared element-wise; arithmetic stays
float64.  Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

C = 32
LEVELS = 5
LAYERS = 17
TAPS = 25
ERR = TAPS       # level-0 logit channel of the predicted log relative squared error
BLEND = TAPS + 1  # logit channel of the blend with the coarser level's output
HEADS = (12, 13, 14, 15, 16)  # the filter-predicting layer of levels 0..4
LAYER_WEIGHTS = 9 * C * C
LAYER_PARAMS = LAYER_WEIGHTS + C
PARAMS = LAYERS * LAYER_PARAMS


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


E4M3_MAX = 448.0  # largest finite OCP E4M3 value, 1.75 * 2^8


def to_e4m3(x: np.ndarray) -> np.ndarray:
    """Clamp to +-448, then round to the nearest OCP FP8 E4M3 value (ties to even), as float64.

    E4M3 has a sign, 4 exponent bits (bias 7) and 3 fraction bits: normals 2^-6 .. 448 in steps of
    2^(e - 3) within binade [2^e, 2^(e+1)), subnormals k * 2^-9 (k = 1..7), no infinities."""
    a = np.clip(np.asarray(x, np.float64), -E4M3_MAX, E4M3_MAX)
    _, e = np.frexp(np.abs(a))  # |a| in [2^(e-1), 2^e)
    q = np.ldexp(1.0, np.maximum(e - 4, -9))
    return np.copysign(np.round(np.abs(a) / q) * q, a)  # np.round rounds half to even


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


def filter5(logits, radiance):
    """The 5x5 softmax filter of the radiance alone (no blend)."""
    return np.einsum("hwj,jhwc->hwc", softmax25(logits), _taps(radiance))


def _bilinear_axis(n_fine, n_coarse):
    """Source indices and weights of one axis of 2x bilinear upsampling: fine
    pixel x samples the coarse grid at (x + 0.5) / 2 - 0.5, indices clamped."""
    x = np.arange(n_fine)
    k = x // 2
    odd = (x % 2) == 1
    i0 = np.where(odd, k, k - 1)
    w0 = np.where(odd, 0.75, 0.25)
    return np.clip(i0, 0, n_coarse - 1), np.clip(i0 + 1, 0, n_coarse - 1), w0, 1.0 - w0


def upsample_bilinear(low, H, W):
    """[H][W][c] from the coarser level low ([ceil(H/2)][ceil(W/2)][c])."""
    h, w = low.shape[:2]
    y0, y1, wy0, wy1 = _bilinear_axis(H, h)
    x0, x1, wx0, wx1 = _bilinear_axis(W, w)
    wx0, wx1 = wx0[None, :, None], wx1[None, :, None]
    top = wx0 * low[y0][:, x0] + wx1 * low[y0][:, x1]
    bot = wx0 * low[y1][:, x0] + wx1 * low[y1][:, x1]
    return wy0[:, None, None] * top + wy1[:, None, None] * bot


def upsample_bilinear_backward(d):
    """The transpose of upsample_bilinear: the coarser level's gradient."""
    H, W = d.shape[:2]
    h, w = (H + 1) // 2, (W + 1) // 2
    y0, y1, wy0, wy1 = _bilinear_axis(H, h)
    x0, x1, wx0, wx1 = _bilinear_axis(W, w)
    out = np.zeros((h, w) + d.shape[2:])
    for ys, wy in ((y0, wy0), (y1, wy1)):
        for xs, wx in ((x0, wx0), (x1, wx1)):
            np.add.at(out, (ys[:, None], xs[None, :]), wy[:, None, None] * wx[None, :, None] * d)
    return out


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def apply(logits, radiance, coarse=None):
    """(out [H][W][3], err [H][W]) as sgi_kpcn_apply: the 5x5 softmax filter f
    of the radiance, blended with the coarser level's output when given:
    out = f + sigmoid(logits[BLEND]) (up(coarse) - f)."""
    f = filter5(logits, radiance)
    if coarse is not None:
        b = sigmoid(logits[..., BLEND])[..., None]
        f = f + b * (upsample_bilinear(coarse, *radiance.shape[:2]) - f)
    return f, logits[..., ERR]


def apply_backward(logits, radiance, coarse, dout, derr=None):
    """(dlogits [H][W][32], dup [H][W][3] or None) as sgi_kpcn_apply_backward:
    dup is the gradient of up(coarse), dlogits[ERR] = derr."""
    taps = _taps(radiance)
    w = softmax25(logits)
    dl = np.zeros(logits.shape)
    dup = None
    df = dout
    if coarse is not None:
        f = np.einsum("hwj,jhwc->hwc", w, taps)
        u = upsample_bilinear(coarse, *radiance.shape[:2])
        b = sigmoid(logits[..., BLEND])
        df = dout * (1.0 - b)[..., None]
        dup = dout * b[..., None]
        dl[..., BLEND] = (dout * (u - f)).sum(-1) * b * (1.0 - b)
    dw = np.einsum("hwc,jhwc->hwj", df, taps)
    dl[..., :TAPS] = w * (dw - (w * dw).sum(-1, keepdims=True))
    if derr is not None:
        dl[..., ERR] = derr
    return dl, dup


def loss(out, err_logit, target, err_weight, alt=None):
    """(loss, dout [H][W][3], derr [H][W]) as sgi_kpcn_loss with scale 1:
    loss = mean_{p,c} (out - t)^2 / (t^2 + 0.01) + err_weight mean_p (err - log(r_p + 1e-4))^2
    with r_p = mean_c (out - t)^2 / (t^2 + 0.01) held constant; with alt (sgi_kpcn_loss_ratio) the
    error target is log(r_p + 1e-4) - log(a_p + 1e-4), a_p the same mean for alt."""
    H, W = out.shape[:2]
    den = target * target + 0.01
    e = out - target
    r = e * e / den
    err_t = np.log(r.mean(-1) + 1e-4)
    if alt is not None:
        ea = alt - target
        err_t = err_t - np.log((ea * ea / den).mean(-1) + 1e-4)
    de = err_logit - err_t
    N, P = 3.0 * H * W, float(H * W)
    L = r.sum() / N + err_weight * (de * de).sum() / P
    return L, 2.0 * e / den / N, 2.0 * err_weight * de / P


def pyramid(radiance):
    """The radiance at levels 0..4 (each level the 2x2 mean of the finer one)."""
    out = [radiance]
    for _ in range(1, LEVELS):
        out.append(pool(out[-1]))
    return out


def reconstruct(logits, radiance):
    """Coarse-to-fine reconstruction: outputs of levels 0..4 (o[0] is the
    denoised radiance) and the level-0 error logit, from the per-level logits."""
    rad = pyramid(radiance)
    o = [None] * LEVELS
    o[LEVELS - 1], _ = apply(logits[LEVELS - 1], rad[LEVELS - 1])
    for l in range(LEVELS - 2, -1, -1):
        o[l], _ = apply(logits[l], rad[l], o[l + 1])
    return o, logits[0][..., ERR]


def reconstruct_backward(logits, radiance, o, dout, derr):
    """Per-level dlogits of reconstruct for the gradient of o[0] (and of the error logit)."""
    rad = pyramid(radiance)
    dl = [None] * LEVELS
    d = dout
    for l in range(LEVELS):
        coarse = o[l + 1] if l + 1 < LEVELS else None
        dl[l], dup = apply_backward(logits[l], rad[l], coarse, d, derr if l == 0 else None)
        if dup is not None:
            d = upsample_bilinear_backward(dup)
    return dl


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
        a["e0"] = self._conv(0, feat)
        for l in range(1, LEVELS):
            a[f"p{l}"] = self._r(pool(a[f"e{l - 1}"]))
            a[f"c{l}"] = self._conv(2 * l - 1, a[f"p{l}"])
            a[f"e{l}"] = self._conv(2 * l, a[f"c{l}"])
        low = a[f"e{LEVELS - 1}"]
        for l, layer in zip(range(LEVELS - 2, 0, -1), range(9, 12)):
            e = a[f"e{l}"]
            a[f"u{l}"] = self._r(e + upsample(low, *e.shape[:2]))
            a[f"d{l}"] = self._conv(layer, a[f"u{l}"])
            low = a[f"d{l}"]
        a["u0"] = self._r(a["e0"] + upsample(low, *a["e0"].shape[:2]))
        # the filter heads: level 0 reads u0, levels 1..3 d_l, level 4 e4
        a["logits"] = []
        for l in range(LEVELS):
            lg = conv3x3(self.head_input_of(a, l), self.W[HEADS[l]], self.b[HEADS[l]])
            a["logits"].append(lg.astype(np.float32).astype(np.float64) if self.bf16 else lg)
        self.a = a
        return a["logits"]

    @staticmethod
    def head_input_of(a, l):
        if l == 0:
            return a["u0"]
        return a[f"d{l}"] if l < LEVELS - 1 else a[f"e{LEVELS - 1}"]

    def head_input(self, l):
        return self.head_input_of(self.a, l)

    def backward(self, dlogits):
        """Gradient of every layer's parameters [17][9248] for the per-level logits' gradients."""
        a, r = self.a, self._r
        grads = np.zeros((LAYERS, LAYER_PARAMS))

        def layer_back(l, x, y, dy, relu=True, need_dx=True):
            dz = dy * (y > 0) if relu else dy
            dW, db, dx = conv3x3_backward(x, dz, self.W[l])
            grads[l, :LAYER_WEIGHTS] = dW.reshape(-1)
            grads[l, LAYER_WEIGHTS:] = db
            return r(dx) if need_dx else None

        # the heads' input gradients (layer outputs, stored BF16); u0's is also e0's skip gradient
        dh = [layer_back(HEADS[l], self.head_input(l), None, r(dlogits[l]), relu=False) for l in range(LEVELS)]
        de = {0: dh[0]}
        g = r(dh[1] + upsample_backward(dh[0]))  # the gradient of d1
        for l, layer in zip(range(1, LEVELS - 1), range(11, 8, -1)):
            du = layer_back(layer, a[f"u{l}"], a[f"d{l}"], g)
            de[l] = du
            # d_{l+1} (e4 at the last step) feeds the decoder and its own head
            g = r(dh[l + 1] + upsample_backward(du))
        # g is now the gradient of e4 from the decoder and its head
        for l in range(LEVELS - 1, 0, -1):
            gc = layer_back(2 * l, a[f"c{l}"], a[f"e{l}"], g)
            gp = layer_back(2 * l - 1, a[f"p{l}"], a[f"c{l}"], gc)
            H, W = a[f"e{l - 1}"].shape[:2]
            g = r(de[l - 1] + pool_backward(gp, H, W))
        layer_back(0, a["feat"], a["e0"], g, need_dx=False)
        return grads



def _pow2_floor(x):
    """Largest power of two <= x (x > 0)."""
    return 2.0 ** np.floor(np.log2(np.asarray(x, np.float64)))


class NetF8:
    """FP8 (E4M3) inference restatement of the network, following the layout that
    include/sgi_kpcn.h's sgi_kpcn_net documents.

    S [4][32]: per-channel power-of-two scales of levels 0..3 (shared by e_l, the
    decoder output added to it and their sum u_l; e_4 uses S_3; a pooled tensor
    keeps its source's scale); Cs [5][32]: those of c_1..c_4 (rows 1..4). Layer 0
    convolves the BF16 features with BF16 weights; every other layer convolves
    E4M3 activations (stored scaled: e4m3(a s)) with E4M3 weights
    e4m3(W / s_in * w_o), w_o the largest power of two keeping the output
    channel's weights within 448. Products and sums are exact here (float64); a
    channel's value is rounded once to FP32 as acc * k1 + k2 (the GPU's fma) and
    an activation is then stored as e4m3(max(., 0)). Pooling and the skip sums
    are exact before their single E4M3 rounding."""

    def __init__(self, params, S, Cs):
        S, Cs = np.asarray(S, np.float64), np.asarray(Cs, np.float64)
        io = [(None, S[0])]
        for l in range(1, LEVELS):
            io += [(S[l - 1], Cs[l]), (Cs[l], S[min(l, LEVELS - 2)])]
        io += [(S[l], S[l - 1]) for l in range(LEVELS - 2, 0, -1)]
        io += [(s, None) for s in (S[0], S[0], S[1], S[2], S[3])]
        self.W, self.k1, self.k2 = [], [], []
        for l, (s_in, s_out) in enumerate(io):
            Wt, b = unpack(params, l)
            if s_in is None:
                self.W.append(to_bf16(Wt))
                self.k1.append(s_out)
                self.k2.append(b * s_out)
                continue
            Ws = Wt / s_in[None, None, :]
            peak = np.abs(Ws).max(axis=(0, 2))
            w_o = np.where(peak > 0, _pow2_floor(448.0 / np.where(peak > 0, peak, 1.0)), 1.0)
            self.W.append(to_e4m3(Ws * w_o[None, :, None]))
            self.k1.append(1.0 / w_o if s_out is None else s_out / w_o)
            self.k2.append(b if s_out is None else b * s_out)

    def _layer(self, l, x, act=True):
        acc = conv3x3(x, self.W[l], np.zeros(C))
        y = (acc * self.k1[l] + self.k2[l]).astype(np.float32).astype(np.float64)
        return to_e4m3(np.maximum(y, 0.0)) if act else y

    def forward(self, feat):
        """Per-level FP32 logits from the BF16 features [H][W][32] (channels >= 10 zero)."""
        a = {"e0": self._layer(0, to_bf16(feat))}
        for l in range(1, LEVELS):
            a[f"p{l}"] = to_e4m3(pool(a[f"e{l - 1}"]))
            a[f"c{l}"] = self._layer(2 * l - 1, a[f"p{l}"])
            a[f"e{l}"] = self._layer(2 * l, a[f"c{l}"])
        low = a[f"e{LEVELS - 1}"]
        for l, layer in zip(range(LEVELS - 2, 0, -1), range(9, 12)):
            e = a[f"e{l}"]
            a[f"u{l}"] = to_e4m3(e + upsample(low, *e.shape[:2]))
            a[f"d{l}"] = self._layer(layer, a[f"u{l}"])
            low = a[f"d{l}"]
        a["u0"] = to_e4m3(a["e0"] + upsample(low, *a["e0"].shape[:2]))
        self.a = a
        return [self._layer(HEADS[l], Net.head_input_of(a, l), act=False) for l in range(LEVELS)]
