"""FP8 (OCP E4M3) inference form of the learned denoiser: calibration of
per-channel activation scales, quantised weights and the per-layer epilogues
that include/sgi_kpcn.h's sgi_kpcn_net describes.

Scales are powers of two, so every rescaling is exact and the only roundings
are the E4M3 stores. Tensors that are added to each other share a scale: at
level l the skip e_l, the decoder output d_{l+1} (e_4 at level 3) and their sum
u_l use S_l, and a pooled tensor p_{l+1} keeps e_l's; c_l has its own C_l.
Calibration runs the BF16 network (the training kernels) on sample frames and
takes each channel's largest activation; the scale leaves a factor of two of
headroom below the E4M3 maximum (448).
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .kpcn import HEADS, LAYER_WEIGHTS, LAYERS, LEVELS, C, Weights, _p

E4M3_MAX = 448.0
HEADROOM = 2.0


def encode_e4m3(x: np.ndarray) -> np.ndarray:
    """FP32 values -> E4M3 bytes: clamp to +-448, round to nearest even (sign, 4 exponent
    bits with bias 7, 3 fraction bits; subnormals below 2^-6 in steps of 2^-9), as gfx950's
    conversion does (a negative value that rounds to zero gives -0, 0x80)."""
    v = np.clip(np.asarray(x, np.float64), -E4M3_MAX, E4M3_MAX)
    a = np.abs(v)
    _, e = np.frexp(a)  # a in [2^(e-1), 2^e)
    step = np.ldexp(1.0, np.maximum(e - 4, -9))
    r = np.round(a / step) * step  # half to even
    _, e = np.frexp(r)
    normal = r >= 2.0 ** -6
    exp_field = np.where(normal, e + 6, 0)
    frac = np.where(normal, np.ldexp(r, 4 - e) - 8.0, np.ldexp(r, 9))
    bits = (exp_field.astype(np.int64) << 3) | frac.astype(np.int64)
    bits = np.where(r == 0.0, 0, bits)
    return (bits | np.where(np.signbit(v), 0x80, 0)).astype(np.uint8)  # a negative value rounding to 0 keeps its sign


def decode_e4m3(b: np.ndarray) -> np.ndarray:
    """E4M3 bytes -> float64 (0x7f / 0xff, the NaN codes, never produced by encode_e4m3)."""
    b = np.asarray(b, np.uint8).astype(np.int64)
    exp_field, frac = (b >> 3) & 0xF, b & 7
    mag = np.where(exp_field == 0, np.ldexp(frac.astype(np.float64), -9),
                   np.ldexp(8.0 + frac, exp_field - 10))
    return np.where(b & 0x80, -mag, mag)


def pow2_floor(x: np.ndarray) -> np.ndarray:
    """Largest power of two <= x (x > 0)."""
    _, e = np.frexp(np.asarray(x, np.float64))
    return np.ldexp(1.0, e - 1)


def activation_scale(maxabs: np.ndarray) -> np.ndarray:
    """Per-channel power-of-two scale s with maxabs * s <= 448 / HEADROOM (1 for an all-zero channel)."""
    m = np.asarray(maxabs, np.float64)
    return np.where(m > 0.0, pow2_floor(E4M3_MAX / HEADROOM / np.where(m > 0.0, m, 1.0)), 1.0)


def layer_scales(S: np.ndarray, Cs: np.ndarray) -> list[tuple[np.ndarray | None, np.ndarray | None]]:
    """(input scale, output scale) of every layer; None: the BF16 features (layer 0's input) or a
    head's FP32 logits. S [4][32] (levels 0..3), Cs [5][32] (C_1..C_4 at rows 1..4)."""
    io: list[tuple[np.ndarray | None, np.ndarray | None]] = [(None, S[0])]
    for l in range(1, LEVELS):
        io.append((S[l - 1], Cs[l]))                     # 2l - 1: p_l -> c_l
        io.append((Cs[l], S[min(l, LEVELS - 2)]))        # 2l: c_l -> e_l (e_4 shares S_3)
    for l in range(LEVELS - 2, 0, -1):
        io.append((S[l], S[l - 1]))                      # 9, 10, 11: u_l -> d_l
    head_in = [S[0], S[0], S[1], S[2], S[3]]             # u_0, d_1, d_2, d_3, e_4
    io.extend((head_in[l], None) for l in range(LEVELS))
    assert len(io) == LAYERS
    return io


def group_scales(maxima: dict[str, np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    """S [4][32] and C [5][32] from per-tensor channel maxima (keys e0..e4, u0..u3, d1..d3, c1..c4)."""
    S = np.ones((LEVELS - 1, C))
    for l in range(LEVELS - 1):
        m = np.maximum(maxima[f"e{l}"], maxima[f"u{l}"])
        m = np.maximum(m, maxima[f"d{l + 1}"] if l + 1 < LEVELS - 1 else maxima[f"e{LEVELS - 1}"])
        S[l] = activation_scale(m)
    Cs = np.ones((LEVELS, C))
    for l in range(1, LEVELS):
        Cs[l] = activation_scale(maxima[f"c{l}"])
    return S, Cs


def quantize(params: np.ndarray, S: np.ndarray, Cs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """E4M3 weights [17][9][32][32] (layer 0 unused: zeros) and epilogues [17][2][32] (k1, k2)."""
    params = np.asarray(params, np.float64)
    w8 = np.zeros((LAYERS, 9, C, C), np.uint8)
    epi = np.zeros((LAYERS, 2, C), np.float32)
    for l, (s_in, s_out) in enumerate(layer_scales(S, Cs)):
        p = params[l * (LAYER_WEIGHTS + C):(l + 1) * (LAYER_WEIGHTS + C)]
        Wt, b = p[:LAYER_WEIGHTS].reshape(9, C, C), p[LAYER_WEIGHTS:]
        if s_in is None:  # layer 0: BF16 weights, E4M3 output
            epi[l, 0], epi[l, 1] = s_out, b * s_out
            continue
        Ws = Wt / s_in[None, None, :]
        peak = np.abs(Ws).max(axis=(0, 2))
        w_o = np.where(peak > 0.0, pow2_floor(E4M3_MAX / np.where(peak > 0.0, peak, 1.0)), 1.0)
        w8[l] = encode_e4m3(Ws * w_o[None, :, None])
        if s_out is None:  # a filter head: FP32 logits
            epi[l, 0], epi[l, 1] = 1.0 / w_o, b
        else:
            epi[l, 0], epi[l, 1] = s_out / w_o, b * s_out
    return w8, epi


def calibrate(weights: Weights, frames) -> dict[str, np.ndarray]:
    """Per-channel maxima of every activation of the BF16 network over frames, an iterable of
    (features [H][W][32] BF16, radiance [H][W][3] FP32) device tensors of one size."""
    import torch

    from .kpcn import Trainer
    maxima: dict[str, np.ndarray] = {}
    tr = None
    for feat, rad in frames:
        if tr is None:
            tr = Trainer(weights, feat.shape[0], feat.shape[1])
        tr.forward(feat, rad)
        for k, t in tr.a.items():
            m = t.float().abs().amax(dim=(0, 1)).cpu().numpy().astype(np.float64)
            maxima[k] = np.maximum(maxima.get(k, 0.0), m)
        torch.cuda.synchronize()
    return maxima


class Fp8Net:
    """The FP8 inference form on the device (with the BF16 weights layer 0 reads)."""

    def __init__(self, weights: Weights, w8: np.ndarray, epi: np.ndarray, device="cuda"):
        import torch

        self.wts = weights
        self.w8 = torch.from_numpy(np.ascontiguousarray(w8, np.uint8).reshape(-1)).to(device)
        self.epi = torch.from_numpy(np.ascontiguousarray(epi, np.float32).reshape(-1)).to(device)

    def net(self) -> _abi.KpcnNet:
        return _abi.KpcnNet(_p(self.wts.w), _p(self.wts.biases), _p(self.w8), _p(self.epi))


def save(path: str, S: np.ndarray, Cs: np.ndarray, w8: np.ndarray, epi: np.ndarray) -> None:
    np.savez(path, S=S, C=Cs, weights_fp8=w8, epilogue=epi)


def load(path: str) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    with np.load(path, allow_pickle=False) as z:
        return z["S"], z["C"], z["weights_fp8"], z["epilogue"]


__all__ = ["encode_e4m3", "decode_e4m3", "activation_scale", "layer_scales", "group_scales", "quantize",
           "calibrate", "Fp8Net", "save", "load", "HEADS"]
