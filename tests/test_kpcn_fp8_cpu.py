"""FP8 (E4M3) inference form of the learned denoiser, on the CPU.

The oracle's E4M3 rounding (oracle/kpcn_oracle.py to_e4m3) is pinned by the
format's closed forms: 127 non-negative values (zero, 7 subnormals k 2^-9,
119 normals with 8 values per binade in steps of 2^(e-3), the top code of the
last binade being NaN), 448 the largest, ties to even, saturation above 448.
The product encoder (stratusgfx_amd/kpcn_fp8.py) is checked against it value
for value and code for code, the product quantiser against the oracle's own
restatement (NetF8) weight for weight, and the FP8 network restatement against
the exact network on a synthetic frame: its reconstruction stays within the
E4M3 activation error (3.8 % here with the shipped parameters; on rendered city
frames the FP8 network's relMSE is within a few % of the BF16 network's),
halving every activation scale moves it by 1.3 % (the exponent range's low end,
where the smallest activations turn subnormal, is a minor part of the error),
and one layer writing at the wrong scale moves it by 18 %."""
import numpy as np

import kpcn_oracle as ko
from stratusgfx_amd import kpcn, kpcn_fp8


def test_e4m3_closed_forms():
    v = np.unique(ko.to_e4m3(np.linspace(0.0, 500.0, 2_000_001)))
    assert len(v) == 127
    assert v[0] == 0.0 and v[1] == 2.0 ** -9 and v[7] == 7 * 2.0 ** -9 and v[8] == 2.0 ** -6
    assert v[-1] == 448.0 and v[-2] == 416.0
    for e in range(-6, 8):  # binade [2^e, 2^(e+1)): 8 values 2^e (1 + k/8)
        inb = v[(v >= 2.0 ** e) & (v < 2.0 ** (e + 1))]
        np.testing.assert_array_equal(inb, 2.0 ** e * (1 + np.arange(8) / 8))
    top = v[v >= 256.0]
    np.testing.assert_array_equal(top, 256.0 * (1 + np.arange(7) / 8))  # 1.875 * 2^8 is the NaN code
    # ties to even, both in the normal range and among the subnormals
    ties = {1.0 + 1 / 16: 1.0, 1.0 + 3 / 16: 1.25, 2.0 ** -10: 0.0, 3 * 2.0 ** -10: 2.0 ** -8,
            5 * 2.0 ** -10: 2.0 ** -8, 440.0: 448.0, 400.0: 384.0, 17.0: 16.0, 19.0: 20.0}
    for x, want in ties.items():
        assert ko.to_e4m3(np.array([x]))[0] == want, x
    np.testing.assert_array_equal(ko.to_e4m3(np.array([464.0, 1e9, -1e9, np.inf, -np.inf])),
                                  [448.0, 448.0, -448.0, 448.0, -448.0])
    x = np.random.default_rng(0).standard_normal(20000) * np.exp(np.random.default_rng(1).uniform(-10, 7, 20000))
    q = ko.to_e4m3(x)
    np.testing.assert_array_equal(ko.to_e4m3(q), q)          # idempotent
    np.testing.assert_array_equal(ko.to_e4m3(-x), -q)        # odd
    s = np.sort(x)
    assert np.all(np.diff(ko.to_e4m3(s)) >= 0)              # monotone
    inside = np.abs(x) <= 448.0                              # half a step at most
    _, e = np.frexp(np.abs(x[inside]))
    step = np.ldexp(1.0, np.maximum(e - 4, -9))
    assert np.all(np.abs(q[inside] - x[inside]) <= step / 2 + 1e-300)


def test_encoder_matches_oracle():
    rng = np.random.default_rng(2)
    x = np.concatenate([np.linspace(-500, 500, 100001), rng.standard_normal(50000) * np.exp(rng.uniform(-12, 6, 50000)),
                        [0.0, -0.0, -1e-9, 448.0, 464.0, -1e9]])
    b = kpcn_fp8.encode_e4m3(x)
    np.testing.assert_array_equal(kpcn_fp8.decode_e4m3(b), ko.to_e4m3(x))
    assert not np.any((b & 0x7F) == 0x7F)  # never a NaN code
    assert b[-4] == 0x80  # a negative value that rounds to zero keeps its sign
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F], np.uint8)
    np.testing.assert_array_equal(kpcn_fp8.encode_e4m3(kpcn_fp8.decode_e4m3(codes)), codes)


def test_quantizer_matches_oracle():
    rng = np.random.default_rng(3)
    params = kpcn.init_params(5)
    S = 2.0 ** rng.integers(-3, 7, (4, 32))
    Cs = 2.0 ** rng.integers(-3, 7, (5, 32))
    w8, epi = kpcn_fp8.quantize(params, S, Cs)
    ref = ko.NetF8(params, S, Cs)
    for l in range(1, ko.LAYERS):
        np.testing.assert_array_equal(kpcn_fp8.decode_e4m3(w8[l]), ref.W[l])
    for l in range(ko.LAYERS):
        np.testing.assert_array_equal(epi[l, 0], np.asarray(ref.k1[l], np.float32))
        np.testing.assert_array_equal(epi[l, 1], np.asarray(ref.k2[l], np.float32))
    np.testing.assert_array_equal(w8[0], 0)


def _synthetic_frame(H=48, W=64):
    rng = np.random.default_rng(0)
    yy, xx = np.mgrid[0:H, 0:W] / 16.0
    rad = np.stack([0.5 + 0.4 * np.sin(xx + c) * np.cos(yy) for c in range(3)], -1) * rng.uniform(0.5, 1.5, (H, W, 1))
    feat = np.zeros((H, W, 32))
    feat[..., 0:3] = np.log1p(rad * 64.0)  # the radiance feature log(1 + 64 radiance)
    feat[..., 3:6] = 0.5 + 0.3 * np.sin(xx[..., None] * np.array([1.0, 2.0, 3.0]))
    n = np.stack([np.sin(xx), np.cos(yy), np.ones_like(xx)], -1)
    feat[..., 6:9] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    feat[..., 9] = np.log1p(10 + yy) * 0.125
    return feat, rad


def test_fp8_network_close_to_exact():
    params = kpcn.load_params()
    feat, rad = _synthetic_frame()
    exact = ko.Net(params)
    o, _ = ko.reconstruct(exact.forward(feat), rad)
    maxima = {k: np.abs(v).max(axis=(0, 1)) for k, v in exact.a.items() if k not in ("logits", "feat")}
    S, Cs = kpcn_fp8.group_scales(maxima)
    assert np.all(S * np.stack([maxima[f"e{l}"] for l in range(4)]) <= 224.0)

    def recon(S_, Cs_):
        o8, _ = ko.reconstruct(ko.NetF8(params, S_, Cs_).forward(feat), rad)
        return o8[0]

    o8 = recon(S, Cs)
    err = np.linalg.norm(o8 - o[0]) / np.linalg.norm(o[0])
    assert err < 0.08, err
    moved = np.linalg.norm(recon(S / 2, Cs / 2) - o8) / np.linalg.norm(o8)
    assert moved < 0.02, moved
    # a misapplied scale (layer 2 writing e_1 at 4x the scale its readers assume) is far off
    bad = ko.NetF8(params, S, Cs)
    bad.k1[2] = bad.k1[2] * 4.0
    ob, _ = ko.reconstruct(bad.forward(feat), rad)
    assert np.linalg.norm(ob[0] - o[0]) / np.linalg.norm(o[0]) > 2 * err
