"""Pins of the oracle's FXAA (fxaa_luminance.fs + fxaa_smoothing.fs,
PerformFxaaPostFx_ at StratusRendererBackend.cpp:2293-2321) and of its
history store for TAA (PerformTaaPostFx_'s copy, :2342-2349), against what
the shaders fix in closed form.

FXAA, with the reference's uniforms (contrast 0.0625, relative 0.125,
sub-pixel 1):
- the luminance pass keeps every rgb code and stores the luminance
  0.2126 r + 0.7152 g + 0.0722 b rounded to the nearest code;
- a flat or low-contrast neighbourhood is passed through unchanged;
- a straight black / white edge: the pixels on both sides blend across the
  edge by the sub-pixel factor smoothstep(1/3)^2 = (7/27)^2 and the rest is
  unchanged (a swapped horizontal / vertical test would blend along the edge
  and change nothing);
- the end of a horizontal edge at distance D: the walk (steps 1, 1.5, 2 x 7,
  4, guess jump 8) stops at the first offset >= D - 1/2 and the blend is
  0.5 - d / (d + 28.5), which moves the code by (1 - g) times that factor.
"""
import numpy as np

import oracle as orc

L = orc.lib()


def _gray(codes):
    c = np.asarray(codes, np.uint32)
    return c | (c << 8) | (c << 16) | np.uint32(0xFF000000)


def test_luminance_pass_codes():
    rng = np.random.default_rng(0)
    t = rng.integers(0, 2 ** 32, size=(64, 64), dtype=np.uint64).astype(np.uint32)
    lum, _ = orc.fxaa(t)
    assert np.array_equal(lum & 0xFFFFFF, t & 0xFFFFFF)  # k / 255 re-stored is k, for every code
    r, g, b = [((t >> (8 * k)) & 0xFF).astype(np.float64) / 255 for k in range(3)]
    y = (0.2126 * r + 0.7152 * g + 0.0722 * b) * 255
    a = (lum >> 24).astype(np.float64)
    assert np.all(np.abs(a - y) <= 0.5 + 1e-4)
    for rgb, code in (((255, 0, 0), 54), ((0, 255, 0), 182), ((0, 0, 255), 18), ((255, 255, 255), 255),
                      ((0, 0, 0), 0)):
        px = np.array([[rgb[0] | (rgb[1] << 8) | (rgb[2] << 16)]], np.uint32)
        assert int(orc.fxaa(px)[0][0, 0] >> 24) == code, rgb


def test_flat_and_low_contrast_pass_through():
    img = _gray(np.full((20, 24), 100))
    img[10, 10] = _gray(110)  # contrast 10/255 < max(0.0625, 0.125 * 110/255)
    _, out = orc.fxaa(img & 0x00FFFFFF)
    assert np.array_equal(out, img)


def _edge_codes(vertical: bool, W=40, H=30, X0=20):
    img = np.zeros((H, W), np.uint32)
    if vertical:
        img[:, X0:] = 255
    else:
        img[X0 - 5:, :] = 255  # rows from X0 - 5 up are white (row 0 is the bottom row)
    return _gray(img)


def test_straight_edges_blend_by_the_subpixel_factor():
    f = (7 / 27) ** 2  # smoothstep(0, 1, 1/3)^2: normalised centre-average contrast 1/3
    lo, hi = int(np.floor(f * 255 + 0.5)), int(np.floor((1 - f) * 255 + 0.5))
    assert (lo, hi) == (17, 238)
    for vertical in (True, False):
        img = _edge_codes(vertical)
        _, out = orc.fxaa(img)
        want = img.copy()
        if vertical:
            want[:, 19], want[:, 20] = _gray(lo), _gray(hi)
        else:
            want[14, :], want[15, :] = _gray(lo), _gray(hi)
        assert np.array_equal(out, want), vertical


def test_edge_end_walk():
    """A white region (x < X0, y < Y0) under a grey one (code 51, g = 0.2), black to its right: the
    top white row's pixel at distance D from X0 blends toward the grey by max((7/27)^2,
    0.5 - d / (d + 28.5)), d the first walk offset >= D - 1/2 (28.5 when none: the guess jump)."""
    W, H, X0, Y0, g = 80, 12, 60, 6, 51
    img = np.zeros((H, W), np.uint32)
    img[Y0:, :] = g
    img[:Y0, :X0] = 255
    img[:Y0, X0:] = 0
    _, out = orc.fxaa(_gray(img))
    offsets = np.cumsum([1, 1.5, 2, 2, 2, 2, 2, 2, 2, 4])
    for D in (2, 3, 5, 9, 12, 17, 30, 40):
        hit = offsets[offsets >= D - 0.5]
        d = hit[0] if len(hit) else 28.5
        blend = max((7 / 27) ** 2, 0.5 - d / (d + 28.5))
        v = 1 - blend * (1 - g / 255)
        assert abs(v * 255 - np.floor(v * 255) - 0.5) > 0.01  # far from a rounding tie
        want = int(np.floor(v * 255 + 0.5))
        got = int(out[Y0 - 1, X0 - D] & 0xFF)
        assert got == want, (D, got, want)


def test_half_rounding_matches_numpy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 20000).astype(np.float32),
                        (rng.normal(0, 1, 5000) * 1e-6).astype(np.float32),    # binary16 subnormals
                        (rng.normal(0, 1, 2000) * 7e4).astype(np.float32),     # overflow to infinity
                        np.float32([0.0, -0.0, 65504.0, 65520.0, 65519.996, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                                    1 + 2 ** -11, 1 + 3 * 2 ** -11, np.inf, -np.inf])])
    got = np.array([L.orc_half_rn(float(v)) for v in x], np.uint16)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def test_taa_history_store():
    v = np.array([[[0.0, 0.5, 1.0], [1.7, 20.0, -0.3], [0.2, 0.00196, 0.998]]], np.float32)
    got = orc.taa_history(v)
    h = v.astype(np.float16).astype(np.float32)
    want = (np.floor(np.clip(h, 0, 1) * np.float32(255) + np.float32(0.5)) / np.float32(255)).astype(np.float32)
    assert np.array_equal(got, want)
    assert got.max() <= 1.0 and np.all(got[0, 1, :2] == 1.0) and got[0, 1, 2] == 0.0  # HDR clamps to the RGB8 range
    # every result is a k / 255 value
    k = np.rint(got * 255)
    assert np.array_equal((k / 255).astype(np.float32), got)
