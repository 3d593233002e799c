"""The oracle's tonemap (gammaTonemap.fs) pinned by closed forms of the ACES
film fit f(x) = x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14): f(0) = 0,
f(1) = 2.54 / 3.16, f(x) < 1 exactly below the root (0.56 + sqrt(0.3584)) / 0.16
of 0.08 x^2 - 0.56 x - 0.14 and saturated at 1 above it, a small negative
input clamps to 0 and a large one (f(-5) = 62.6 / 57.94 > 1) saturates; the gamma step is v^(1 / gamma) (numpy's float64 pow on the
gamma-1 output, within float32 rounding); codes are floor(v * 255 + 0.5),
non-decreasing in x, r, g, b in bytes 0..2 and alpha 255."""
import numpy as np

import oracle as orc


def _run(x, gamma):
    lit = np.repeat(np.asarray(x, np.float32)[None, :, None], 3, axis=2)
    out, film = orc.tonemap(lit, gamma)
    return out[0], film[0, :, 0]


def test_aces_closed_forms():
    _, v = _run([0.0, 1.0, -0.01, -5.0], 1.0)
    assert v[0] == 0.0 and v[2] == 0.0 and v[3] == 1.0
    assert abs(v[1] - 2.54 / 3.16) <= 2 * np.spacing(np.float32(0.8))
    root = (0.56 + np.sqrt(0.3584)) / 0.16
    _, v = _run([root * 0.98, root * 0.999, root * 1.001, root * 1.05, 1e4], 1.0)
    assert v[0] < 1.0 and v[1] < 1.0 and np.all(v[2:] == 1.0)


def test_gamma_and_codes():
    x = np.concatenate([[0.0], np.geomspace(1e-4, 50.0, 4000)]).astype(np.float32)
    c1, v1 = _run(x, 1.0)
    c22, v22 = _run(x, 2.2)
    inv_gamma = np.float64(np.float32(1.0) / np.float32(2.2))  # the shader's 1.0 / gamma in FP32
    want = np.power(v1.astype(np.float64), inv_gamma)
    assert np.all(np.abs(v22 - want) <= 4 * np.spacing(np.maximum(want, 1e-30).astype(np.float32)))
    for c, v in ((c1, v1), (c22, v22)):
        codes = c & 0xFF
        np.testing.assert_array_equal(codes, np.floor(v.astype(np.float64) * 255.0 + 0.5).astype(np.uint32))
        assert np.all(np.diff(codes.astype(np.int64)) >= 0)
        assert np.all((c >> 24) == 255) and np.all(((c >> 8) & 0xFF) == codes) and np.all(((c >> 16) & 0xFF) == codes)
    assert c22[0] & 0xFF == 0 and c22[-1] & 0xFF == 255
    assert np.all(v22[1:] >= v1[1:])  # gamma 2.2 brightens every value in (0, 1)


def test_channel_order():
    lit = np.array([[[0.0, 1.0, 100.0]]], np.float32)
    out, _ = orc.tonemap(lit, 2.2)
    t = int(out[0, 0])
    assert t & 0xFF == 0 and (t >> 16) & 0xFF == 255 and 0 < (t >> 8) & 0xFF < 255 and t >> 24 == 255
