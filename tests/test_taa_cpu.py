"""The oracle's TAA (taa.fs with aa_common.glsl's tone curve), pinned by what
its semantics fix: with no motion and history equal to the frame it returns
the frame (to rounding); history outside the range of the current texel and
its four neighbours (a plus, not the 3x3 square) is clamped back to it, so a
constant frame ignores any history; inside the range the result is
0.1 current + 0.9 history; a one-texel velocity (u - 1/W) reads exactly the
history texel to the left, so a frame moved right by one texel over its
history reproduces the frame, and the opposite sign does not."""
import numpy as np

import oracle as orc


def test_static_frame_is_kept():
    rng = np.random.default_rng(0)
    cur = rng.uniform(0.0, 4.0, (16, 24, 3)).astype(np.float32)
    out = orc.taa(cur, cur)
    np.testing.assert_allclose(out, cur, rtol=2e-6)


def test_history_clamped_to_plus_neighbourhood():
    cur = np.full((9, 9, 3), 0.5, np.float32)
    out = orc.taa(cur, np.full_like(cur, 3.0))
    np.testing.assert_allclose(out, cur, rtol=2e-6)                 # clamped to the constant frame
    cur[4, 4] = 2.0                                                   # one bright texel
    hist = np.full_like(cur, 1.0)
    out = orc.taa(cur, hist)
    for (y, x) in [(3, 4), (5, 4), (4, 3), (4, 5)]:                   # its 4 neighbours see range [0.5, 2]
        np.testing.assert_allclose(out[y, x], 0.1 * 0.5 + 0.9 * 1.0, rtol=1e-5)
    for (y, x) in [(3, 3), (5, 5), (3, 5), (5, 3), (0, 0)]:           # diagonals and far texels do not
        np.testing.assert_allclose(out[y, x], 0.5, rtol=1e-5)
    np.testing.assert_allclose(out[4, 4], 0.1 * 2.0 + 0.9 * 1.0, rtol=1e-5)


def test_velocity_reads_the_moved_texel():
    rng = np.random.default_rng(1)
    H, W = 8, 64
    hist = rng.uniform(0.0, 3.0, (H, W, 3)).astype(np.float32)
    cur = np.concatenate([hist[:, :1], hist[:, :-1]], axis=1)        # cur[x] = hist[x - 1]
    vel = np.zeros((H, W, 2), np.float32)
    vel[..., 0] = 1.0 / W
    out = orc.taa(cur, hist, vel)
    np.testing.assert_allclose(out, cur, rtol=2e-6)
    wrong = orc.taa(cur, hist, -vel)
    assert np.abs(wrong - cur).max() > 0.1
