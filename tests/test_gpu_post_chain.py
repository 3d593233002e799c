"""The rest of the reference's post chain on the GPU, bit for bit the oracle:

- sgi_fxaa (fxaa_luminance.fs + fxaa_smoothing.fs, fused into one kernel)
  against the oracle's two passes, on random RGBA8 noise (edges everywhere),
  on straight and ending edges, on a ragged viewport and on the Cornell frame
  after the tonemap; with the reference's uniforms and with others;
- sgi_taa_history (PerformTaaPostFx_'s copy into the RGB8 history) on HDR
  values, and four frames of sgi_taa -> sgi_taa_history with HDR input and
  motion, against the oracle's chain;
- aliasing, NULL and non-finite arguments are refused."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import Case, require_gpu
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import FrameBuffers, GiPass

pytestmark = pytest.mark.gpu


def _gray(codes):
    c = np.asarray(codes, np.uint32)
    return c | (c << 8) | (c << 16) | np.uint32(0xFF000000)


def _fxaa_device(gi, img, params=None):
    torch = require_gpu()
    src = torch.from_numpy(img.view(np.int32).copy()).cuda()
    out = torch.zeros_like(src)
    gi.fxaa(src, out, params)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def _edges(H, W):
    img = np.zeros((H, W), np.uint32)
    img[:, W // 3:] = 200
    img[H // 2:, : W // 2] = 90
    img[5:9, 10:60] = 255
    yy, xx = np.mgrid[0:H, 0:W]
    img[(xx - 0.7 * yy) % 37 < 9] = 30  # slanted stripes: staircases and edge ends
    return _gray(img)


@pytest.mark.parametrize("which", ["noise", "edges", "ragged"])
def test_fxaa_bit_exact(which):
    require_gpu()  # torch's device before the library's context
    rng = np.random.default_rng(11)
    if which == "noise":
        H, W = 96, 160
        img = rng.integers(0, 2 ** 32, size=(H, W), dtype=np.uint64).astype(np.uint32)
    elif which == "edges":
        H, W = 120, 200
        img = _edges(H, W)
    else:
        H, W = 37, 53
        img = _edges(H, W)
        img[rng.random((H, W)) < 0.1] = _gray(255)
    gi = GiPass(W, H, 16, 12)
    got = _fxaa_device(gi, img)
    _, want = orc.fxaa(img)
    diff = got != want
    changed = int((want != (img | np.uint32(0xFF000000))).sum())
    print(f"[fxaa {which}] {W}x{H}: {int(diff.sum())} texels differ; the filter changed {changed}", flush=True)
    assert not diff.any()
    assert changed > 0


def test_fxaa_parameters():
    require_gpu()
    H, W = 64, 96
    img = _edges(H, W)
    gi = GiPass(W, H, 16, 12)
    for c, r, s in ((0.0312, 0.063, 0.75), (0.0833, 0.333, 0.0), (0.0, 0.0, 1.0)):
        got = _fxaa_device(gi, img, _abi.FxaaParams(c, r, s))
        _, want = orc.fxaa(img, c, r, s)
        assert np.array_equal(got, want), (c, r, s)


def test_fxaa_cornell_frame_after_tonemap():
    torch = require_gpu()
    case = Case(sc.CORNELL_640x360)
    w = case.w
    gi = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    gi.set_frame(case.dev.frame)
    gi.set_csm(case.dev.csm_struct())
    fb = FrameBuffers.allocate(w.width, w.height, w.num_vpls, gi.num_tiles, w.max_vpls_per_tile)
    gi.render_frame(fb.frame_io(case.dev.gbuffer(), None, case.dev.vpls, w.num_vpls, case.dev.entries,
                                case.dev.cubes(), True))
    tm = torch.zeros((w.height, w.width), dtype=torch.int32, device="cuda")
    gi.tonemap(fb.combined, tm)
    out = torch.zeros_like(tm)
    gi.fxaa(tm, out)
    torch.cuda.synchronize()
    img = tm.cpu().numpy().view(np.uint32)
    _, want = orc.fxaa(img)
    got = out.cpu().numpy().view(np.uint32)
    changed = int((want != img).sum())
    print(f"[fxaa cornell] {changed} texels changed by the filter", flush=True)
    assert np.array_equal(got, want)
    assert changed > 100
    with pytest.raises(_abi.SgiError):
        gi.fxaa(tm, tm)
    with pytest.raises(_abi.SgiError):
        gi.fxaa(tm, out, _abi.FxaaParams(float("nan"), 0.1, 1.0))


def test_taa_history_bit_exact():
    torch = require_gpu()
    H, W = 61, 87
    rng = np.random.default_rng(4)
    v = np.exp(rng.uniform(-8, 3.5, (H, W, 3))).astype(np.float32)
    v[rng.random((H, W, 3)) < 0.05] *= -1
    gi = GiPass(W, H, 16, 12)
    src = torch.from_numpy(v).cuda()
    out = torch.full_like(src, -1.0)
    gi.taa_history(src, out)
    gi.taa_history(src, src)  # in place
    torch.cuda.synchronize()
    want = orc.taa_history(v)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))
    assert np.array_equal(src.cpu().numpy().view(np.uint32), want.view(np.uint32))
    assert (v > 1).any() and want.max() == 1.0


def test_taa_chain_four_frames():
    """sgi_taa -> sgi_taa_history per frame, HDR frames moving one texel per frame with matching
    velocity: every frame's output and history equal the oracle's chain."""
    torch = require_gpu()
    H, W = 72, 128
    rng = np.random.default_rng(9)
    base = np.exp(rng.uniform(-4, 3, (H, W + 8, 3))).astype(np.float32)
    vel = np.zeros((H, W, 2), np.float32)
    vel[..., 0] = 1.0 / W  # content moves +1 texel in x per frame
    gi = GiPass(W, H, 16, 12)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    hist_d = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    out_d = torch.zeros_like(hist_d)
    vel_d = d(vel)
    hist_o = np.zeros((H, W, 3), np.float32)
    for f in range(4):
        cur = base[:, 8 - f: 8 - f + W]
        gi.taa(d(cur), hist_d, vel_d, out_d)
        gi.taa_history(out_d, hist_d)
        torch.cuda.synchronize()
        out_o = orc.taa(cur, hist_o, vel)
        hist_o = orc.taa_history(out_o)
        assert np.array_equal(out_d.cpu().numpy().view(np.uint32), out_o.view(np.uint32)), f
        assert np.array_equal(hist_d.cpu().numpy().view(np.uint32), hist_o.view(np.uint32)), f
        print(f"[taa chain] frame {f}: output and RGB8 history bit-exact; history max {hist_o.max():.3f}, "
              f"output max {out_o.max():.2f}", flush=True)
    assert out_o.max() > 1.0 and hist_o.max() == 1.0
