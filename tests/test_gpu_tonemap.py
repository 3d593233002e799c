"""sgi_tonemap (gammaTonemap.fs) on the GPU against the oracle.

The ACES film fit and the store are exact restatements; pow is the device
library's (within a few ulp of glibc's), so a code may differ by one only
where the oracle's value v lies within 8 ulp of a rounding tie (k + 0.5) / 255;
such codes are at most 0.1 % of the channels.  Inputs: log-uniform values over
1e-4 .. 100 with zeros at a ragged viewport, and the Cornell lit frame."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import Case, require_gpu
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import FrameBuffers, GiPass

pytestmark = pytest.mark.gpu


def _check(lit_np, got, gamma):
    want, film = orc.tonemap(lit_np, gamma)
    gc = np.stack([(got >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.int64)
    wc = np.stack([(want >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.int64)
    assert np.all(gc[..., 3] == 255)
    diff = gc[..., :3] != wc[..., :3]
    v = film.astype(np.float64)
    tie = np.abs(v * 255.0 - (np.floor(v * 255.0 + 0.5) - 0.5))
    tie = np.minimum(tie, np.abs(v * 255.0 - (np.floor(v * 255.0 + 0.5) + 0.5)))
    near = tie <= 8 * 255.0 * np.spacing(np.maximum(film, 1e-30))
    print(f"[tonemap] {int(diff.sum())} of {diff.size} channel codes differ (all within 8 ulp of a tie: "
          f"{bool(np.all(near[diff]))})", flush=True)
    assert np.all(np.abs(gc[..., :3] - wc[..., :3])[diff] == 1)
    assert np.all(near[diff])
    assert diff.sum() <= 1e-3 * diff.size


@pytest.mark.parametrize("gamma", [2.2, 1.0])
def test_tonemap_random(gamma):
    torch = require_gpu()
    H, W = 123, 317
    rng = np.random.default_rng(3)
    lit = np.exp(rng.uniform(np.log(1e-4), np.log(100.0), (H, W, 3))).astype(np.float32)
    lit[rng.random((H, W, 3)) < 0.05] = 0.0
    gi = GiPass(W, H, 16, 12)
    out = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    gi.tonemap(torch.from_numpy(lit).cuda(), out, gamma=gamma)
    torch.cuda.synchronize()
    _check(lit, out.cpu().numpy().view(np.uint32), gamma)


def test_tonemap_cornell_frame():
    torch = require_gpu()
    case = Case(sc.CORNELL_640x360)
    w = case.w
    gi = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    gi.set_frame(case.dev.frame)
    gi.set_csm(case.dev.csm_struct())
    fb = FrameBuffers.allocate(w.width, w.height, w.num_vpls, gi.num_tiles, w.max_vpls_per_tile)
    gi.render_frame(fb.frame_io(case.dev.gbuffer(), None, case.dev.vpls, w.num_vpls, case.dev.entries,
                                case.dev.cubes(), True))
    out = torch.zeros((w.height, w.width), dtype=torch.int32, device="cuda")
    gi.tonemap(fb.combined, out)
    torch.cuda.synchronize()
    _check(fb.combined.cpu().numpy(), out.cpu().numpy().view(np.uint32), 2.2)
    with pytest.raises(_abi.SgiError):
        gi.tonemap(fb.combined, out, gamma=0.0)
