"""sgi_taa (taa.fs) on the GPU: bit for bit the oracle's restatement on random
frames, histories and velocities (some reaching past the edges, where the
bilinear history taps clamp), and without a velocity plane; an output that
aliases an input is refused."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import require_gpu
from stratusgfx_amd import _abi
from stratusgfx_amd.gi import GiPass

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_velocity", [True, False])
def test_taa_bit_exact(with_velocity):
    torch = require_gpu()
    H, W = 97, 203
    rng = np.random.default_rng(5)
    cur = np.exp(rng.uniform(-6, 3, (H, W, 3))).astype(np.float32)
    hist = np.exp(rng.uniform(-6, 3, (H, W, 3))).astype(np.float32)
    vel = (rng.standard_normal((H, W, 2)) * 0.02).astype(np.float32) if with_velocity else None
    gi = GiPass(W, H, 16, 12)
    d = lambda a: None if a is None else torch.from_numpy(a).cuda()  # noqa: E731
    out = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    gi.taa(d(cur), d(hist), d(vel), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = orc.taa(cur, hist, vel)
    bad = got.view(np.uint32) != want.view(np.uint32)
    print(f"[taa] {W}x{H} velocity={with_velocity}: {int(bad.sum())} of {bad.size} values differ; "
          f"max rel {float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-30))):.2e}", flush=True)
    assert not bad.any()
    with pytest.raises(_abi.SgiError):
        gi.taa(out, d(hist), None, out)
