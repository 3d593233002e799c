"""The packed-FP16 shading form of the gather (SGI_PRECISION_FP16: the tile-list
and reference estimators of sgi_gather, and the fixed-count sampled estimator
of sgi_gather_ex) against the FP64 twin of the oracle on the same inputs and
tile lists.

The FP16 form keeps the FP32 accept / shadow decisions, so the colour output
and the sample weights (reservoir.w) must equal the FP32 GPU path bit for bit;
the shaded radiance (reservoir.rgb) carries FP16 rounding.  Its error is stated
per output channel: rel-L2 against the FP64 oracle, rel-L2 against the FP32 GPU
path (the FP16 arithmetic alone) and the 99th percentile of the per-pixel
relative error.  Bounds: rel-L2 <= 1e-2 per channel, p99 <= 1e-2 (FP16 has an
11-bit significand, ~5e-4 relative per rounding, and the shading chains about
twenty roundings per VPL).  The random mode sums one to six VPLs per pixel, so
a single badly conditioned VPL is not averaged away there: its case is a light
opposite a grazing viewer (V.Ls near -1), where 1 + V.Ls cancels; the FP16
form takes N.Ls, 1 + V.Ls and N.V + N.Ls from FP32 for that reason."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import Case, require_gpu
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import GiPass

pytestmark = pytest.mark.gpu

REL_L2_FP16 = 1e-2
P99_FP16 = 1e-2
CITY_BANDS = [(0, 8), (536, 552), (1064, 1080)]


def _per_channel(got, want, ref32):
    """rel-L2 vs FP64, rel-L2 vs the FP32 GPU path and the p99 relative error, per channel."""
    out = []
    for c, name in enumerate("RGB"):
        g, w, r = got[..., c].astype(np.float64), want[..., c], ref32[..., c].astype(np.float64)
        rel = np.abs(g - w) / np.maximum(np.abs(w), 1e-3 * max(np.abs(w).max(), 1e-30))
        out.append((name, orc.rel_l2(g, w), orc.rel_l2(g, r), float(np.quantile(rel, 0.99))))
    return out


SAMPLES = 8


def _run(gi, case, upd, vis, ent, tiles, precision, mode):
    torch = require_gpu()
    H, W = case.w.height, case.w.width
    col = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    res = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    if mode == "samples":  # the fixed-count estimator (the learned form's input), pairs of draws in FP16
        gi.gather_ex(case.dev.gbuffer(), upd, vis, ent, case.dev.cubes(), None, col, res, mode=_abi.GATHER_SAMPLES,
                     samples=SAMPLES, rejection_limit=-1, precision=precision)
    else:
        gi.gather(case.dev.gbuffer(), upd, vis, ent, case.dev.cubes(), tiles, col, res, precision=precision)
    torch.cuda.synchronize()
    return col.cpu().numpy(), res.cpu().numpy()


def _device_lists(case, gi):
    torch = require_gpu()
    w = case.w
    n = w.num_vpls
    upd = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    vis = torch.zeros((1 + n,), dtype=torch.int32, device="cuda")
    gi.cull_global(case.dev.vpls, n, upd, vis)
    ent = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    gi.visible_atlas_entries(case.dev.entries, vis, n, ent)
    tiles = torch.zeros((gi.num_tiles, 1 + gi.K), dtype=torch.int32, device="cuda")
    gi.cull_tiles(case.dev.gbuffer(), upd, vis, tiles)
    torch.cuda.synchronize()
    return upd, vis, ent, tiles


def _check(case, gi, mode, bands):
    w, h = case.w, case.host
    upd, vis, ent, tiles = _device_lists(case, gi)
    t = tiles if mode == "tiles" else None
    c32, r32 = _run(gi, case, upd, vis, ent, t, _abi.PRECISION_FP32, mode)
    c16, r16 = _run(gi, case, upd, vis, ent, t, _abi.PRECISION_FP16, mode)
    m = r32[..., :3] > 1e-3 * r32[..., :3].max()
    q = np.quantile(r16[..., :3][m] / r32[..., :3][m], [0.0, 0.01, 0.5, 0.99, 1.0])
    print(f"[{w.name} {mode}] FP16 / FP32 ratio quantiles 0, .01, .5, .99, 1: {np.array2string(q, precision=4)}")
    assert np.array_equal(c16, c32), "colour output"
    assert np.array_equal(r16[..., 3], r32[..., 3]), "sample weights"
    assert np.all(np.isfinite(r16))
    o_upd, o_vis, o_ent = case.oracle_cull()
    o_tiles = tiles.cpu().numpy() if mode == "tiles" else None
    got, want, ref = [], [], []
    for y0, y1 in bands:
        o_mode = {"tiles": orc.GATHER_TILES, "random": orc.GATHER_RANDOM, "samples": orc.GATHER_SAMPLES}[mode]
        _, orr = orc.gather(h.frame, case.hg, o_upd, int(o_vis[0]), o_ent, case.hcubes, tiles=o_tiles,
                            tile_size=w.tile_size, K=w.max_vpls_per_tile, mode=o_mode, precision=1, y0=y0, y1=y1,
                            samples=SAMPLES, rejection_limit=-1)
        got.append(r16[y0:y1, :, :3])
        want.append(orr[..., :3])
        ref.append(r32[y0:y1, :, :3])
    got, want, ref = (np.concatenate(a) for a in (got, want, ref))
    assert np.abs(want).sum() > 0
    e32 = orc.rel_l2(ref.astype(np.float64), want)
    print(f"[{w.name} {mode}] FP32 path vs FP64 oracle rel-L2 {e32:.3e}")
    for name, e64, e_32, p99 in _per_channel(got, want, ref):
        print(f"[{w.name} {mode}] FP16 reservoir.{name}: rel-L2 vs FP64 {e64:.3e}, vs FP32 path {e_32:.3e}, "
              f"p99 rel {p99:.3e}")
        assert e64 <= REL_L2_FP16 and e_32 <= REL_L2_FP16 and p99 <= P99_FP16, name


@pytest.fixture(scope="module")
def cornell():
    return Case(sc.CORNELL_640x360)


@pytest.fixture(scope="module")
def city():
    return Case(sc.CITY_1080P)


def _gi(case):
    w = case.w
    p = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    p.set_frame(case.dev.frame)
    p.set_csm(case.dev.csm_struct())
    return p


@pytest.mark.parametrize("mode", ["tiles", "random", "samples"])
def test_fp16_cornell(cornell, mode):
    _check(cornell, _gi(cornell), mode, [(0, cornell.w.height)])


@pytest.mark.parametrize("mode", ["tiles", "samples"])
def test_fp16_city_bands(city, mode):
    _check(city, _gi(city), mode, CITY_BANDS)


def test_fp16_refused_outside_the_fixed_count_mode(cornell):
    require_gpu()
    gi = _gi(cornell)
    upd, vis, ent, _ = _device_lists(cornell, gi)
    with pytest.raises(_abi.SgiError):
        _run_ex_all_fp16(gi, cornell, upd, vis, ent)


def _run_ex_all_fp16(gi, case, upd, vis, ent):
    torch = require_gpu()
    H, W = case.w.height, case.w.width
    col = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    res = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    gi.gather_ex(case.dev.gbuffer(), upd, vis, ent, case.dev.cubes(), None, col, res, mode=_abi.GATHER_ALL,
                 precision=_abi.PRECISION_FP16)
