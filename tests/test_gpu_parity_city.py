"""Parity on the procedural city block at 1920x1080 (1024 VPLs, 6x256x256 FP16
cubes, 16x16 tiles, K = 64) — the benchmark workload.

Lists are compared bit-exactly on every tile; the gather and direct light are
compared with the oracle on bands of rows (the oracle takes minutes for the
whole frame), the denoiser on the whole frame; size-independent invariants
are checked on the whole frame."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import Case, assert_tiles_equal, require_gpu
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import GiPass

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4
BANDS = [(0, 8), (536, 552), (1064, 1080)]


@pytest.fixture(scope="module")
def case():
    return Case(sc.CITY_1080P)


@pytest.fixture(scope="module")
def gi(case):
    w = case.w
    p = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    p.set_frame(case.dev.frame)
    p.set_csm(case.dev.csm_struct())
    return p


@pytest.fixture(scope="module")
def run(case, gi):
    """The GPU's culling and gathers on the case, plus the oracle's culling."""
    torch = require_gpu()
    w = case.w
    n = w.num_vpls
    upd = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    vis = torch.zeros((1 + n,), dtype=torch.int32, device="cuda")
    gi.cull_global(case.dev.vpls, n, upd, vis)
    ent = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    gi.visible_atlas_entries(case.dev.entries, vis, n, ent)
    tiles = torch.zeros((gi.num_tiles, 1 + gi.K), dtype=torch.int32, device="cuda")
    s1 = torch.zeros((gi.num_tiles, 8), dtype=torch.float32, device="cuda")
    g = case.dev.gbuffer()
    gi.cull_tiles(g, upd, vis, tiles, stage1=s1)
    H, W = w.height, w.width
    out = {}
    for mode, t in (("tiles", tiles), ("random", None)):
        col = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
        res = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        gi.gather(g, upd, vis, ent, case.dev.cubes(), t, col, res)
        out[mode] = (col, res)
    lit = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    gi.direct_light(g, lit)
    torch.cuda.synchronize()
    o_upd, o_vis, o_ent = case.oracle_cull()
    o_s1, o_tiles = orc.cull_tiles(case.host.frame, case.hg, w.tile_size, w.max_vpls_per_tile, o_upd, o_vis)
    return dict(upd=upd, vis=vis, ent=ent, tiles=tiles, s1=s1, out=out, lit=lit, o_upd=o_upd, o_vis=o_vis,
                o_ent=o_ent, o_s1=o_s1, o_tiles=o_tiles)


def test_cull_global_bit_exact(run):
    c = int(run["o_vis"][0])
    assert c > 100
    assert np.array_equal(run["vis"].cpu().numpy()[:1 + c], run["o_vis"])
    assert np.array_equal(run["upd"].cpu().numpy()[:c].view(np.uint32), run["o_upd"].view(np.uint32))
    assert np.array_equal(run["ent"].cpu().numpy()[:c], run["o_ent"])


def test_cull_tiles_bit_exact_every_tile(run):
    assert np.array_equal(run["s1"].cpu().numpy().view(np.uint32), run["o_s1"].view(np.uint32))
    assert_tiles_equal(run["tiles"].cpu().numpy(), run["o_tiles"])
    counts = run["o_tiles"][:, 0]
    assert counts.max() == 64, "the city should fill some tiles to K"


def test_tile_lists_invariants(case, run):
    """Size-independent checks on the GPU lists: sorted by ratio, within radius, facing the tile plane."""
    tiles = run["tiles"].cpu().numpy()
    s1 = run["s1"].cpu().numpy()
    upd = run["upd"].cpu().numpy()
    c = int(run["vis"][0].item())
    for t in range(0, tiles.shape[0], 7):
        n = tiles[t, 0]
        idx = tiles[t, 1:1 + n]
        assert np.all((idx >= 0) & (idx < c)) and len(set(idx.tolist())) == n
        P, N = s1[t, :3].astype(np.float32), s1[t, 4:7].astype(np.float32)
        lmf = (upd[idx, :3] - P).astype(np.float32)
        ratio = np.sqrt((lmf * lmf).sum(1)) / upd[idx, 12]
        assert np.all(np.diff(ratio) >= -1e-6) and np.all(ratio <= 1.0 + 1e-6)
        assert np.all((lmf * N).sum(1) >= -1e-4)


@pytest.mark.parametrize("mode", ["tiles", "random"])
def test_gather_bands(case, run, mode):
    w, h = case.w, case.host
    col, res = (a.cpu().numpy() for a in run["out"][mode])
    tiles = run["o_tiles"] if mode == "tiles" else None
    worst = 0.0
    for y0, y1 in BANDS:
        oc, orr = orc.gather(h.frame, case.hg, run["o_upd"], int(run["o_vis"][0]), run["o_ent"], case.hcubes,
                             tiles=tiles, tile_size=w.tile_size, K=w.max_vpls_per_tile,
                             mode=orc.GATHER_TILES if tiles is not None else orc.GATHER_RANDOM, y0=y0, y1=y1)
        e = max(orc.rel_l2(col[y0:y1], oc), orc.rel_l2(res[y0:y1], orr))
        worst = max(worst, e)
        print(f"gather[{mode}] rows {y0}..{y1} rel-L2 {e:.3e}")
    assert worst <= REL_L2
    # whole-frame invariants: colour = albedo + 1e-5, weights >= 1, finite
    alb = np.maximum(h.albedo[..., :3], 0.01) + np.float32(1e-5)
    assert np.array_equal(col, alb.astype(np.float32))
    assert np.all(res[..., 3] >= 1.0) and np.all(np.isfinite(res))


def test_direct_light_bands(case, run):
    lit = run["lit"].cpu().numpy()
    for y0, y1 in BANDS:
        o = orc.direct_light(case.host.frame, case.hcsm, case.hg, y0, y1)
        e = orc.rel_l2(lit[y0:y1], o)
        print(f"direct rows {y0}..{y1} rel-L2 {e:.3e}")
        assert e <= REL_L2
    assert np.all(np.isfinite(lit)) and lit.max() > 0


def test_denoise_whole_frame(case, run, gi):
    torch = require_gpu()
    W, H = case.w.width, case.w.height
    col, res = run["out"]["tiles"]
    lit = run["lit"]
    outs = [torch.zeros((H, W, c), dtype=torch.float32, device="cuda") for c in (3, 3, 4)]
    hist = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    gi.denoise(case.dev.gbuffer(), None, run["vis"], lit, col, res, None, None, outs[0], outs[1], outs[2], hist)
    torch.cuda.synchronize()
    comb, gio, _, histo = orc.denoise(case.host.frame, case.hg, None, int(run["o_vis"][0]), lit.cpu().numpy(),
                                      col.cpu().numpy(), res.cpu().numpy())
    e = orc.rel_l2(outs[0].cpu().numpy(), comb)
    e_gi = orc.rel_l2(outs[1].cpu().numpy(), gio)
    print(f"denoise 1080p rel-L2 combined {e:.3e} gi {e_gi:.3e}")
    assert e <= REL_L2 and e_gi <= REL_L2
    assert np.array_equal(hist.cpu().numpy(), histo)
