"""Parity of every HIP kernel with the FP32 oracle on the Cornell-box workload
(640x360, 64 VPLs, 6x64x64 FP32 cubes, 1024^2 cascades, 16x16 tiles, K = 12).

Integer outputs (visible list, per-tile lists) must be bit-exact; FP32 images
within rel-L2 1e-4 (the bound the plain form is held to)."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import Case, assert_tiles_equal, require_gpu
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import FrameBuffers, GiPass

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4


@pytest.fixture(scope="module")
def case():
    return Case(sc.CORNELL_640x360)


@pytest.fixture(scope="module")
def ora(case):
    """The oracle's whole frame on the case's inputs."""
    w, h = case.w, case.host
    upd, vis, ent = case.oracle_cull()
    s1, tiles = orc.cull_tiles(h.frame, case.hg, w.tile_size, w.max_vpls_per_tile, upd, vis)
    col_r, res_r = orc.gather(h.frame, case.hg, upd, int(vis[0]), ent, case.hcubes, mode=orc.GATHER_RANDOM)
    col_t, res_t = orc.gather(h.frame, case.hg, upd, int(vis[0]), ent, case.hcubes, tiles=tiles,
                              tile_size=w.tile_size, K=w.max_vpls_per_tile, mode=orc.GATHER_TILES)
    lit = orc.direct_light(h.frame, case.hcsm, case.hg)
    den = orc.denoise(h.frame, case.hg, None, int(vis[0]), lit, col_t, res_t)
    return dict(upd=upd, vis=vis, ent=ent, stage1=s1, tiles=tiles, col_r=col_r, res_r=res_r, col_t=col_t,
                res_t=res_t, lit=lit, den=den)


@pytest.fixture(scope="module")
def gi(case):
    w = case.w
    p = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    p.set_frame(case.dev.frame)
    p.set_csm(case.dev.csm_struct())
    return p


def test_cull_global_bit_exact(case, ora, gi):
    torch = require_gpu()
    n = case.w.num_vpls
    out = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    vis = torch.full((1 + n,), -7, dtype=torch.int32, device="cuda")
    gi.cull_global(case.dev.vpls, n, out, vis)
    torch.cuda.synchronize()
    c = int(ora["vis"][0])
    assert c > 0, "the scene should leave some VPLs visible"
    got = vis.cpu().numpy()
    assert np.array_equal(got[:1 + c], ora["vis"])
    assert np.array_equal(out.cpu().numpy()[:c].view(np.uint32), ora["upd"].view(np.uint32))


def _device_visible(case, ora):
    torch = require_gpu()
    upd = torch.from_numpy(ora["upd"]).cuda()
    vis = torch.from_numpy(ora["vis"]).cuda()
    ent = torch.from_numpy(np.ascontiguousarray(ora["ent"])).cuda()
    return upd, vis, ent


def test_cull_tiles_bit_exact(case, ora, gi):
    torch = require_gpu()
    upd, vis, _ = _device_visible(case, ora)
    tiles = torch.full((gi.num_tiles, 1 + gi.K), -1, dtype=torch.int32, device="cuda")
    s1 = torch.zeros((gi.num_tiles, 8), dtype=torch.float32, device="cuda")
    gi.cull_tiles(case.dev.gbuffer(), upd, vis, tiles, stage1=s1)
    torch.cuda.synchronize()
    assert np.array_equal(s1.cpu().numpy().view(np.uint32), ora["stage1"].view(np.uint32)), "stage-1 tile means"
    assert_tiles_equal(tiles.cpu().numpy(), ora["tiles"])
    assert ora["tiles"][:, 0].max() > 0


def test_cull_tiles_remap_to_input_indices(case, ora, gi):
    torch = require_gpu()
    upd, vis, _ = _device_visible(case, ora)
    tiles = torch.zeros((gi.num_tiles, 1 + gi.K), dtype=torch.int32, device="cuda")
    gi.cull_tiles(case.dev.gbuffer(), upd, vis, tiles, remap=1)
    torch.cuda.synchronize()
    want = ora["tiles"].copy()
    for t in range(want.shape[0]):
        n = want[t, 0]
        want[t, 1:1 + n] = ora["vis"][1 + want[t, 1:1 + n]]
    assert_tiles_equal(tiles.cpu().numpy(), want)


@pytest.mark.parametrize("mode", ["random", "tiles"])
def test_gather(case, ora, gi, mode):
    torch = require_gpu()
    upd, vis, ent = _device_visible(case, ora)
    W, H = case.w.width, case.w.height
    color = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    res = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    tiles = torch.from_numpy(ora["tiles"]).cuda() if mode == "tiles" else None
    gi.gather(case.dev.gbuffer(), upd, vis, ent, case.dev.cubes(), tiles, color, res)
    torch.cuda.synchronize()
    want_c = ora["col_r" if mode == "random" else "col_t"]
    want_r = ora["res_r" if mode == "random" else "res_t"]
    e_c = orc.rel_l2(color.cpu().numpy(), want_c)
    e_r = orc.rel_l2(res.cpu().numpy(), want_r)
    print(f"gather[{mode}] rel-L2 color {e_c:.3e} reservoir {e_r:.3e}")
    assert e_c <= REL_L2 and e_r <= REL_L2
    assert np.abs(want_r[..., :3]).sum() > 0


def test_direct_light(case, ora, gi):
    torch = require_gpu()
    W, H = case.w.width, case.w.height
    lit = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    gi.direct_light(case.dev.gbuffer(), lit)
    torch.cuda.synchronize()
    e = orc.rel_l2(lit.cpu().numpy(), ora["lit"])
    print(f"direct rel-L2 {e:.3e}")
    assert e <= REL_L2
    assert ora["lit"].sum() > 0


def test_denoise(case, ora, gi):
    torch = require_gpu()
    W, H = case.w.width, case.w.height
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    vis = torch.from_numpy(ora["vis"]).cuda()
    outs = [torch.zeros((H, W, c), dtype=torch.float32, device="cuda") for c in (3, 3, 4)]
    hist = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    gi.denoise(case.dev.gbuffer(), None, vis, t(ora["lit"]), t(ora["col_t"]), t(ora["res_t"]), None, None,
               outs[0], outs[1], outs[2], hist)
    torch.cuda.synchronize()
    comb, gio, reso, histo = ora["den"]
    for name, g, w_ in (("combined", outs[0], comb), ("gi", outs[1], gio), ("reservoir", outs[2], reso),
                        ("history", hist, histo)):
        e = orc.rel_l2(g.cpu().numpy(), w_)
        print(f"denoise {name} rel-L2 {e:.3e}")
        assert e <= REL_L2, name


@pytest.mark.parametrize("use_tiles", [False, True])
def test_render_frame_chain(case, ora, gi, use_tiles):
    """The whole frame through sgi_render_frame against the oracle's chain."""
    torch = require_gpu()
    w = case.w
    fb = FrameBuffers.allocate(w.width, w.height, w.num_vpls, gi.num_tiles, w.max_vpls_per_tile)
    g, cubes = case.dev.gbuffer(), case.dev.cubes()
    io = fb.frame_io(g, None, case.dev.vpls, w.num_vpls, case.dev.entries, cubes, use_tiles=use_tiles)
    gi.render_frame(io)
    torch.cuda.synchronize()
    h = case.host
    col, res = (ora["col_t"], ora["res_t"]) if use_tiles else (ora["col_r"], ora["res_r"])
    comb, gio, _, _ = (ora["den"] if use_tiles else
                       orc.denoise(h.frame, case.hg, None, int(ora["vis"][0]), ora["lit"], col, res))
    e = orc.rel_l2(fb.combined.cpu().numpy(), comb)
    e_gi = orc.rel_l2(fb.gi.cpu().numpy(), gio)
    print(f"frame[tiles={use_tiles}] lit rel-L2 {e:.3e}, gi rel-L2 {e_gi:.3e}")
    assert e <= REL_L2 and e_gi <= REL_L2


@pytest.mark.parametrize("mode,acc64,rej", [("all", False, 5), ("all", True, 5), ("samples", False, 5),
                                            ("samples", False, -1)])
def test_gather_ex_modes(case, ora, gi, mode, acc64, rej):
    """sgi_gather_ex: every visible VPL (FP32 or FP64 sums of the FP32 per-VPL terms, both against the
    FP32 restatement: the decisions are the same, the sums differ by rounding) and 8 hashed draws per pixel
    with the reference's 5-rejection rule and with failing draws skipped."""
    from stratusgfx_amd import _abi
    torch = require_gpu()
    upd, vis, ent = _device_visible(case, ora)
    W, H = case.w.width, case.w.height
    color = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    res = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    m = _abi.GATHER_ALL if mode == "all" else _abi.GATHER_SAMPLES
    gi.gather_ex(case.dev.gbuffer(), upd, vis, ent, case.dev.cubes(), None, color, res, mode=m, samples=8,
                 accumulate_fp64=acc64, rejection_limit=rej)
    torch.cuda.synchronize()
    h = case.host
    omode = orc.GATHER_CONVERGED if mode == "all" else orc.GATHER_SAMPLES
    oc, orr = orc.gather(h.frame, case.hg, ora["upd"], int(ora["vis"][0]), ora["ent"], case.hcubes, mode=omode,
                         precision=0, samples=8, rejection_limit=rej)
    e = orc.rel_l2(res.cpu().numpy(), orr)
    # the weight channel sums each accepted draw's reweighting factor (0.95 ... 1): one side / radius
    # or shadow decision that differs from the restatement's moves a pixel's weight by far more than
    # the rounding of the continuous factors
    dw = float(np.abs(res[..., 3].cpu().numpy().astype(np.float64) - orr[..., 3]).max())
    print(f"gather_ex[{mode}, fp64={acc64}, rejection limit {rej}] rel-L2 {e:.3e}, max weight difference {dw:.2e}")
    assert e <= REL_L2
    assert orc.rel_l2(color.cpu().numpy(), oc) <= REL_L2
    assert dw < 1e-3


def test_sampled_gather_outside_binary16(case, ora, gi):
    """The compacting gather decides each draw's side / radius test from the VPL's position and radius
    rounded to binary16 wherever the rounding cannot change it, and otherwise from the FP32 record. Four
    visible VPLs are moved to where binary16 fails: one coordinate at 7e4 and one at -7e4 (beyond its
    largest finite value, 65504: the rounded position is infinite), a radius of 1e5 (infinite when
    rounded) and a radius of -1 (a ratio test every draw fails). Against the oracle on the same records,
    16 draws per pixel with failing draws skipped: within 1e-4 rel-L2, every draw's decision equal
    (the weight channel)."""
    from stratusgfx_amd import _abi
    torch = require_gpu()
    upd_h = ora["upd"].copy()
    count = int(ora["vis"][0])
    assert count >= 4
    upd_h[0, 0] = 7.0e4
    upd_h[1, 2] = -7.0e4
    upd_h[2, 12] = 1.0e5
    upd_h[3, 12] = -1.0
    upd = torch.from_numpy(upd_h).cuda()
    vis = torch.from_numpy(ora["vis"]).cuda()
    ent = torch.from_numpy(np.ascontiguousarray(ora["ent"])).cuda()
    W, H = case.w.width, case.w.height
    color = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    res = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    gi.gather_ex(case.dev.gbuffer(), upd, vis, ent, case.dev.cubes(), None, color, res, mode=_abi.GATHER_SAMPLES,
                 samples=16, rejection_limit=-1)
    torch.cuda.synchronize()
    h = case.host
    oc, orr = orc.gather(h.frame, case.hg, upd_h, count, ora["ent"], case.hcubes, mode=orc.GATHER_SAMPLES,
                         precision=0, samples=16, rejection_limit=-1)
    got = res.cpu().numpy()
    e = orc.rel_l2(got, orr)
    dw = float(np.abs(got[..., 3].astype(np.float64) - orr[..., 3]).max())
    print(f"sampled gather with VPLs outside binary16: rel-L2 {e:.3e}, max weight difference {dw:.2e}")
    assert e <= REL_L2 and orc.rel_l2(color.cpu().numpy(), oc) <= REL_L2
    assert dw < 1e-3
