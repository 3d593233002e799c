"""Batched views (sgi_views_*): four views of one city block at different
viewport sizes, rendered with one launch per stage over all views, must
equal the same views rendered one by one with sgi_render_frame (tile lists)
bit for bit; the captured graph must reproduce them on every replay."""
import dataclasses

import numpy as np
import pytest

from gpu_util import assert_tiles_equal, learned_single, require_gpu, shipped_nets
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import FrameBuffers, GiPass, Views, view_io

pytestmark = pytest.mark.gpu

SIZES = [(320, 180), (256, 256), (400, 224), (96, 160)]
TILE, K = 16, 32


@pytest.fixture(scope="module")
def batch():
    torch = require_gpu()
    base = dataclasses.replace(sc.CITY_1080P, name="city-views", num_vpls=256, cube_res=64, csm_res=512,
                               max_vpls_per_tile=K)
    views = []
    for i, (w, h) in enumerate(SIZES):
        wl = dataclasses.replace(base, width=w, height=h)
        s = sc.Scene(wl)
        v = s.view
        v.yaw_deg += 25.0 * i
        v.pitch_deg -= 2.0 * i
        s.set_view(v)
        views.append((wl, s, sc.generate_device(s)))
        print(f"[case] view {i}: {w}x{h}", flush=True)
    shared = views[0][2]
    f32 = dict(dtype=torch.float32, device="cuda")
    refs, bufs = [], []
    for wl, s, inp in views:
        gi = GiPass(wl.width, wl.height, TILE, K)
        gi.set_frame(inp.frame)
        gi.set_csm(inp.csm_struct())
        fb = FrameBuffers.allocate(wl.width, wl.height, wl.num_vpls, gi.num_tiles, K)
        g, cubes = inp.gbuffer(), shared.cubes()
        io = fb.frame_io(g, None, shared.vpls, wl.num_vpls, shared.entries, cubes, use_tiles=True)
        gi.render_frame(io)
        torch.cuda.synchronize()
        refs.append({k: getattr(fb, k).clone() for k in ("tiles", "direct", "reservoir", "combined", "gi",
                                                        "reservoir_out", "history")})
        H, W = wl.height, wl.width
        bufs.append(dict(tiles=torch.full((gi.num_tiles, 1 + K), -5, dtype=torch.int32, device="cuda"),
                         direct=torch.zeros((H, W, 3), **f32), color=torch.zeros((H, W, 3), **f32),
                         reservoir=torch.zeros((H, W, 4), **f32), combined=torch.zeros((H, W, 3), **f32),
                         gi=torch.zeros((H, W, 3), **f32), reservoir_out=torch.zeros((H, W, 4), **f32),
                         history=torch.zeros((H, W), **f32)))
    ios = [view_io(inp, **b) for (_, _, inp), b in zip(views, bufs)]
    total = sum(w * h for w, h in SIZES)
    vw = Views(len(SIZES), total, TILE, K)
    vw.set(ios, shared.vpls, base.num_vpls, shared.entries, shared.cubes())
    return dict(views=views, refs=refs, bufs=bufs, vw=vw, ios=ios, shared=shared)


def _check(batch, what):
    for i, (ref, b) in enumerate(zip(batch["refs"], batch["bufs"])):
        for k, want in ref.items():
            got = b[k]
            if k == "tiles":  # slots past a tile's count are not written
                assert_tiles_equal(got.cpu().numpy(), want.cpu().numpy())
            else:
                assert torch_equal(got, want), f"{what}: view {i} {k}"


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a, b))


def test_batched_views_equal_single_views(batch):
    torch = require_gpu()
    batch["vw"].run()
    torch.cuda.synchronize()
    _check(batch, "run")
    assert any(int(r["tiles"][:, 0].max()) > 0 for r in batch["refs"])


def test_graph_replays(batch):
    torch = require_gpu()
    s = torch.cuda.Stream()
    for b in batch["bufs"]:
        for t in b.values():
            t.zero_()
    torch.cuda.synchronize()
    batch["vw"].capture(s.cuda_stream)
    for rep in range(3):
        with torch.cuda.stream(s):
            batch["vw"].replay(s.cuda_stream)
        s.synchronize()
        _check(batch, f"replay {rep}")
        # the program's own work between replays
        x = torch.randn(1024, device="cuda").sum()
        torch.cuda.synchronize()
        del x


def test_replay_refuses_a_changed_batch(batch):
    require_gpu()
    vw, ios, shared = batch["vw"], batch["ios"], batch["shared"]
    vw.set(ios[:2], shared.vpls, 256, shared.entries, shared.cubes())
    with pytest.raises(_abi.SgiError):
        vw.replay()
    vw.set(ios, shared.vpls, 256, shared.entries, shared.cubes())
    vw.replay()


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
def test_learned_batch_equals_single_views(batch, precision):
    """sgi_views_set_learned: every view's direct light, sampled gather (12 draws) and learned
    denoise (shipped weights and threshold) equal the view rendered alone, bit for bit, on a run and
    on graph replays; a replay after the form changed is refused."""
    torch = require_gpu()
    wts, f8, thr = shipped_nets()
    net = wts.net() if precision == "bf16" else f8.net()
    vw, ios, shared = batch["vw"], batch["ios"], batch["shared"]
    vw.set(ios, shared.vpls, 256, shared.entries, shared.cubes())
    vw.set_learned(12, net, thr)
    keys = ("direct", "color", "reservoir", "combined", "gi", "reservoir_out", "history")
    refs = [learned_single(inp, shared, 256, net, thr, 12, TILE, K) for (_, _, inp) in batch["views"]]
    s = torch.cuda.Stream()
    for rep in range(3):
        for b in batch["bufs"]:
            for t in b.values():
                t.zero_()
        torch.cuda.synchronize()
        if rep == 0:
            vw.run()
        else:
            if rep == 1:
                vw.capture(s.cuda_stream)
            with torch.cuda.stream(s):
                vw.replay(s.cuda_stream)
            s.synchronize()
        torch.cuda.synchronize()
        for i, (ref, b) in enumerate(zip(refs, batch["bufs"])):
            for k in keys:
                assert torch.equal(b[k], ref[k]), f"{precision} pass {rep}: view {i} {k}"
    assert any(float(r["gi"].abs().sum()) > 0 for r in refs)
    vw.set_learned(0)  # back to the tile-list form: the graph captured with the learned form is refused
    with pytest.raises(_abi.SgiError):
        vw.replay(s.cuda_stream)
