"""The 8-frame path at 3840x2160 with 4096 VPLs: a camera dollying 0.5 m per
frame and turning 2 degrees per frame over the city block, every frame
rendered through sgi_render_frame (tile lists, gather, a-trous + temporal)
with the previous frame's G-buffer, GI and history, the G-buffer velocity
taken from the camera motion.

Per frame: the visible list and every tile list bit-exact against the
oracle; the gather and direct light on the whole frame within 1e-4 rel-L2
(the oracle's rows on up to 16 host threads).
On frames 1 and 7 the whole-frame denoiser (spatial passes, reprojection,
history rule, combine) against the oracle run on the GPU's own inputs of
that frame: outputs within 1e-4 rel-L2, history bit-exact.  Across the
sequence: histories grow where the reprojection holds, everything finite.

The learned form along the same path: 16 draws per pixel, the network with
FP8 (E4M3) inference and with BF16, temporal accumulation from the previous
frame's G-buffer, GI and history; relMSE against each frame's converged render
(every visible VPL, FP64 sums) reported per frame beside the noisy input's
and the a-trous form's on the same input (with its own temporal chain): the
FP8 network's within 10 % of the BF16 network's, and both below the a-trous
form's, on every frame.  With the shipped fallback threshold (the network's
error channel predicting the ratio of its error to the a-trous filter's),
tiles of the path fall back to the a-trous filter (the threshold keeps the
fallback under 1 % of held-out tiles, so in distribution it is rare; this path
is outside it), never every tile, and each network's relMSE stays within 0.1 %
of its own relMSE without the fallback on every frame (the tolerance the
threshold was chosen with on held-out blocks)."""
import ctypes
import dataclasses
import json
import math
import os

import numpy as np
import pytest

import oracle as orc
from gpu_util import assert_tiles_equal, require_gpu
from stratusgfx_amd import _abi, kpcn, kpcn_fp8, quantize_kpcn
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import FrameBuffers, GiPass

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4
FRAMES = 8
THREADS = min(16, os.cpu_count() or 1)
DENOISE_FRAMES = (1, 7)
W4K = sc.Workload("city-3840x2160-4096vpl", _abi.SCENE_CITY, 3, 3840, 2160, 4096, 256, _abi.TEXEL_F16, 1024, 16,
                  64)


def _step(v: _abi.View) -> _abi.View:
    """0.5 m along the view direction (horizontal part), then 2 degrees of yaw."""
    n = _abi.View()
    ctypes.memmove(ctypes.byref(n), ctypes.byref(v), ctypes.sizeof(v))
    yaw = math.radians(v.yaw_deg)
    n.eye[0] = v.eye[0] + 0.5 * math.sin(yaw)
    n.eye[2] = v.eye[2] - 0.5 * math.cos(yaw)
    n.yaw_deg = v.yaw_deg + 2.0
    return n


def _copy(st):
    c = type(st)()
    ctypes.memmove(ctypes.byref(c), ctypes.byref(st), ctypes.sizeof(st))
    return c


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    orc.set_threads(THREADS)
    yield
    orc.set_threads(1)


@pytest.fixture(scope="module")
def seq():
    torch = require_gpu()
    w = W4K
    scene = sc.Scene(w)
    shared = sc.generate_device(scene)  # frame 0's view; VPL records and cubes serve every frame
    # frame 0 keeps its own copies of the frame / cascade uniforms (the scene's change with each view)
    shared = dataclasses.replace(shared, frame=_copy(scene.frame), csm=_copy(scene.csm))
    hcubes_host = sc.to_host(dataclasses.replace(shared, depth=None, normal=None, albedo=None, refl=None, rma=None,
                                                 velocity=None, ids=None, csm_depth=None))
    frames = []
    views = [scene.view]
    prev_fb = prev_g = None
    for k in range(FRAMES):
        if k > 0:
            v = _step(views[-1])
            scene.set_view(v, views[-1])
            views.append(v)
            inp = sc.generate_view_device(scene, shared)
        else:
            inp = shared
        gi = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
        gi.set_frame(inp.frame)
        gi.set_csm(inp.csm_struct())
        fb = FrameBuffers.allocate(w.width, w.height, w.num_vpls, gi.num_tiles, w.max_vpls_per_tile)
        g = inp.gbuffer()
        io = fb.frame_io(g, prev_g, shared.vpls, w.num_vpls, shared.entries, shared.cubes(), True,
                         prev_gi=prev_fb.gi if prev_fb is not None else None,
                         prev_history=prev_fb.history if prev_fb is not None else None)
        gi.render_frame(io)
        torch.cuda.synchronize()
        print(f"[case] 4k frame {k}: eye ({views[-1].eye[0]:.2f}, {views[-1].eye[2]:.2f}) yaw {views[-1].yaw_deg:.1f}",
              flush=True)
        frames.append(dict(inp=inp, gb=g, fb=fb, gi=gi))
        prev_fb, prev_g = fb, g
    return dict(frames=frames, shared=shared, hcubes=hcubes_host)


def _host(inp):
    return sc.to_host(dataclasses.replace(inp, atlases=[]))


@pytest.mark.parametrize("k", range(FRAMES))
def test_frame_lists_gather_direct(seq, k):
    w = W4K
    f = seq["frames"][k]
    h = _host(f["inp"])
    hg = h.gbuffer()
    hcsm = h.csm_struct()
    upd, vis = orc.cull_global(h.frame, hcsm, _vpls(seq))
    fb = f["fb"]
    c = int(vis[0])
    assert np.array_equal(fb.visible.cpu().numpy()[:1 + c], vis)
    _, o_tiles = orc.cull_tiles(h.frame, hg, w.tile_size, w.max_vpls_per_tile, upd, vis)
    assert_tiles_equal(fb.tiles.cpu().numpy(), o_tiles)
    ent = _entries(seq)[vis[1:]]
    col, res = fb.color.cpu().numpy(), fb.reservoir.cpu().numpy()
    lit = fb.direct.cpu().numpy()
    oc, orr = orc.gather(h.frame, hg, upd, c, ent, seq["hcubes"].cubes(), tiles=o_tiles, tile_size=w.tile_size,
                         K=w.max_vpls_per_tile, mode=orc.GATHER_TILES)
    e = max(orc.rel_l2(col, oc), orc.rel_l2(res, orr))
    od = orc.direct_light(h.frame, hcsm, hg)
    ed = orc.rel_l2(lit, od)
    print(f"frame {k} 3840x2160: gather rel-L2 {e:.3e}, direct rel-L2 {ed:.3e}")
    assert e <= REL_L2 and ed <= REL_L2
    assert np.all(np.isfinite(fb.gi.cpu().numpy())) and np.all(np.isfinite(fb.combined.cpu().numpy()))


def _vpls(seq):
    if "hvpls" not in seq:
        seq["hvpls"] = seq["shared"].vpls.cpu().numpy()
        seq["hent"] = seq["shared"].entries.cpu().numpy()
    return seq["hvpls"]


def _entries(seq):
    _vpls(seq)
    return seq["hent"]


@pytest.mark.parametrize("k", DENOISE_FRAMES)
def test_frame_denoise_whole_frame(seq, k):
    f, p = seq["frames"][k], seq["frames"][k - 1]
    h, hp = _host(f["inp"]), _host(p["inp"])
    fb, pfb = f["fb"], p["fb"]
    comb, gio, reso, histo = orc.denoise(h.frame, h.gbuffer(), hp.gbuffer(), int(fb.visible[0].item()),
                                         fb.direct.cpu().numpy(), fb.color.cpu().numpy(), fb.reservoir.cpu().numpy(),
                                         pfb.gi.cpu().numpy(), pfb.history.cpu().numpy())
    e = (orc.rel_l2(fb.combined.cpu().numpy(), comb), orc.rel_l2(fb.gi.cpu().numpy(), gio),
         orc.rel_l2(fb.reservoir_out.cpu().numpy(), reso))
    hist = fb.history.cpu().numpy()
    print(f"frame {k} denoise rel-L2 combined {e[0]:.3e} gi {e[1]:.3e} reservoir {e[2]:.3e}; "
          f"history > 1 on {(hist > 1).mean():.3f} of pixels")
    assert max(e) <= REL_L2
    assert np.array_equal(hist, histo)


def test_history_grows_over_the_sequence(seq):
    hs = [f["fb"].history.cpu().numpy() for f in seq["frames"]]
    assert np.all(hs[0] == 1.0)
    means = [float(h.mean()) for h in hs]
    print("mean history per frame", np.round(means, 3))
    # the camera moves slowly: most of the frame reprojects, so the mean history rises
    assert means[-1] > 2.0 and all(b >= a - 0.5 for a, b in zip(means, means[1:]))


def test_learned_sequence_fp8(seq):
    torch = require_gpu()
    w = W4K
    H, W = w.height, w.width
    wts = kpcn.Weights(kpcn.load_params())
    _, _, w8, epi = kpcn_fp8.load(quantize_kpcn.DEFAULT_FP8)
    f8 = kpcn_fp8.Fp8Net(wts, w8, epi)
    nets = {"bf16": wts.net(), "fp8": f8.net()}
    with open(os.path.splitext(kpcn.DEFAULT_WEIGHTS)[0] + ".json") as fh:
        thr = float(json.load(fh)["threshold"])
    assert math.isfinite(thr) and thr > 0
    # each network without the fallback and with the shipped threshold
    chains = {name: (net, t) for name, net in nets.items()
              for name, t in ((name, float("inf")), (name + "+fallback", thr))}
    f32 = dict(dtype=torch.float32, device="cuda")
    cubes = seq["shared"].cubes()
    rel = {name: [] for name in list(chains) + ["noisy", "atrous"]}
    prev = {name: None for name in list(chains) + ["atrous"]}
    n_tiles = ((W + 15) // 16) * ((H + 15) // 16)
    fbt = torch.zeros((1 + n_tiles,), dtype=torch.int32, device="cuda")
    fallback = {name: [] for name in chains}

    def relmse(x, target):
        x = x.double()
        return float(((x - target) ** 2 / (target * target + 1e-2)).mean())
    ws = None
    for k, f in enumerate(seq["frames"]):
        gi, fb, g = f["gi"], f["fb"], f["gb"]
        col, res = torch.zeros((H, W, 3), **f32), torch.zeros((H, W, 4), **f32)
        gi.gather_ex(g, fb.vpl_updated, fb.visible, fb.visible_entries, cubes, None, col, res,
                     mode=_abi.GATHER_SAMPLES, samples=16)
        cc, cr = torch.zeros((H, W, 3), **f32), torch.zeros((H, W, 4), **f32)
        gi.gather_ex(g, fb.vpl_updated, fb.visible, fb.visible_entries, cubes, None, cc, cr, mode=_abi.GATHER_ALL,
                     accumulate_fp64=True)
        target = (cr[..., :3] / cr[..., 3:4].clamp_min(1e-5)).double()
        del cc, cr
        if ws is None:
            ws = gi.kpcn_workspace()
        rel["noisy"].append(relmse(res[..., :3] / res[..., 3:4].clamp_min(1e-5), target))
        for name in list(chains) + ["atrous"]:
            outs = [torch.zeros((H, W, c), **f32) for c in (3, 3, 4)] + [torch.zeros((H, W), **f32)]
            pg, ph, pgb = prev[name] if prev[name] is not None else (None, None, None)
            if name == "atrous":
                gi.denoise(g, pgb, fb.visible, fb.direct, col, res, pg, ph, *outs)
            else:
                net, t = chains[name]
                gi.denoise_learned(g, pgb, fb.visible, fb.direct, col, res, pg, ph, net, t, ws, *outs, fbt)
                torch.cuda.synchronize()
                fallback[name].append(int(fbt[0].item()))
            torch.cuda.synchronize()
            assert bool(torch.isfinite(outs[1]).all()) and bool(torch.isfinite(outs[0]).all())
            rel[name].append(relmse(outs[1], target))
            prev[name] = (outs[1], outs[3], g)
        del col, res, target
    for name in rel:
        print(f"[4k {name}] relMSE per frame", " ".join(f"{r:.8f}" for r in rel[name]),
              f"fallback tiles {fallback[name]}" if name in fallback else "", flush=True)
    assert all(a <= 1.1 * b for a, b in zip(rel["fp8"], rel["bf16"]))
    assert all(a < b for a, b in zip(rel["fp8"], rel["atrous"]))
    assert all(a < b for a, b in zip(rel["bf16"], rel["atrous"]))
    # the fallback: off at +inf; at the shipped threshold it takes the a-trous filter on part of
    # this path (4K, temporal, outside the training views); the threshold is the smallest whose
    # held-out relMSE was within 0.1 % of the best, and the fallback stays within that
    for name in nets:
        assert all(n == 0 for n in fallback[name])
        assert sum(fallback[name + "+fallback"]) > 0 and all(n < n_tiles for n in fallback[name + "+fallback"])
        assert all(a <= b * (1.0 + 1e-3) for a, b in zip(rel[name + "+fallback"], rel[name])), name
