#!/usr/bin/env python3
"""GI frames/s of the deferred VPL GI pass on one MI355X: the learned form
against the a-trous form at equal relative MSE to the converged render.

A step is one GI frame on resident inputs, in RenderScene's order
(StratusRendererBackend.cpp:1982, 2010-2045), on the procedural city block at
1920x1080 with 4096 VPLs (6x256x256 FP16 distance cubes):

  learned   global VPL sun-visibility cull and compaction, the visible slots'
            atlas entries, deferred direct sun light with cascaded PCF, an
            S-sample-per-pixel gather over the visible VPLs (the reference's
            hashed estimator, vpl_pbr_gi.fs:145-200, at a fixed draw count
            with failing draws skipped) and the learned denoiser (network +
            multi-scale predicted filters, a-trous on the tiles whose
            predicted error exceeds the trained threshold, the temporal pass
            and combine), S swept over --learned-spp;
  a-trous   the same steps with S samples per pixel and the 4-pass a-trous /
            temporal denoiser, S swept over --atrous-spp;
  tile list per-tile VPL culling (16x16 tiles, up to 64 VPLs) and the
            tile-list gather + a-trous: the parity-tested plain pipeline, timed
            for its own throughput and roofline.

For each learned sample count, the a-trous form at equal relMSE is
interpolated log-log (sample count and frames/s against relMSE) between the two
swept a-trous points that bracket the learned relMSE; both bracketing points are
reported beside it.  The learned form runs with the BF16 network and with its
FP8 (E4M3) inference form when that file exists; the operating point is the
(precision, sample count) with the highest ratio of frames/s at equal relMSE,
then timed over exactly --steps frames, and `value` is its frames/s.  Every
form's frames/s comes from its frame captured in two hipGraphs (one per
history buffer set) and replayed, checked bit-equal to the eager frame first;
the per-stage `kernels_ms` come from an eager run with HIP events between the
stages.  The learned form is also run on the 1024-VPL scene (same geometry),
whose own converged render lies `converged_1024_vs_4096_relmse` from the
target: each sample count against the 4096-VPL a-trous sweep, and
(`same_scene`) against an a-trous sweep on the 1024-VPL scene, both forms
scored against the 1024-VPL converged render.  The operating point is also run
without the fallback and on the packed G-buffer attachments.  relMSE = mean of
(x - r)^2 / (r^2 + 1e-2) in FP64 over the GI output after one frame, against
the converged render (every visible VPL of the 4096-VPL scene, FP64 sums).
Inputs are
generated on the GPU before the timed regions; the denoiser history alternates
between two buffer sets as the reference swaps its FBOs.

Prints one JSON line.  With the default flags it also times the CPU oracle
(16 host threads) on a bounded sample of the a-trous frame at the learned
operating point's equal-relMSE sample count as the CPU baseline.
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "1080p GI frames/s, learned vs à-trous denoise at equal relMSE to converged"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--gather", choices=["tiles", "random"], default="tiles", help="the plain form's gather")
    ap.add_argument("--learned-spp", type=int, nargs="+", default=[8, 12, 16, 24, 32],
                    help="gather samples per pixel of the learned form (the operating point is chosen among them)")
    ap.add_argument("--atrous-spp", type=int, nargs="+",
                    default=[2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 160, 192, 224, 256, 320, 384, 448],
                    help="samples per pixel of the a-trous form's sweep")
    ap.add_argument("--sweep-steps", type=int, default=20, help="timed frames per a-trous sweep point")
    ap.add_argument("--direct-fork", choices=["start", "after-cull", "after-gather"], default="start",
                    help="where the frame forks the direct-light pass onto the second stream (it must end before the "
                         "denoiser's temporal pass, which waits for it)")
    ap.add_argument("--rejection-limit", type=int, default=-1,
                    help="sampled gathers: rejections before draws are taken unconditionally (5 = the reference's "
                         "rule); -1 = fixed draw count with failing draws skipped, which converges to the "
                         "every-VPL render")
    ap.add_argument("--views", type=int, default=32,
                    help="batched views of the city block at uneven resolutions (0 = skip)")
    ap.add_argument("--net-precision", choices=["all", "bf16", "fp8"], default="all",
                    help="the learned form's network: BF16, its FP8 inference form (when present), or both")
    ap.add_argument("--forms", choices=["all", "learned", "tiles"], default="all",
                    help="run every form (default) or one of them alone (for profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing")
    ap.add_argument("--no-learned-extras", action="store_true",
                    help="time only the learned sweep and operating point (no 1024-VPL, packed G-buffer or "
                         "no-fallback runs): counter runs then see only the 4096-VPL frame's kernels")
    ap.add_argument("--cpu-rows", type=int, default=120, help="image rows of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the CPU baseline (the oracle's rows split over them)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="instead of the benchmark, run a plain device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def copy_probe(mb: int) -> None:
    import torch
    n = mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 5}))


def gather_algorithmic_bytes(tiles_np, tile_size, W, H, K, texel_bytes):
    """Algorithmic HBM bytes of one tile-list gather launch.

    per pixel: depth 4 + normal 12 + albedo rgb 12 + reflectance 4 + roughness/metallic 8
               read, colour 12 + reservoir 16 written (68 B);
    per tile:  the list record 4 + 4K, and for each listed VPL its 64 B record and 8 B atlas entry;
    per (pixel, listed VPL): one cube texel (texel_bytes).
    """
    import numpy as np
    tx = (W + tile_size - 1) // tile_size
    counts = tiles_np[:, 0].astype(np.int64)
    ty_idx = np.arange(counts.shape[0]) // tx
    tx_idx = np.arange(counts.shape[0]) % tx
    pw = np.minimum(tile_size, W - tx_idx * tile_size)
    ph = np.minimum(tile_size, H - ty_idx * tile_size)
    pix = (pw * ph).astype(np.int64)
    total = W * H * 68 + counts.shape[0] * (4 + 4 * K) + int(counts.sum()) * (64 + 8)
    total += int((counts * pix).sum()) * texel_bytes
    return int(total), int((counts * pix).sum())


def samples_algorithmic_bytes(W, H, S, visible, texel_bytes):
    """Algorithmic HBM bytes of one sampled-gather launch (sgi_gather_ex SAMPLES, S draws per pixel):
    per pixel the 68 B of the tile-list gather (G-buffer in, colour and reservoir out), per draw one
    cube texel; the visible VPLs' 48 B packed records once."""
    return int(W * H * (68 + S * texel_bytes) + visible * 48)


def load_profile(name):
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def load_traffic(name):
    d = load_profile(name)
    return None if d is None else d.get("traffic_bytes_per_launch")


def valu_estimate(name, avg_launch_ms, draws_per_wave):
    """VALU issue of a kernel from its SQ counter summary (profiles/*_valu.json): wave instructions
    x 4 cycles / (SIMDs x clock), against the launch time measured here."""
    d = load_profile(name)
    if d is None:
        return None
    c = d["counters_per_launch"]
    issue_ms = c["SQ_INSTS_VALU"] * d["cycles_per_wave_valu_instruction"] / (d["simds"] * d["clock_ghz"] * 1e9) * 1e3
    return {"wave_instructions_per_launch": int(c["SQ_INSTS_VALU"]),
            "per_wave": round(c["SQ_INSTS_VALU"] / c["SQ_WAVES"], 1),
            "per_draw_and_wave": round(c["SQ_INSTS_VALU"] / c["SQ_WAVES"] / draws_per_wave, 1),
            "issue_estimate_ms": round(issue_ms, 4), "issue_frac_estimate": round(issue_ms / avg_launch_ms, 3),
            "source": d["source"]}


def culling_roofline(per_plain, W, H, tiles, K, visible, num_vpls):
    """Both culling stages of the plain frame against HBM: algorithmic bytes per launch
    (tile culling: depth 4 + normal 12 B per pixel, the visible records 64 B each, the tile
    records 4 + 4K B written; the global cull: 64 B per input record read, the compacted record
    and its index, 68 B per visible VPL, written), HIP-event launch times, counter traffic."""
    d = load_profile("cull_city1080p_traffic.json")
    tr = (lambda k: d["kernels"][k]["traffic_bytes_per_launch"]) if d else (lambda k: None)
    out = {}
    for name, ms, alg, traffic in (
            ("cull_global", per_plain[0], 64 * num_vpls + 68 * visible,
             None if d is None else tr("cull_visibility") + tr("cull_compact")),
            ("cull_tiles", per_plain[1], 16 * W * H + 64 * visible + tiles * (4 + 4 * K), tr("cull_tiles"))):
        ach = alg / (ms * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": int(alg),
                     "avg_launch_ms": round(ms, 4)}
    out["cull_global"]["kernels"] = "cullVisibilityKernel + cullCompactKernel"
    out["cull_tiles"]["kernels"] = "cullTilesKernel<16>"
    out["note"] = ("neither stage is HBM-bound: the global cull is two launches of 64 and 16 waves over 4096 VPLs "
                   "(latency), the tile cull one wave per tile against every visible VPL (VALU for the estimated "
                   "ratios and the exact ones below the histogram bound, and the serial stage-1 sums; "
                   "profiles/valu_counters_gather_samples_city1080p.md)")
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(host, w, rows, visible_count, spp, rejection_limit, threads):
    """Times the oracle (the FP32 restatement, rows split over `threads` host threads) on a bounded
    sample of the frame: the a-trous form the learned form is compared with (global culling, the
    sampled gather at `spp` draws per pixel over every visible VPL, direct light, the a-trous
    denoiser), and beside it the reference's tile-list form (tile culling + tile-list gather)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))

    import numpy as np

    import oracle as orc

    used = orc.set_threads(threads)
    H, W = w.height, w.width
    y0 = max(0, H // 2 - rows // 2)
    y1 = min(H, y0 + rows)
    scale = H / (y1 - y0)
    g, cubes, csm = host.gbuffer(), host.cubes(), host.csm_struct()
    t = {}
    t0 = time.perf_counter()
    upd, vis = orc.cull_global(host.frame, csm, host.vpls)
    t["cull_global"] = time.perf_counter() - t0
    ent = host.entries[vis[1:]]
    t0 = time.perf_counter()
    lit = orc.direct_light(host.frame, csm, g, y0, y1)
    t["direct_light"] = (time.perf_counter() - t0) * scale
    t0 = time.perf_counter()
    col, res = orc.gather(host.frame, g, upd, int(vis[0]), ent, cubes, mode=orc.GATHER_SAMPLES, samples=spp,
                          rejection_limit=rejection_limit, y0=y0, y1=y1)
    t["gather"] = (time.perf_counter() - t0) * scale
    # the denoiser on the sampled rows as an image of their own (same per-pixel work)
    crop = lambda a: np.ascontiguousarray(a[y0:y1])  # noqa: E731
    planes = [crop(host.depth), crop(host.normal), crop(host.albedo), crop(host.refl), crop(host.rma),
              crop(host.velocity), crop(host.ids)]
    from stratusgfx_amd import _abi
    cg = _abi.Gbuffer(W, y1 - y0, *[p.ctypes.data for p in planes], 0)
    t0 = time.perf_counter()
    orc.denoise(host.frame, cg, None, int(vis[0]), lit, col, res)
    t["denoise"] = (time.perf_counter() - t0) * scale
    # the tile-list form's own stages
    tl = {}
    t0 = time.perf_counter()
    _, tiles = orc.cull_tiles(host.frame, g, w.tile_size, w.max_vpls_per_tile, upd, vis)
    tl["cull_tiles"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.gather(host.frame, g, upd, int(vis[0]), ent, cubes, tiles=tiles, tile_size=w.tile_size,
               K=w.max_vpls_per_tile, mode=orc.GATHER_TILES, y0=y0, y1=y1)
    tl["gather"] = (time.perf_counter() - t0) * scale
    orc.set_threads(1)
    total = sum(t.values())
    tl_total = t["cull_global"] + t["direct_light"] + t["denoise"] + sum(tl.values())
    return {
        "value": round(1.0 / total, 4), "unit": "frames/s", "cores": used, "kind": "port",
        "cpu": cpu_model(), "host_threads_available": os.cpu_count(),
        "sample": (f"FP32 oracle on {used} threads (rows split over threads): the a-trous form at {spp} draws "
                   f"per pixel (the learned form's equal-relMSE point) - global culling ({visible_count} "
                   f"visible of {w.num_vpls}) on the whole frame, direct light, sampled gather and denoiser on "
                   f"rows {y0}..{y1} of {H}, scaled by {scale:.2f} to a frame"),
        "seconds_per_frame": {k: round(v, 3) for k, v in t.items()},
        "tile_list_form": {"frames_per_s": round(1.0 / tl_total, 4),
                           "seconds_per_frame": {k: round(v, 3) for k, v in tl.items()},
                           "note": "cull_global, direct light and denoiser as above; tile culling on the whole "
                                   "frame, the tile-list gather on the same rows"},
    }


def rel_mse(x, ref) -> float:
    """mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2), in FP64."""
    x, r = x.double(), ref.double()
    return float(((x - r) ** 2 / (r * r + 1e-2)).mean())


def atrous_at_equal(sweep, learned):
    """The a-trous form at exactly the learned form's relMSE: sample count and frames/s interpolated
    log-log against relMSE between the two swept points that bracket it. The bracketing points are
    reported beside it: the cheapest one at or below the learned relMSE (an upper bound of the
    a-trous cost, so its speedup overstates) and the dearest one above it (a lower bound, so its
    speedup understates). Outside the swept range the nearest point gives a bound only."""
    if not sweep:
        return {}
    rel, fps_l = learned["relmse"], learned["frames_per_s"]
    pts = sorted(sweep, key=lambda r: r["spp"])
    below = [r for r in pts if r["relmse"] <= rel]
    above = [r for r in pts if r["relmse"] > rel]
    out = {}
    if below:
        b = below[0]
        out["atrous_at_or_below_relmse"] = {**b, "speedup": round(fps_l / b["frames_per_s"], 3)}
    if above:
        a = above[-1]
        out["atrous_above_relmse"] = {**a, "speedup": round(fps_l / a["frames_per_s"], 3)}
    for a, b in zip(pts, pts[1:]):
        if a["relmse"] >= rel >= b["relmse"] and a["relmse"] > b["relmse"]:
            t = (math.log(a["relmse"]) - math.log(rel)) / (math.log(a["relmse"]) - math.log(b["relmse"]))
            fps = math.exp(math.log(a["frames_per_s"]) + t * (math.log(b["frames_per_s"]) - math.log(a["frames_per_s"])))
            spp = math.exp(math.log(a["spp"]) + t * (math.log(b["spp"]) - math.log(a["spp"])))
            out["atrous_at_equal_relmse"] = {"spp": round(spp, 2), "frames_per_s": round(fps, 2), "relmse": rel,
                                             "interpolated": True}
            out["speedup_vs_atrous_at_equal_relmse"] = round(fps_l / fps, 3)
            return out
    # not bracketed: below every swept relMSE (the a-trous form would need more samples than swept,
    # the ratio to the last point is a lower bound) or above every one (an upper bound)
    edge = pts[-1] if rel < pts[-1]["relmse"] else pts[0]
    out["atrous_at_equal_relmse"] = {**edge, "interpolated": False,
                                     "bound": "lower" if edge is pts[-1] else "upper"}
    out["speedup_vs_atrous_at_equal_relmse"] = round(fps_l / edge["frames_per_s"], 3)
    return out


def timed_loop(step, steps, warmup, nstages):
    """Runs warmup + steps frames; returns (elapsed seconds of the timed steps, mean ms per stage of
    the frame's stream, mean ms of the direct-light pass on its concurrent stream). step(k, ev, evd)
    records ev[0..nstages] on the frame's stream and evd[0..1] around the concurrent pass."""
    import torch
    for k in range(warmup):
        step(k, None, None)
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(nstages + 1)] for _ in range(steps)]
    side = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(warmup + k, events[k], side[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per = [sum(events[k][i].elapsed_time(events[k][i + 1]) for k in range(steps)) / steps for i in range(nstages)]
    direct = sum(side[k][0].elapsed_time(side[k][1]) for k in range(steps)) / steps
    return elapsed, per, direct


def graph_loop(step, steps, warmup, state, outputs):
    """The frame captured in two hipGraphs (torch.cuda.graph; frame parity 0 and 1 alternate the two
    history buffer sets, as the eager loop does) and replayed: warmup replays, then exactly `steps`
    replays between two synchronisations; returns the elapsed seconds.  One launch per frame, so the
    host's per-call argument checks stay out of the frame time.  Before timing, frame 0 runs eagerly
    and then from its graph on the same state (`state`: the tensors frame 0 reads and no frame
    rewrites before it, restored in between) and `outputs` must come out bit-equal."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graphs = []
    for par in range(2):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            step(par, None, None)
        graphs.append(g)
    torch.cuda.synchronize()
    saved = [t.clone() for t in state]
    step(0, None, None)
    torch.cuda.synchronize()
    ref = [t.clone() for t in outputs]
    for t, v in zip(state, saved):
        t.copy_(v)
    torch.cuda.synchronize()
    graphs[0].replay()
    torch.cuda.synchronize()
    if not all(torch.equal(a, b) for a, b in zip(outputs, ref)):
        raise SystemExit("bench.py: a graph replay differs from the eager frame")
    for k in range(warmup):
        graphs[k % 2].replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        graphs[(warmup + k) % 2].replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    del graphs
    return elapsed


def stage_ms(names, per, direct):
    """Per-stage mean ms of the frame's stream, plus the concurrent direct-light pass."""
    out = {k: round(v, 4) for k, v in zip(names, per)}
    out["direct_light_concurrent"] = round(direct, 4)
    return out


_SIDE_STREAM = []


def fork_direct(p, fb, evd):
    """Issues the direct-light pass on a second stream of default priority: it reads only the
    G-buffer and the shadow cascades, so it runs beside global culling, the VPL gather and the
    denoiser's spatial passes. The pass context's screen event (sgi_set_screen_event) is recorded
    after it: the denoiser waits for it just before its temporal pass, the only one that reads the
    direct light. The fork is recorded after the previous frame's denoiser, which read the same
    buffer."""
    import torch
    if not _SIDE_STREAM:
        _SIDE_STREAM.append(torch.cuda.Stream())
    side = _SIDE_STREAM[0]
    fork = torch.cuda.Event()
    fork.record(torch.cuda.current_stream())
    side.wait_event(fork)
    if evd is not None:
        evd[0].record(side)
    p.gi.direct_light(p.g, fb.direct, stream=side.cuda_stream)
    if evd is not None:
        evd[1].record(side)
    if getattr(p, "screen_event", None) is None:
        p.screen_event = torch.cuda.Event()
    p.screen_event.record(side)
    p.gi.set_screen_event(p.screen_event)


class FormState:
    """One form's scene, pass context and frame buffers."""

    def __init__(self, w):
        import torch

        from stratusgfx_amd import scene as sc
        from stratusgfx_amd.gi import FrameBuffers, GiPass

        print(f"[bench] generating {w.name} on the GPU", file=sys.stderr, flush=True)
        self.w = w
        self.scene = sc.Scene(w)
        self.dev = sc.generate_device(self.scene)
        self.gi = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
        self.gi.set_frame(self.dev.frame)
        self.gi.set_csm(self.dev.csm_struct())
        self.g, self.cubes = self.dev.gbuffer(), self.dev.cubes()
        self.n = w.num_vpls
        self.fb = FrameBuffers.allocate(w.width, w.height, self.n, self.gi.num_tiles, w.max_vpls_per_tile)
        self.hist = [(torch.zeros_like(self.fb.gi), torch.zeros_like(self.fb.history)) for _ in range(2)]

    def cull(self):
        from stratusgfx_amd import _abi
        fb = self.fb
        self.gi.cull_global(self.dev.vpls, self.n, fb.vpl_updated, fb.visible)
        self.gi.visible_atlas_entries(self.dev.entries, fb.visible, min(self.n, _abi.MAX_TOTAL_VPLS_PER_FRAME),
                                      fb.visible_entries)


def conv_roofline(net_w, H, W, reps=20):
    """HIP-event timing of the learned denoiser's convolution building block (sgi_kpcn_conv3x3, 32 -> 32
    channels, BF16) at the full frame size: the kernel the network's full-resolution layers run with
    their fused staging and epilogues."""
    import torch

    from stratusgfx_amd import _abi, kpcn
    lib = _abi.load()
    x = torch.randn((H, W, 32), device="cuda").clamp_min(0).to(torch.bfloat16)
    y = torch.empty((H, W, 32), dtype=torch.bfloat16, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    args = (x.data_ptr(), 0, net_w.layer_w(1).data_ptr(), net_w.layer_b(1).data_ptr(), 1, H, W, y.data_ptr(), 0, s)
    for _ in range(3):
        _abi.call(lib, "sgi_kpcn_conv3x3", *args)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record()
    for i in range(reps):
        _abi.call(lib, "sgi_kpcn_conv3x3", *args)
        ev[i + 1].record()
    torch.cuda.synchronize()
    ms = sum(ev[i].elapsed_time(ev[i + 1]) for i in range(reps)) / reps
    P = H * W
    alg_bytes = P * 32 * 2 * 2 + kpcn.LAYER_WEIGHTS * 2 + 32 * 4
    flops = 2 * 9 * 32 * 32 * P
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    return {"kernel": "conv3x3Kernel<plain in, bf16 out> (32->32, 1920x1080)", "bound": "hbm",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("conv_city1080p_traffic.json"),
            "algorithmic_bytes": alg_bytes, "flops": flops, "avg_launch_ms": round(ms, 4),
            "mfma_tflops": round(flops / (ms * 1e-3) / 1e12, 1)}


def vpl_colouring(state, reps=50):
    """HIP-event timing of sgi_color_vpls (vpl_light_color.cs: colour and specular position of every
    visible VPL from its diffuse and distance cubes) on the frame's visible list, writing a scratch copy
    of the records (the timed frames keep the generator's colours).  A queued matmul keeps the launches
    back to back behind it, and one event pair brackets all of them (an event between every two
    launches would add its own packet's time to each)."""
    import torch

    from stratusgfx_amd import _abi
    from stratusgfx_amd import scene as sc
    dif, cd = sc.diffuse_cubes(state.scene, state.dev)
    state.cull()
    fb = state.fb
    recs = fb.vpl_updated.clone()
    cap = min(state.n, _abi.MAX_TOTAL_VPLS_PER_FRAME)

    def run():
        state.gi.color_vpls(fb.visible, fb.visible_entries, cd, state.cubes, cap, recs)

    for _ in range(3):
        run()
    a = torch.randn((4096, 4096), device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    for _ in range(3):
        a = a @ a
        a = a / a.abs().amax()
    ev[0].record()
    for _ in range(reps):
        run()
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    visible = int(fb.visible[0].item())
    atlas_gb = sum(t.numel() for t in dif) / 1e9
    del dif, a
    torch.cuda.empty_cache()
    return {"kernel": "colorVplsKernel<F16>", "avg_launch_ms": round(ms, 4), "visible_vpls": visible,
            "algorithmic_bytes": visible * (64 + 8 + 4 + 2 + 32),
            "diffuse_cubes": f"6x{state.w.cube_res}x{state.w.cube_res} rgba8, {atlas_gb:.1f} GB",
            "note": "one thread per visible VPL (two texels, one record); outside the timed GI frame, whose "
                    "records arrive coloured"}


def gbuffer_unpack_roofline(state, reps=20):
    """HIP-event timing of sgi_gbuffer_unpack at the frame size: the reference's RGB8 / RGBA8 / RG8 / RGB8
    attachments and RG16F velocity (18 B per pixel) into the FP32 planes (56 B per pixel)."""
    import torch

    from stratusgfx_amd import _abi
    W, H = state.w.width, state.w.height
    i32 = dict(dtype=torch.int32, device="cuda")
    src = [torch.randint(-2**31, 2**31 - 1, (H, W), **i32) for _ in range(4)]
    refl = torch.randint(-2**15, 2**15 - 1, (H, W), dtype=torch.int16, device="cuda")
    packed = _abi.GbufferPacked(W, H, src[0].data_ptr(), src[1].data_ptr(), refl.data_ptr(), src[2].data_ptr(),
                                src[3].data_ptr())
    out = [torch.empty((H, W, c), dtype=torch.float32, device="cuda") for c in (3, 4, 2, 3, 2)]
    for _ in range(3):
        state.gi.gbuffer_unpack(packed, *out)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        state.gi.gbuffer_unpack(packed, *out)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    alg = W * H * (18 + 56)
    achieved = alg / (ms * 1e-3) / 1e9
    del src, refl, out
    return {"kernel": "unpackKernel (packed G-buffer attachments -> FP32 planes)", "bound": "hbm",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": alg, "avg_launch_ms": round(ms, 4)}


def post_fx_roofline(state, reps=20):
    """HIP-event timing of the lit buffer's consumers at the frame size on the plain form's last frame:
    sgi_taa (current frame, history, velocity -> FP32; 44 B per pixel), sgi_tonemap (FP32 -> RGBA8;
    16 B per pixel) and sgi_fxaa on the tonemapped target (RGBA8 in and out; 8 B per pixel)."""
    import torch
    W, H = state.w.width, state.w.height
    lit = state.fb.combined
    hist = lit.clone()
    vel = torch.zeros((H, W, 2), dtype=torch.float32, device="cuda")
    taa_out = torch.empty_like(lit)
    rgba = torch.empty((H, W), dtype=torch.int32, device="cuda")
    out = {}
    aa = torch.empty_like(rgba)
    for name, run, per_px in (("taa", lambda: state.gi.taa(lit, hist, vel, taa_out), 44),
                              ("tonemap", lambda: state.gi.tonemap(lit, rgba), 16),
                              ("fxaa", lambda: state.gi.fxaa(rgba, aa), 8)):
        for _ in range(3):
            run()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            run()
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / reps
        achieved = W * H * per_px / (ms * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": W * H * per_px,
                     "avg_launch_ms": round(ms, 4)}
    del hist, vel, taa_out, rgba, aa
    return out


def _event_ms(run, reps):
    """Mean ms of `reps` back-to-back calls of run() between one pair of HIP events (3 calls first)."""
    import torch
    for _ in range(3):
        run()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        run()
    ev[1].record()
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def shadow_passes(state, reps=10):
    """HIP-event timing of the two shadow pre-passes on the block's meshes (12 triangles per box):
    sgi_render_vpl_cubes for the reference's per-frame budget of 5 lights (maxShadowUpdatesPerFrame,
    StratusRendererBackend.h:378) and for 256 lights, into scratch atlases (distance in the frame's
    texel format and the RGBA8 colour layer, both written as the reference's pass does), and
    sgi_render_csm (the frame's 4 cascades)."""
    import numpy as np
    import torch

    from stratusgfx_amd import _abi
    w = state.w
    pos, col = state.scene.triangles()
    pos_d = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    col_d = torch.from_numpy(np.ascontiguousarray(col.view(np.int32))).cuda()
    res = w.cube_res
    out = {"triangles": int(pos.shape[0])}
    vis = state.fb.visible[1:1 + int(state.fb.visible[0].item())]
    for n in (5, 256):
        layers = n
        dt = torch.float32 if w.cube_format == _abi.TEXEL_F32 else torch.float16
        z = torch.empty((layers, 6, res, res), dtype=dt, device="cuda")
        c = torch.empty((layers, 6, res, res), dtype=torch.int32, device="cuda")
        cz, cc = _abi.CubeAtlases(), _abi.CubeAtlases()
        cz.resolution, cz.format, cz.layers_per_atlas, cz.num_atlases = res, w.cube_format, layers, 1
        cc.resolution, cc.format, cc.layers_per_atlas, cc.num_atlases = res, _abi.TEXEL_RGBA8, layers, 1
        cz.atlas[0], cc.atlas[0] = z.data_ptr(), c.data_ptr()
        ups = vis[:n].contiguous()
        ent = torch.zeros((state.n, 2), dtype=torch.int32, device="cuda")
        ent[ups.long(), 1] = torch.arange(ups.numel(), dtype=torch.int32, device="cuda")
        ms = _event_ms(lambda: state.gi.render_vpl_cubes(pos_d, col_d, state.dev.vpls, ups, ent, cz, cc),
                       reps if n <= 5 else 2)
        texels = ups.numel() * 6 * res * res
        out[f"vpl_cubes_{n}"] = {"lights": int(ups.numel()), "avg_ms": round(ms, 4),
                                 "texels_per_s": round(texels / (ms * 1e-3), 1),
                                 "written_bytes": texels * (z.element_size() + 4)}
        del z, c
    csm = state.dev.csm_struct()
    r = csm.resolution
    depth = torch.empty((4, r, r), dtype=torch.float32, device="cuda")
    ms = _event_ms(lambda: state.gi.render_csm(pos_d, csm, depth), reps)
    out["csm"] = {"cascades": 4, "resolution": r, "avg_ms": round(ms, 4), "written_bytes": 4 * r * r * 4}
    out["kernels"] = "cubeFaceListKernel + cubeRasterKernel; csmListKernel + csmRasterKernel"
    del depth
    return out


def batched_views(args, learned=None):
    """Many views of the city block at uneven resolutions, one launch per stage over all views
    in a captured graph, launch shape from the autotuner."""
    import numpy as np
    import torch

    from stratusgfx_amd import autotune
    from stratusgfx_amd import scene as sc
    from stratusgfx_amd.train_kpcn import random_view

    rng = np.random.default_rng(2)
    sizes = [(1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]
    scene = sc.Scene(sc.CITY_1080P)
    shared = sc.generate_device(scene)
    inputs = []
    for i in range(args.views):
        W, H = sizes[int(rng.integers(len(sizes)))]
        v = random_view(scene, rng, W, H)
        v.csm_resolution = sc.CITY_1080P.csm_res
        scene.set_view(v)
        inputs.append(sc.generate_view_device(scene, shared))
    mp = sum(i.width * i.height for i in inputs) / 1e6
    print(f"[bench] batched views: {len(inputs)} views, {mp:.1f} MP; autotuning", file=sys.stderr, flush=True)
    tuned = autotune.tune(inputs, shared, sc.CITY_1080P.num_vpls, reps=5)
    c = tuned["choice"]
    ms, _ = autotune.measure(inputs, shared, sc.CITY_1080P.num_vpls, c["tile_size"], c["max_vpls_per_tile"],
                             reps=args.steps)
    out = {"views": len(inputs), "megapixels": round(mp, 2), "tile_size": c["tile_size"],
           "max_vpls_per_tile": c["max_vpls_per_tile"], "ms_per_batch": round(ms, 3),
           "views_per_s": round(len(inputs) / (ms * 1e-3), 1), "megapixels_per_s": round(mp / (ms * 1e-3), 1),
           "relmse": c["relmse"], "reference_shape": tuned["reference"], "autotune": tuned["table"],
           "note": "tile-list form: one launch per stage over every view, replayed from a captured graph"}
    if learned is not None:
        out["learned"] = batched_learned(inputs, shared, sc.CITY_1080P.num_vpls, learned, args.steps)
    return out


def batched_learned(inputs, shared, num_vpls, learned, reps):
    """The same batch in its learned form (sgi_views_set_learned): batched culling and direct light,
    then each view's sampled gather and network with the fallback threshold, in one captured graph;
    HIP-event time per batch and the mean relMSE of the views' GI against their converged renders."""
    import torch

    from stratusgfx_amd import autotune
    from stratusgfx_amd.gi import Views, view_io
    S, net, thr, label = learned
    f32 = dict(dtype=torch.float32, device="cuda")
    bufs, ios = [], []
    for inp in inputs:
        H, W = inp.height, inp.width
        nt = ((W + 15) // 16) * ((H + 15) // 16)
        b = dict(tiles=torch.zeros((nt, 65), dtype=torch.int32, device="cuda"), direct=torch.zeros((H, W, 3), **f32),
                 color=torch.zeros((H, W, 3), **f32), reservoir=torch.zeros((H, W, 4), **f32),
                 combined=torch.zeros((H, W, 3), **f32), gi=torch.zeros((H, W, 3), **f32),
                 reservoir_out=torch.zeros((H, W, 4), **f32), history=torch.zeros((H, W), **f32))
        bufs.append(b)
        ios.append(view_io(inp, **b))
    vw = Views(len(inputs), sum(i.width * i.height for i in inputs), 16, 64)
    s = torch.cuda.Stream()
    vw.set(ios, shared.vpls, num_vpls, shared.entries, shared.cubes(), stream=s.cuda_stream)
    vw.set_learned(S, net, thr)
    vw.capture(s.cuda_stream)
    for _ in range(2):
        vw.replay(s.cuda_stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    with torch.cuda.stream(s):
        ev[0].record()
        for i in range(reps):
            vw.replay(s.cuda_stream)
            ev[i + 1].record()
    s.synchronize()
    ms = sum(ev[i].elapsed_time(ev[i + 1]) for i in range(reps)) / reps
    rel = sum(autotune.rel_mse(b["gi"], autotune.converged(inp, shared, num_vpls))
              for b, inp in zip(bufs, inputs)) / len(inputs)
    vw.close()
    mp = sum(i.width * i.height for i in inputs) / 1e6
    return {"samples_per_pixel": S, "network": label, "fallback_threshold": thr, "ms_per_batch": round(ms, 3),
            "views_per_s": round(len(inputs) / (ms * 1e-3), 1), "megapixels_per_s": round(mp / (ms * 1e-3), 1),
            "relmse": round(rel, 6),
            "note": "culling and direct light one launch over every view; per view the sampled gather and the "
                    "network (their sizes differ), all in one captured graph"}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import dataclasses

    import torch

    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)
        return

    from stratusgfx_amd import _abi, kpcn, kpcn_fp8
    from stratusgfx_amd import scene as sc

    w_learned = sc.CITY_1080P
    w_plain = dataclasses.replace(sc.CITY_1080P, name="city-1920x1080-4096vpl", num_vpls=4096)
    plain = FormState(w_plain)
    learned = FormState(w_learned)
    texel_bytes = 4 if w_plain.cube_format == _abi.TEXEL_F32 else 2
    H, W = w_plain.height, w_plain.width
    f32 = dict(dtype=torch.float32, device="cuda")

    # the converged render: every visible VPL of the 4096-VPL scene, FP64 sums
    plain.cull()
    conv_col, conv_res = torch.zeros((H, W, 3), **f32), torch.zeros((H, W, 4), **f32)
    plain.gi.gather_ex(plain.g, plain.fb.vpl_updated, plain.fb.visible, plain.fb.visible_entries, plain.cubes, None,
                       conv_col, conv_res, mode=_abi.GATHER_ALL, accumulate_fp64=True)
    target = (conv_res[..., :3] / conv_res[..., 3:4].clamp_min(1e-5)).contiguous()
    # how far the 1024-VPL scene's own converged render is from it (the learned form's floor)
    learned.cull()
    learned.gi.gather_ex(learned.g, learned.fb.vpl_updated, learned.fb.visible, learned.fb.visible_entries,
                         learned.cubes, None, conv_col, conv_res, mode=_abi.GATHER_ALL, accumulate_fp64=True)
    target_1024 = (conv_res[..., :3] / conv_res[..., 3:4].clamp_min(1e-5)).contiguous()
    floor_rel = rel_mse(target_1024, target)
    del conv_col, conv_res

    # ---- plain forms ---------------------------------------------------------
    # (1) the pipeline of the reference with tile lists (the parity-tested path): timed, roofline
    use_tiles = args.gather == "tiles"

    def fork_at(where, p, fb, evd):
        """Forks the direct-light pass at the frame stage --direct-fork names."""
        if where == args.direct_fork:
            fork_direct(p, fb, evd)

    def plain_step(k, ev, evd, hist=True):
        p, fb = plain, plain.fb
        prev_gi, prev_h = p.hist[k % 2] if hist else (None, None)
        out_gi, out_h = p.hist[(k + 1) % 2]
        rec = (lambda i: ev[i].record()) if ev is not None else (lambda i: None)
        rec(0)
        fork_at("start", p, fb, evd)
        p.cull()
        fork_at("after-cull", p, fb, evd)
        rec(1)
        if use_tiles:
            p.gi.cull_tiles(p.g, fb.vpl_updated, fb.visible, fb.tiles)
        rec(2)
        p.gi.gather(p.g, fb.vpl_updated, fb.visible, fb.visible_entries, p.cubes, fb.tiles if use_tiles else None,
                    fb.color, fb.reservoir)
        fork_at("after-gather", p, fb, evd)
        rec(3)
        p.gi.denoise(p.g, None, fb.visible, fb.direct, fb.color, fb.reservoir, prev_gi, prev_h, fb.combined, out_gi,
                     fb.reservoir_out, out_h)
        rec(4)
        return out_gi

    gi_plain = plain_step(0, None, None, hist=False)
    torch.cuda.synchronize()
    plain_rel = rel_mse(gi_plain, target)
    el_plain, per_plain = None, None
    if args.forms in ("all", "tiles"):
        print(f"[bench] timing {args.steps} plain frames", file=sys.stderr, flush=True)
        _, per_plain, direct_plain = timed_loop(plain_step, args.steps, args.warmup, 4)
        el_plain = graph_loop(plain_step, args.steps, args.warmup, list(plain.hist[0]),
                              [plain.hist[1][0], plain.fb.combined])
    tiles_np = plain.fb.tiles.cpu().numpy()
    visible_plain = int(plain.fb.visible[0].item())

    # (2) the a-trous form at equal relMSE: the reference's hashed estimator over every
    # visible VPL at S samples per pixel + the a-trous denoiser
    def atrous_step(S, p=plain):
        def step(k, ev, evd, hist=True):
            fb = p.fb
            prev_gi, prev_h = p.hist[k % 2] if hist else (None, None)
            out_gi, out_h = p.hist[(k + 1) % 2]
            rec = (lambda i: ev[i].record()) if ev is not None else (lambda i: None)
            rec(0)
            fork_at("start", p, fb, evd)
            p.cull()
            fork_at("after-cull", p, fb, evd)
            rec(1)
            p.gi.gather_ex(p.g, fb.vpl_updated, fb.visible, fb.visible_entries, p.cubes, None, fb.color,
                           fb.reservoir, mode=_abi.GATHER_SAMPLES, samples=S, rejection_limit=args.rejection_limit)
            fork_at("after-gather", p, fb, evd)
            rec(2)
            p.gi.denoise(p.g, None, fb.visible, fb.direct, fb.color, fb.reservoir, prev_gi, prev_h, fb.combined,
                         out_gi, fb.reservoir_out, out_h)
            rec(3)
            return out_gi
        return step

    # (3) the reference's own live estimator (vpl_pbr_gi.fs:139-200: 1-6 hashed draws per pixel by
    # roughness and distance, at most 5 rejections) over every visible VPL + the a-trous denoiser
    def random_step(k, ev, evd, hist=True):
        p, fb = plain, plain.fb
        prev_gi, prev_h = p.hist[k % 2] if hist else (None, None)
        out_gi, out_h = p.hist[(k + 1) % 2]
        rec = (lambda i: ev[i].record()) if ev is not None else (lambda i: None)
        rec(0)
        fork_at("start", p, fb, evd)
        p.cull()
        fork_at("after-cull", p, fb, evd)
        rec(1)
        p.gi.gather(p.g, fb.vpl_updated, fb.visible, fb.visible_entries, p.cubes, None, fb.color, fb.reservoir)
        fork_at("after-gather", p, fb, evd)
        rec(2)
        p.gi.denoise(p.g, None, fb.visible, fb.direct, fb.color, fb.reservoir, prev_gi, prev_h, fb.combined,
                     out_gi, fb.reservoir_out, out_h)
        rec(3)
        return out_gi

    random_res = None
    if args.forms == "all":
        g_r = random_step(0, None, None, hist=False)
        torch.cuda.synchronize()
        rm = rel_mse(g_r, target)
        draws = float(plain.fb.reservoir[..., 3].double().mean())
        _, per_r, direct_r = timed_loop(random_step, args.sweep_steps, 2, 3)
        el_r = graph_loop(random_step, args.sweep_steps, 2, list(plain.hist[0]), [plain.hist[1][0], plain.fb.combined])
        random_res = {"relmse": round(rm, 6), "frames_per_s": round(args.sweep_steps / el_r, 2),
                      "mean_reservoir_weight": round(draws, 3),
                      "kernels_ms": stage_ms(["cull_global", "gather", "denoise"], per_r, direct_r),
                      "note": "the reference's per-pixel estimator as shipped (gatherRandomKernel), 4096 VPLs"}
        print(f"[bench] reference estimator: relMSE {rm:.5f}, {args.sweep_steps / el_r:.1f} frames/s",
              file=sys.stderr, flush=True)

    def atrous_sweep(p, tgt, label):
        out = []
        for S in (args.atrous_spp if args.forms == "all" else []):
            st = atrous_step(S, p)
            g_s = st(0, None, None, hist=False)
            torch.cuda.synchronize()
            rm = rel_mse(g_s, tgt)
            _, per_s, direct_s = timed_loop(st, args.sweep_steps, 2, 3)
            el_s = graph_loop(st, args.sweep_steps, 2, list(p.hist[0]), [p.hist[1][0], p.fb.combined])
            out.append({"spp": S, "relmse": round(rm, 6), "frames_per_s": round(args.sweep_steps / el_s, 2),
                        "kernels_ms": stage_ms(["cull_global", "gather", "denoise"], per_s, direct_s)})
            print(f"[bench] a-trous ({label}) at {S} spp: relMSE {rm:.5f}, {args.sweep_steps / el_s:.1f} frames/s",
                  file=sys.stderr, flush=True)
        return out

    sweep = atrous_sweep(plain, target, "4096 VPLs")
    # the same sweep on the 1024-VPL scene against its own converged render
    sweep_1024 = atrous_sweep(learned, target_1024, "1024 VPLs")

    # ---- learned form ----------------------------------------------------------
    # on the a-trous form's 4096-VPL scene (same VPLs, same sampler: only the sample count and the
    # denoiser differ), and on the 1024-VPL scene (whose own converged render sits
    # `converged_1024_vs_4096_relmse` from the target)
    weights_path = kpcn.DEFAULT_WEIGHTS
    learned_res = learned_1024 = learned_packed = learned_no_fallback = None
    learned_sweep = []
    learned_1024_sweep = []
    el_learned = None
    wts = None
    if os.path.exists(weights_path) and args.forms in ("all", "learned"):
        meta_path = os.path.splitext(weights_path)[0] + ".json"
        threshold = float("inf")
        if os.path.exists(meta_path):
            with open(meta_path) as f:
                threshold = float(json.load(f).get("threshold", float("inf")))
        wts = kpcn.Weights(kpcn.load_params(weights_path))
        nets = {"bf16": wts.net()}
        fp8_path = os.path.splitext(weights_path)[0] + "_fp8.npz"
        if os.path.exists(fp8_path) and args.net_precision in ("all", "fp8"):
            _, _, w8, epi = kpcn_fp8.load(fp8_path)
            f8net = kpcn_fp8.Fp8Net(wts, w8, epi)
            nets["fp8"] = f8net.net()
        if args.net_precision == "fp8" and "fp8" in nets:
            del nets["bf16"]
        n_tiles = ((W + 15) // 16) * ((H + 15) // 16)
        fbt = torch.zeros((1 + n_tiles,), dtype=torch.int32, device="cuda")

        def learned_form(state, label, S, steps, warmup, precision, own_target=None, form_threshold=None):
            thr_form = threshold if form_threshold is None else form_threshold
            net = nets[precision]
            ws = state.gi.kpcn_workspace()

            def learned_step(k, ev, evd, hist=True, thr=thr_form):
                p, fb = state, state.fb
                prev_gi, prev_h = p.hist[k % 2] if hist else (None, None)
                out_gi, out_h = p.hist[(k + 1) % 2]
                rec = (lambda i: ev[i].record()) if ev is not None else (lambda i: None)
                rec(0)
                fork_at("start", p, fb, evd)
                p.cull()
                fork_at("after-cull", p, fb, evd)
                rec(1)
                p.gi.gather_ex(p.g, fb.vpl_updated, fb.visible, fb.visible_entries, p.cubes, None, fb.color,
                               fb.reservoir, mode=_abi.GATHER_SAMPLES, samples=S,
                               rejection_limit=args.rejection_limit)
                fork_at("after-gather", p, fb, evd)
                rec(2)
                p.gi.denoise_learned(p.g, None, fb.visible, fb.direct, fb.color, fb.reservoir, prev_gi, prev_h, net,
                                     thr, ws, fb.combined, out_gi, fb.reservoir_out, out_h, fbt)
                rec(3)
                return out_gi

            gi_l = learned_step(0, None, None, hist=False)
            torch.cuda.synchronize()
            rel = rel_mse(gi_l, target)
            rel_own = rel_mse(gi_l, own_target) if own_target is not None else None
            fallback_tiles = int(fbt[0].item())
            fb = state.fb
            noisy = rel_mse(fb.reservoir[..., :3] / fb.reservoir[..., 3:4].clamp_min(1e-5), target)
            gi_nf = learned_step(0, None, None, hist=False, thr=float("inf"))
            torch.cuda.synchronize()
            rel_nf = rel_mse(gi_nf, target)
            # the a-trous filter alone on the same input
            tmp = [torch.zeros((H, W, c), **f32) for c in (3, 3, 4)] + [torch.zeros((H, W), **f32)]
            state.gi.denoise(state.g, None, fb.visible, fb.direct, fb.color, fb.reservoir, None, None, *tmp)
            torch.cuda.synchronize()
            atrous_same = rel_mse(tmp[1], target)
            del tmp
            _, per, direct_ms = timed_loop(learned_step, steps, warmup, 3)
            el = graph_loop(learned_step, steps, warmup, list(state.hist[0]), [state.hist[1][0], state.fb.combined])
            res = {
                "frames_per_s": round(steps / el, 2), "ms_per_step": round(el * 1e3 / steps, 4),
                "relmse": round(rel, 6), "relmse_without_fallback": round(rel_nf, 6),
                "relmse_noisy_input": round(noisy, 6), "relmse_atrous_on_same_input": round(atrous_same, 6),
                "fallback_threshold": thr_form, "fallback_tiles": fallback_tiles, "tiles": n_tiles,
                "kernels_ms": stage_ms(["cull_global", "gather", "denoise_learned"], per, direct_ms),
                "vpls": state.w.num_vpls, "visible_vpls": int(fb.visible[0].item()),
                "samples_per_pixel": S, "network_precision": precision,
                "weights": os.path.relpath(fp8_path if precision == "fp8" else weights_path, HERE),
            }
            if rel_own is not None:
                res["relmse_vs_own_converged"] = round(rel_own, 6)
            print(f"[bench] learned ({label}, {precision}, {S} spp, {steps} frames): relMSE {rel:.5f}, "
                  f"{res['frames_per_s']} frames/s", file=sys.stderr, flush=True)
            return res, el

        # the learned form's sample counts, each against the a-trous form at equal relMSE
        learned_sweep = []
        for precision in nets:
            for S in args.learned_spp:
                r, _ = learned_form(plain, "4096 VPLs", S, args.sweep_steps, 2, precision)
                r.update(atrous_at_equal(sweep, r))
                learned_sweep.append(r)
        # the operating point: the best ratio at equal relMSE among bracketed (interpolated) points
        ranked = [r for r in learned_sweep if r.get("atrous_at_equal_relmse", {}).get("interpolated")]
        if ranked:
            best = max(ranked, key=lambda r: r["speedup_vs_atrous_at_equal_relmse"])
        else:
            best = learned_sweep[len(learned_sweep) // 2]
        S_op, P_op = best["samples_per_pixel"], best["network_precision"]
        print(f"[bench] operating point {P_op} network, {S_op} spp; timing {args.steps} learned frames",
              file=sys.stderr, flush=True)
        learned_res, el_learned = learned_form(plain, "4096 VPLs", S_op, args.steps, args.warmup, P_op)
        learned_res.update(atrous_at_equal(sweep, learned_res))
    if learned_res is not None and not args.no_learned_extras:
        # the operating point with the fallback switched off (threshold +inf): the fallback's own cost
        nofb, _ = learned_form(plain, "4096 VPLs, no fallback", S_op, args.sweep_steps, 2, P_op,
                               form_threshold=float("inf"))
        nofb.update(atrous_at_equal(sweep, nofb))
        learned_no_fallback = {k: nofb[k] for k in ("frames_per_s", "relmse", "kernels_ms", "fallback_threshold")
                               if k in nofb}
        learned_no_fallback["speedup_vs_atrous_at_equal_relmse"] = nofb.get("speedup_vs_atrous_at_equal_relmse")
        # the same form reading the reference's packed G-buffer attachments (sgi_gbuffer.packed, 22 B
        # per pixel instead of the FP32 planes' 60), written once by sgi_gbuffer_pack outside the
        # timed frames; the relMSE against the same converged render (the attachments quantise
        # normals, albedo and material to 8 bits, as the reference's G-buffer does)
        packed_att = sc.PackedAttachments.pack(plain.gi, plain.dev)
        g_fp32 = plain.g
        plain.g = plain.dev.gbuffer_packed(packed_att)
        try:
            packed_res, _ = learned_form(plain, "4096 VPLs, packed G-buffer", S_op, args.sweep_steps, 2, P_op)
        finally:
            plain.g = g_fp32
        learned_packed = {k: packed_res[k] for k in ("frames_per_s", "ms_per_step", "relmse", "kernels_ms",
                                                     "fallback_tiles")}
        learned_packed["note"] = ("every pass reads RGB8 normals, RGBA8 albedo, RG8 reflectivity, RGB8 "
                                  "roughness / metallic and RG16F velocity directly; depth FP32")
        # the 1024-VPL scene: each sample count (1) against the 1024-VPL a-trous sweep, both forms on
        # the same scene at equal relMSE to its own converged render (same_scene), and (2) against the
        # 4096-VPL a-trous sweep at equal relMSE to the 4096-VPL converged render (whose distance from
        # the 1024-VPL one, converged_1024_vs_4096_relmse, floors the learned form's relMSE there)
        for S in args.learned_spp:
            r, _ = learned_form(learned, "1024 VPLs", S, args.sweep_steps, 2, P_op, own_target=target_1024)
            r.update(atrous_at_equal(sweep, r))
            own = atrous_at_equal(sweep_1024, {"relmse": r["relmse_vs_own_converged"],
                                               "frames_per_s": r["frames_per_s"]})
            r["same_scene"] = {k: own[k] for k in ("atrous_at_or_below_relmse", "atrous_above_relmse",
                                                   "atrous_at_equal_relmse", "speedup_vs_atrous_at_equal_relmse")
                               if k in own}
            learned_1024_sweep.append(r)
        learned_1024 = max(learned_1024_sweep,
                           key=lambda r: r["same_scene"].get("speedup_vs_atrous_at_equal_relmse", 0.0))

    gather_roof = None
    if use_tiles and per_plain is not None:
        alg_bytes, pairs = gather_algorithmic_bytes(tiles_np, w_plain.tile_size, W, H, w_plain.max_vpls_per_tile,
                                                    texel_bytes)
        achieved = alg_bytes / (per_plain[2] * 1e-3) / 1e9
        gather_roof = {"kernel": "gatherTilesKernel<F16,16> (plain form)", "bound": "hbm",
                       "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4),
                       "traffic": load_traffic("gather_city1080p_traffic.json"),
                       "algorithmic_bytes": alg_bytes, "pixel_vpl_pairs": pairs,
                       "avg_launch_ms": round(per_plain[2], 4)}
    plain_res = None
    if per_plain is not None:
        cull_roof = culling_roofline(per_plain, W, H, tiles_np.shape[0], w_plain.max_vpls_per_tile, visible_plain,
                                     w_plain.num_vpls)
        plain_res = {
            "frames_per_s": round(args.steps / el_plain, 2), "ms_per_step": round(el_plain * 1e3 / args.steps, 4),
            "relmse": round(plain_rel, 6),
            "kernels_ms": stage_ms(["cull_global", "cull_tiles", "gather", "denoise"], per_plain, direct_plain),
            "vpls": w_plain.num_vpls, "visible_vpls": visible_plain, "gather": args.gather,
        }
    conv_roof = None
    if learned_res is not None:
        value, el, form = learned_res["frames_per_s"], el_learned, "learned"
        if plain_res is not None:
            learned_res["speedup_vs_tile_list_plain"] = round(
                learned_res["frames_per_s"] / plain_res["frames_per_s"], 3)
        # the learned frame's dominant kernel: the sampled gather at the operating point
        S = learned_res["samples_per_pixel"]
        alg = samples_algorithmic_bytes(W, H, S, learned_res["visible_vpls"], texel_bytes)
        g_ms = learned_res["kernels_ms"]["gather"]
        achieved = alg / (g_ms * 1e-3) / 1e9
        roofline = {"kernel": f"gatherSamplesCompactKernel<F16> ({S} draws per pixel)", "bound": "hbm",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": load_traffic(f"gather_samples{S}_city1080p_traffic.json"),
                    "algorithmic_bytes": alg, "draws": W * H * S, "avg_launch_ms": g_ms,
                    "valu": valu_estimate(f"gather_samples{S}_city1080p_valu.json", g_ms, S),
                    "note": ("not HBM-bound: VALU issue (estimate beside it) and the vector memory pipeline "
                             "filling L1 lines for random reads (a binary16 VPL position per draw, a VPL record "
                             "and its cube texel per accepted draw) bound it; the HBM fraction is low by "
                             "construction")}
        conv_roof = conv_roofline(wts, H, W)
    elif plain_res is not None:
        value, el, form = plain_res["frames_per_s"], el_plain, "tile-list plain form only"
        roofline = gather_roof
    else:
        raise SystemExit("bench.py: nothing was timed (no trained weights for the learned form?)")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1000.0 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"fp32 shading, {learned_res['network_precision']} network (fp32 accumulation)"
                  if learned_res else "fp32"),
        "data": "synthetic: seeded procedural city block, ray-cast G-buffer / VPL cubes / cascades",
        "config": {"workload": "city-1920x1080, 4096 VPLs: learned (S spp + network) vs a-trous (S' spp) at "
                               "equal relMSE",
                   "learned_spp": learned_res["samples_per_pixel"] if learned_res else None,
                   "width": W, "height": H, "cube": f"6x{w_plain.cube_res}x{w_plain.cube_res} fp16",
                   "form": form, "converged": "all visible VPLs of the 4096-VPL scene, FP64 sums",
                   "converged_1024_vs_4096_relmse": round(floor_rel, 6)},
        "learned": learned_res,
        "learned_sweep": learned_sweep if learned_res is not None else [],
        "learned_without_fallback": learned_no_fallback,
        "learned_packed_gbuffer": learned_packed,
        "learned_1024_vpls": learned_1024,
        "learned_1024_sweep": learned_1024_sweep,
        "atrous_1024_sweep": [{k: v for k, v in r.items() if k != "kernels_ms"} for r in sweep_1024],
        "atrous_sweep": sweep,
        "plain_tile_lists": plain_res,
        "reference_estimator": random_res,
        "roofline": roofline,
        "roofline_network_conv": conv_roof,
        "roofline_plain_gather": gather_roof,
        "roofline_culling": cull_roof if per_plain is not None else None,
        "batched_views": None,
        "vpl_colouring": None,
        "roofline_gbuffer_unpack": None,
        "roofline_post_fx": None,
        "shadow_passes": None,
        "cpu_baseline": None,
    }
    if plain_res is not None:
        print("[bench] timing the VPL colouring pass", file=sys.stderr, flush=True)
        result["vpl_colouring"] = vpl_colouring(plain)
        result["roofline_gbuffer_unpack"] = gbuffer_unpack_roofline(plain)
        result["roofline_post_fx"] = post_fx_roofline(plain)
        print("[bench] timing the shadow pre-passes", file=sys.stderr, flush=True)
        result["shadow_passes"] = shadow_passes(plain)
    if not args.no_cpu_baseline and plain_res is not None:
        print("[bench] timing the CPU oracle on a sample", file=sys.stderr, flush=True)
        host = sc.to_host(plain.dev)
        eq = (learned_res or {}).get("atrous_at_equal_relmse", {})
        spp_cpu = int(round(eq["spp"])) if eq.get("spp") else max(args.atrous_spp)
        result["cpu_baseline"] = cpu_baseline(host, w_plain, args.cpu_rows, visible_plain, spp_cpu,
                                              args.rejection_limit, args.cpu_threads)
        del host
    if args.views > 0 and args.forms == "all":
        del plain, learned
        torch.cuda.empty_cache()
        bl = None
        if learned_res is not None:
            bl = (learned_res["samples_per_pixel"], (f8net.net() if learned_res["network_precision"] == "fp8"
                                                      else wts.net()), threshold, learned_res["network_precision"])
        result["batched_views"] = batched_views(args, bl)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
