#!/usr/bin/env python3
"""GI frames/s of the deferred VPL GI pass on one MI355X.

A step is one GI frame of the pass on resident inputs, in RenderScene's order
(StratusRendererBackend.cpp:1982, 2010-2045): global VPL sun-visibility cull
and compaction, the visible slots' atlas entries, deferred direct sun light
with cascaded PCF, per-tile VPL culling, the per-pixel VPL gather and the
4-pass a-trous / temporal denoiser with the combine.  The workload is the
procedural city block at 1920x1080 with 1024 VPLs (6x256x256 FP16 distance
cubes, 16x16 tiles, up to 64 VPLs per tile).  Inputs are generated on the GPU
before the timed region; the denoiser's history alternates between two buffer
sets as the reference swaps its FBOs.

Prints one JSON line (rank 0).  With the default flags it also times the CPU
oracle on a bounded sample of the same frame as the CPU baseline.
"""
import argparse
import json
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
    ap.add_argument("--gather", choices=["tiles", "random"], default="tiles")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing")
    ap.add_argument("--cpu-rows", type=int, default=24, help="image rows of the CPU baseline sample")
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


def load_traffic(name):
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("traffic_bytes_per_launch")


def cpu_baseline(host, w, rows, visible_count):
    """Times the oracle (single-threaded FP32 port) on a bounded sample of the frame."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import ctypes

    import numpy as np

    import oracle as orc

    H, W = w.height, w.width
    y0 = max(0, H // 2 - rows // 2)
    y1 = min(H, y0 + rows)
    g, cubes, csm = host.gbuffer(), host.cubes(), host.csm_struct()
    t = {}
    t0 = time.perf_counter()
    upd, vis = orc.cull_global(host.frame, csm, host.vpls)
    t["cull_global"] = time.perf_counter() - t0
    ent = host.entries[vis[1:]]
    t0 = time.perf_counter()
    _, tiles = orc.cull_tiles(host.frame, g, w.tile_size, w.max_vpls_per_tile, upd, vis)
    t["cull_tiles"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    lit = orc.direct_light(host.frame, csm, g, y0, y1)
    t["direct_light"] = (time.perf_counter() - t0) * H / (y1 - y0)
    t0 = time.perf_counter()
    col, res = orc.gather(host.frame, g, upd, int(vis[0]), ent, cubes, tiles=tiles, tile_size=w.tile_size,
                          K=w.max_vpls_per_tile, mode=orc.GATHER_TILES, y0=y0, y1=y1)
    t["gather"] = (time.perf_counter() - t0) * H / (y1 - y0)
    # the denoiser on the sampled rows as an image of their own (same per-pixel work)
    crop = lambda a: np.ascontiguousarray(a[y0:y1])
    planes = [crop(host.depth), crop(host.normal), crop(host.albedo), crop(host.refl), crop(host.rma),
              crop(host.velocity), crop(host.ids)]
    from stratusgfx_amd import _abi
    cg = _abi.Gbuffer(W, y1 - y0, *[p.ctypes.data for p in planes], 0)
    t0 = time.perf_counter()
    orc.denoise(host.frame, cg, None, int(vis[0]), lit, col, res)
    t["denoise"] = (time.perf_counter() - t0) * H / (y1 - y0)
    total = sum(t.values())
    return {
        "value": 1.0 / total, "unit": "frames/s", "cores": 1, "kind": "port",
        "sample": (f"FP32 oracle, 1 thread: cull_global ({visible_count} visible of {w.num_vpls}) and tile culling "
                   f"on the whole frame; direct light, tile-list gather and denoiser on rows {y0}..{y1} of {H}, "
                   f"scaled by {H / (y1 - y0):.1f} to a frame"),
        "seconds_per_frame": {k: round(v, 3) for k, v in t.items()},
    }


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)
        return

    from stratusgfx_amd import _abi
    from stratusgfx_amd import scene as sc
    from stratusgfx_amd.gi import FrameBuffers, GiPass

    w = sc.CITY_1080P
    print(f"[bench] generating {w.name} on the GPU", file=sys.stderr, flush=True)
    scene = sc.Scene(w)
    dev = sc.generate_device(scene)
    gi = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    gi.set_frame(dev.frame)
    gi.set_csm(dev.csm_struct())
    g, cubes = dev.gbuffer(), dev.cubes()
    n = w.num_vpls
    cap = min(n, _abi.MAX_TOTAL_VPLS_PER_FRAME)
    fb = FrameBuffers.allocate(w.width, w.height, n, gi.num_tiles, w.max_vpls_per_tile)
    hist = [(torch.zeros_like(fb.gi), torch.zeros_like(fb.history)) for _ in range(2)]
    use_tiles = args.gather == "tiles"
    stages = ["cull_global", "direct_light", "cull_tiles", "gather", "denoise"]

    def step(k, ev=None):
        prev_gi, prev_h = hist[k % 2]
        out_gi, out_h = hist[(k + 1) % 2]
        rec = (lambda i: ev[i].record()) if ev is not None else (lambda i: None)
        rec(0)
        gi.cull_global(dev.vpls, n, fb.vpl_updated, fb.visible)
        gi.visible_atlas_entries(dev.entries, fb.visible, cap, fb.visible_entries)
        rec(1)
        gi.direct_light(g, fb.direct)
        rec(2)
        if use_tiles:
            gi.cull_tiles(g, fb.vpl_updated, fb.visible, fb.tiles)
        rec(3)
        gi.gather(g, fb.vpl_updated, fb.visible, fb.visible_entries, cubes, fb.tiles if use_tiles else None,
                  fb.color, fb.reservoir)
        rec(4)
        gi.denoise(g, None, fb.visible, fb.direct, fb.color, fb.reservoir, prev_gi, prev_h, fb.combined, out_gi,
                   fb.reservoir_out, out_h)
        rec(5)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(args.steps)]
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k, events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_stage = {s: sum(events[k][i].elapsed_time(events[k][i + 1]) for k in range(args.steps)) / args.steps
                 for i, s in enumerate(stages)}
    visible = int(fb.visible[0].item())
    tiles_np = fb.tiles.cpu().numpy()
    texel_bytes = 2 if w.cube_format == _abi.TEXEL_F16 else 4
    gather_ms = per_stage["gather"]
    roofline = None
    if use_tiles:
        alg_bytes, pairs = gather_algorithmic_bytes(tiles_np, w.tile_size, w.width, w.height, w.max_vpls_per_tile,
                                                    texel_bytes)
        achieved = alg_bytes / (gather_ms * 1e-3) / 1e9
        roofline = {"kernel": "gatherTilesKernel", "bound": "hbm", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": load_traffic("gather_city1080p_traffic.json"),
                    "algorithmic_bytes": alg_bytes, "pixel_vpl_pairs": pairs}
    result = {
        "metric": METRIC,
        "value": round(args.steps / elapsed, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000.0 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded procedural city block, ray-cast G-buffer / VPL cubes / cascades",
        "config": {"workload": w.name + ("-tiles16-K64" if use_tiles else "-random"), "width": w.width,
                   "height": w.height, "vpls": n, "visible_vpls": visible, "cube": f"6x{w.cube_res}x{w.cube_res} fp16",
                   "tile_size": w.tile_size, "max_vpls_per_tile": w.max_vpls_per_tile, "gather": args.gather,
                   "form": "plain FP32 restatement + a-trous (learned form not built)"},
        "kernels_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle on a sample", file=sys.stderr, flush=True)
        host = sc.to_host(dev)
        result["cpu_baseline"] = cpu_baseline(host, w, args.cpu_rows, visible)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
