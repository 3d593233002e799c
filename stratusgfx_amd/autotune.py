"""Autotuner of the batched views' launch shape: the tile size (8 / 16 / 32)
and the per-tile VPL cap K of a batch, picked by measurement.

Each candidate renders the batch through a captured graph (sgi_views_capture
/ sgi_views_replay); its time is the mean of HIP-event-timed replays and its
quality the relMSE of every view's GI output against that view's converged
render (every visible VPL, FP64 sums).  The tile lists change which VPLs a
pixel sees, so the cap and the tile size trade time for bias: the tuner picks
the fastest candidate whose relMSE is within `quality_tol` (relative) of the
reference shape's (16x16 tiles, K = 64).
"""
from __future__ import annotations

import json
import os

from . import _abi
from .gi import GiPass, Views, view_io

CANDIDATES = [(ts, k) for ts in (8, 16, 32) for k in (16, 32, 64)]
REFERENCE = (16, 64)


def rel_mse(x, ref) -> float:
    x, r = x.double(), ref.double()
    return float(((x - r) ** 2 / (r * r + 1e-2)).mean())


def converged(inp, shared, num_vpls):
    """The view's every-visible-VPL render (normalised radiance), FP64 sums."""
    import torch

    H, W = inp.height, inp.width
    gi = GiPass(W, H, 16, 64)
    gi.set_frame(inp.frame)
    gi.set_csm(inp.csm_struct())
    f32 = dict(dtype=torch.float32, device="cuda")
    upd = torch.zeros((num_vpls, 16), **f32)
    vis = torch.zeros((1 + num_vpls,), dtype=torch.int32, device="cuda")
    gi.cull_global(shared.vpls, num_vpls, upd, vis)
    ent = torch.zeros((num_vpls, 2), dtype=torch.int32, device="cuda")
    gi.visible_atlas_entries(shared.entries, vis, num_vpls, ent)
    col, res = torch.zeros((H, W, 3), **f32), torch.zeros((H, W, 4), **f32)
    gi.gather_ex(inp.gbuffer(), upd, vis, ent, shared.cubes(), None, col, res, mode=_abi.GATHER_ALL,
                 accumulate_fp64=True)
    torch.cuda.synchronize()
    return (res[..., :3] / res[..., 3:4].clamp_min(1e-5)).contiguous()


def measure(inputs, shared, num_vpls, tile_size, K, targets=None, reps=10):
    """(ms per batch, mean relMSE or None) of one launch shape."""
    import torch

    f32 = dict(dtype=torch.float32, device="cuda")
    bufs, ios = [], []
    for inp in inputs:
        H, W = inp.height, inp.width
        nt = ((W + tile_size - 1) // tile_size) * ((H + tile_size - 1) // tile_size)
        b = dict(tiles=torch.zeros((nt, 1 + K), dtype=torch.int32, device="cuda"), direct=torch.zeros((H, W, 3), **f32),
                 color=torch.zeros((H, W, 3), **f32), reservoir=torch.zeros((H, W, 4), **f32),
                 combined=torch.zeros((H, W, 3), **f32), gi=torch.zeros((H, W, 3), **f32),
                 reservoir_out=torch.zeros((H, W, 4), **f32), history=torch.zeros((H, W), **f32))
        bufs.append(b)
        ios.append(view_io(inp, **b))
    total = sum(inp.width * inp.height for inp in inputs)
    vw = Views(len(inputs), total, tile_size, K)
    s = torch.cuda.Stream()
    vw.set(ios, shared.vpls, num_vpls, shared.entries, shared.cubes(), stream=s.cuda_stream)
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
    rel = None
    if targets is not None:
        rel = sum(rel_mse(b["gi"], t) for b, t in zip(bufs, targets)) / len(targets)
    vw.close()
    return ms, rel


def tune(inputs, shared, num_vpls, candidates=CANDIDATES, quality_tol=0.05, reps=10, cache=None) -> dict:
    """Measures every candidate on the batch and returns the choice with the whole table."""
    targets = [converged(inp, shared, num_vpls) for inp in inputs]
    table = []
    for ts, k in candidates:
        ms, rel = measure(inputs, shared, num_vpls, ts, k, targets, reps)
        table.append({"tile_size": ts, "max_vpls_per_tile": k, "ms_per_batch": round(ms, 4), "relmse": rel})
    ref = next((r for r in table if (r["tile_size"], r["max_vpls_per_tile"]) == REFERENCE), table[0])
    ok = [r for r in table if r["relmse"] <= ref["relmse"] * (1.0 + quality_tol)]
    best = min(ok, key=lambda r: r["ms_per_batch"])
    out = {"choice": best, "reference": ref, "quality_tol": quality_tol, "table": table,
           "views": [[inp.width, inp.height] for inp in inputs]}
    if cache:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(out, f, indent=1)
    return out
