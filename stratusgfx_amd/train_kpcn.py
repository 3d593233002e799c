"""Trains the learned denoiser on pairs the gather itself generates, on the GPU.

    python -m stratusgfx_amd.train_kpcn --scenes 64 --views 2 --steps 12000 --batch 2

Data: seed-generated city blocks (seeds 1000 + i) at the benchmark's frame
size and cube resolution, each block's default pose and random poses over
it; input = a sampled gather of 4096 VPLs (sgi_gather_ex SAMPLES, failing
draws skipped, so that more samples converge to the target) at a sample
count drawn per view from --spp, target = the converged gather over every
visible VPL with FP64 sums (sgi_gather_ex ALL), both normalised as the
denoiser's pass 0 does.  Mixing sample counts lets one network serve every
operating point: it reads the noise level from the input itself.
Training: Adam on the relative-MSE loss plus the error head's regression,
gradients from the project's MFMA kernels.  The error channel learns, per
pixel, the log ratio of the network's relative squared error to that of the
a-trous filter on the same input (sgi_kpcn_loss_ratio; --err-target error:
the network's own log error, sgi_kpcn_loss), so that a tile's mean of
exp(err) predicts whether the a-trous filter would do better there.
--train error-head updates only that channel's taps and bias, which leaves
the network's output unchanged: it calibrates the fallback of a trained
network.  Validation on held-out seeds, at every sample count of --spp,
reports the relative MSE of the noisy input, the a-trous filter
(sgi_denoise), the network alone, and the network with its a-trous fallback
at each threshold, and picks the threshold with the lowest relative MSE.
Writes the FP32 parameters (.npy) and a JSON summary beside them.

    python -m stratusgfx_amd.train_kpcn --init stratusgfx_amd/weights/kpcn_city.npy \
        --train error-head --err-weight 1 --steps 3000 --scenes 256 --views 2
"""
from __future__ import annotations

import argparse
import ctypes
import dataclasses
import json
import math
import os
import time

import numpy as np

from . import _abi, kpcn
from . import scene as sc
from .gi import GiPass


def rel_mse(x, ref) -> float:
    """mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2), in FP64."""
    x = x.double()
    r = ref.double()
    return float(((x - r) ** 2 / (r * r + 1e-2)).mean())


def _inside_box(boxes: np.ndarray, p, margin=0.75) -> bool:
    c, h = boxes[:, 0:3], boxes[:, 4:7]
    d = np.abs(np.asarray(p, np.float32)[None, :] - c)
    return bool(np.any(np.all(d <= h + margin, axis=1)))


def random_view(scene: sc.Scene, rng, width: int, height: int | None = None) -> _abi.View:
    """A pose uniform over the block at 1.5-30 m, random yaw / pitch, outside every box."""
    n = scene.lib.sgi_scene_num_boxes(scene.handle)
    boxes = np.zeros((n, 16), np.float32)
    _abi.call(scene.lib, "sgi_scene_boxes", scene.handle, boxes.ctypes.data)
    v = _abi.View()
    _abi.call(scene.lib, "sgi_scene_default_view", scene.handle, width, height if height else width,
              ctypes.byref(v))
    e = scene.w.city_half_extent * 0.9
    for _ in range(100):
        eye = (rng.uniform(-e, e), rng.uniform(1.5, 30.0), rng.uniform(-e, e))
        if not _inside_box(boxes, eye):
            break
    v.eye[0], v.eye[1], v.eye[2] = eye
    v.yaw_deg = float(rng.uniform(0.0, 360.0))
    v.pitch_deg = float(rng.uniform(-35.0, 5.0))
    return v


class Sample:
    """One view's network inputs and targets on the device."""

    def __init__(self, feat, rad, target, extra=None, spp=0, alt=None):
        self.feat, self.rad, self.target, self.extra, self.spp = feat, rad, target, extra, spp
        self.alt = alt  # the a-trous filter's output on the same input (the error head's reference)


def make_samples(seed: int, views: int, width: int, height: int, num_vpls: int, spp_choices, rng, keep_frame=False,
                 rejection_limit=-1, cube_res=256, default_view=True, with_alt=False):
    """`views` views of one seeded block (the block's default pose first when
    default_view, then random poses); each view's gather takes a sample count
    drawn from spp_choices (every count when keep_frame, for validation)."""
    import torch

    w = sc.Workload(f"city-train-{seed}", _abi.SCENE_CITY, seed, width, height, num_vpls, cube_res, _abi.TEXEL_F16,
                    1024, 16, 64)
    scene = sc.Scene(w)
    out = []
    inp = None
    poses = []
    for k in range(views):
        if default_view and k == 0:
            v = _abi.View()
            _abi.call(scene.lib, "sgi_scene_default_view", scene.handle, width, height, ctypes.byref(v))
        else:
            v = random_view(scene, rng, width, height)
        poses.append((v, list(spp_choices) if keep_frame else [int(rng.choice(spp_choices))]))
    for v, counts in poses:
        scene.set_view(v)
        inp = sc.generate_device(scene)
        for samples_pp in counts:
            out.append(_sample(inp, width, height, num_vpls, samples_pp, rejection_limit, keep_frame, with_alt))
        del inp
    torch.cuda.synchronize()
    return out


def _sample(inp, width, height, num_vpls, samples_pp, rejection_limit, keep_frame, with_alt=False):
    """Network inputs and the converged target of one view at one sample count."""
    import torch

    gi = GiPass(width, height, 16, 64)
    gi.set_frame(inp.frame)
    gi.set_csm(inp.csm_struct())
    f32 = dict(dtype=torch.float32, device="cuda")
    n = num_vpls
    upd = torch.zeros((n, 16), **f32)
    vis = torch.zeros((1 + n,), dtype=torch.int32, device="cuda")
    gi.cull_global(inp.vpls, n, upd, vis)
    ent = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    gi.visible_atlas_entries(inp.entries, vis, n, ent)
    g = inp.gbuffer()
    col, res = torch.zeros((height, width, 3), **f32), torch.zeros((height, width, 4), **f32)
    gi.gather_ex(g, upd, vis, ent, inp.cubes(), None, col, res, mode=_abi.GATHER_SAMPLES, samples=samples_pp,
                 rejection_limit=rejection_limit)
    colT, resT = torch.zeros((height, width, 3), **f32), torch.zeros((height, width, 4), **f32)
    gi.gather_ex(g, upd, vis, ent, inp.cubes(), None, colT, resT, mode=_abi.GATHER_ALL, accumulate_fp64=True)
    feat = torch.empty((height, width, 32), dtype=torch.bfloat16, device="cuda")
    rad = torch.empty((height, width, 3), **f32)
    gi.kpcn_features(g, res, feat, rad)
    target = resT[..., :3] / resT[..., 3:4].clamp_min(1e-5)
    alt = None
    if with_alt:  # the a-trous filter on the same input (its GI output does not read the direct light)
        outs = [torch.zeros((height, width, c), **f32) for c in (3, 3, 4)]
        hist = torch.zeros((height, width), **f32)
        gi.denoise(g, None, vis, torch.zeros((height, width, 3), **f32), col, res, None, None, outs[0], outs[1],
                   outs[2], hist)
        alt = outs[1]
        del outs, hist
    extra = None
    if keep_frame:
        lit = torch.zeros((height, width, 3), **f32)
        gi.direct_light(g, lit)
        # the G-buffer and the gather's outputs stay; the cubes are not needed again
        extra = dict(gi=gi, inp=dataclasses.replace(inp, atlases=[]), vis=vis, col=col, res=res, lit=lit)
    return Sample(feat, rad, target.contiguous(), extra, samples_pp, alt)


def evaluate(wts: kpcn.Weights, val, thresholds):
    """relMSE of noisy / a-trous / network / network + fallback at each threshold, averaged over
    views, and the noisy / a-trous / network relMSE per sample count."""
    import torch

    acc = {"noisy": 0.0, "atrous": 0.0, "network": 0.0}
    per_spp = {}
    fb = {t: [0.0, 0.0] for t in thresholds}
    for s in val:
        before = dict(acc)
        e = s.extra
        gi, inp = e["gi"], e["inp"]
        g = inp.gbuffer()
        H, W = s.rad.shape[:2]
        f32 = dict(dtype=torch.float32, device="cuda")
        outs = [torch.zeros((H, W, c), **f32) for c in (3, 3, 4)]
        hist = torch.zeros((H, W), **f32)
        gi.denoise(g, None, e["vis"], e["lit"], e["col"], e["res"], None, None, outs[0], outs[1], outs[2], hist)
        acc["atrous"] += rel_mse(outs[1], s.target)
        acc["noisy"] += rel_mse(s.rad, s.target)
        ws = gi.kpcn_workspace()
        n_tiles = ((W + 15) // 16) * ((H + 15) // 16)
        fbt = torch.zeros((1 + n_tiles,), dtype=torch.int32, device="cuda")
        gi.denoise_learned(g, None, e["vis"], e["lit"], e["col"], e["res"], None, None, wts.net(), float("inf"), ws,
                           outs[0], outs[1], outs[2], hist, fbt)
        acc["network"] += rel_mse(outs[1], s.target)
        d = per_spp.setdefault(str(s.spp), {"views": 0, "noisy": 0.0, "atrous": 0.0, "network": 0.0})
        d["views"] += 1
        for k in ("noisy", "atrous", "network"):
            d[k] += acc[k] - before[k]
        for t in thresholds:
            gi.denoise_learned(g, None, e["vis"], e["lit"], e["col"], e["res"], None, None, wts.net(), t, ws,
                               outs[0], outs[1], outs[2], hist, fbt)
            fb[t][0] += rel_mse(outs[1], s.target)
            fb[t][1] += int(fbt[0]) / n_tiles
    n = len(val)
    res = {k: v / n for k, v in acc.items()}
    res["per_spp"] = {k: {m: (v[m] / v["views"] if m != "views" else v[m]) for m in v} for k, v in per_spp.items()}
    res["fallback"] = {f"{t:g}": {"relmse": v[0] / n, "tiles": v[1] / n} for t, v in fb.items()}
    return res


def pick_threshold(val: dict, thresholds, tol=1e-3, max_tiles=0.01):
    """The fallback threshold from a validation result (evaluate()): the smallest threshold whose
    relMSE is within `tol` (relative) of the best of the network alone and every threshold, and at
    which at most `max_tiles` of the held-out tiles fall back, so that the fallback acts as early as
    held-out data allows without costing more than `tol` in error, while on data like the training
    data (where the network rarely fails) it stays rare and costs little time; +inf when no
    threshold qualifies.  Returns (threshold, its validation relMSE)."""
    rel = {float("inf"): val["network"]}
    for t in thresholds:
        f = val["fallback"][f"{t:g}"]
        if f["tiles"] <= max_tiles:
            rel[t] = f["relmse"]
    best = min([val["network"]] + [val["fallback"][f"{t:g}"]["relmse"] for t in thresholds])
    ok = [k for k, v in rel.items() if v <= best * (1.0 + tol)]
    t = min(ok) if ok else float("inf")
    return t, rel[t]


def main(argv=None) -> None:
    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--views", type=int, default=2)
    ap.add_argument("--val-scenes", type=int, default=4)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cube-res", type=int, default=256, help="VPL distance-cube resolution of the training scenes")
    ap.add_argument("--vpls", type=int, default=4096)
    ap.add_argument("--spp", type=int, nargs="+", default=[4, 8, 16, 32],
                    help="sample counts of the input gathers (one drawn per training view)")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--err-weight", type=float, default=1e-4,
                    help="weight of the error head's loss; its gradient otherwise swamps the filter's")
    ap.add_argument("--loss-scale", type=float, default=1000.0,
                    help="gradient scale (the relative-MSE gradients are ~1e-9 per logit at these radiances)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these parameters (a weights .npy) instead of a fresh network")
    ap.add_argument("--first-scene", type=int, default=1000, help="seed of the first training block")
    ap.add_argument("--out", default=kpcn.DEFAULT_WEIGHTS)
    ap.add_argument("--log", default=None)
    ap.add_argument("--err-target", choices=["error", "ratio"], default="ratio",
                    help="the error channel's target: the log relative squared error of the output (error), or "
                         "its log ratio to the a-trous filter's on the same input (ratio, sgi_kpcn_loss_ratio)")
    ap.add_argument("--train", choices=["all", "error-head"], default="all",
                    help="update every parameter, or only the error channel of the level-0 head (the rest of "
                         "the network, and so its output, unchanged: calibrating a trained network's fallback)")
    a = ap.parse_args(argv)
    rng = np.random.default_rng(a.seed)
    t0 = time.time()
    train = []
    for i in range(a.scenes):
        train += make_samples(a.first_scene + i, a.views, a.width, a.height, a.vpls, a.spp, rng, cube_res=a.cube_res,
                              with_alt=a.err_target == "ratio")
        if i % 8 == 7:
            print(f"[data] {len(train)} training views, {time.time() - t0:.0f} s", flush=True)
    val = []
    for i in range(a.val_scenes):
        val += make_samples(5000 + i, 1, a.width, a.height, a.vpls, a.spp, rng, keep_frame=True, cube_res=a.cube_res,
                            default_view=False)
    tmed = float(torch.stack([s.target.median() for s in train]).median())
    print(f"[data] {len(train)} train / {len(val)} val views in {time.time() - t0:.0f} s; median target {tmed:.4g}",
          flush=True)

    wts = kpcn.Weights(kpcn.load_params(a.init) if a.init else kpcn.init_params(a.seed))
    tr = kpcn.Trainer(wts, a.height, a.width)
    mask = None
    if a.train == "error-head":  # the error channel's taps and bias in the level-0 head
        m = np.zeros(kpcn.PARAMS, np.float32)
        base_l = kpcn.HEADS[0] * kpcn.LAYER_PARAMS
        for t in range(9):
            o = base_l + t * kpcn.C * kpcn.C + _abi.KPCN_ERR_CHANNEL * kpcn.C
            m[o:o + kpcn.C] = 1.0
        m[base_l + kpcn.LAYER_WEIGHTS + _abi.KPCN_ERR_CHANNEL] = 1.0
        mask = torch.from_numpy(m).cuda()
    base = evaluate(wts, val, [])
    print(f"[val] before training {json.dumps(base)}", flush=True)
    hist = []
    t1 = time.time()
    for step in range(1, a.steps + 1):
        lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * (step - 1) / a.steps))
        tr.zero_grad()
        for j in rng.choice(len(train), a.batch, replace=False):
            s = train[j]
            tr.forward(s.feat, s.rad)
            tr.loss_backward(s.target, err_weight=a.err_weight, scale=a.loss_scale / a.batch, alternative=s.alt)
        if mask is not None:
            tr.grad.mul_(mask)
        tr.step(lr)
        if step % 50 == 0 or step == 1:
            ls = tr.loss_sums.cpu().numpy() / (a.batch * a.width * a.height)
            hist.append({"step": step, "relmse": float(ls[0] / 3.0), "err_loss": float(ls[1]), "lr": lr})
            print(f"[train] step {step} relMSE {ls[0] / 3.0:.5f} err-head {ls[1]:.4f} lr {lr:.2e} "
                  f"{time.time() - t1:.0f} s", flush=True)
    torch.cuda.synchronize()
    train_s = time.time() - t1
    # thresholds on a tile's mean of exp(err): the predicted relative squared error (error), or the
    # predicted ratio of the network's to the a-trous filter's (ratio)
    thresholds = ([0.5, 0.7, 0.85, 1.0, 1.2, 1.5, 2.0, 3.0] if a.err_target == "ratio"
                  else [0.01, 0.02, 0.05, 0.1, 0.2, 0.5])
    after = evaluate(wts, val, thresholds)
    print(f"[val] after training {json.dumps(after)}", flush=True)
    best_t, best = pick_threshold(after, thresholds)
    kpcn.save_params(wts.params.cpu().numpy(), a.out)
    summary = {"train_views": len(train), "val_views": len(val), "width": a.width, "height": a.height, "cube_res": a.cube_res, "vpls": a.vpls, "spp": a.spp,
               "steps": a.steps, "batch": a.batch, "lr": a.lr, "err_weight": a.err_weight,
               "init": os.path.basename(a.init) if a.init else None, "first_scene": a.first_scene,
               "loss_scale": a.loss_scale, "err_target": a.err_target, "trained": a.train,
               "train_seconds": train_s, "data_seconds": t1 - t0, "median_target": tmed,
               "before": base, "after": after, "threshold": best_t, "history_points": len(hist)}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    if a.log:  # the loss curve stays out of the weights folder
        with open(a.log, "w") as f:
            json.dump(dict(summary, history=hist), f, indent=1)
    print(f"[done] threshold {best_t} relMSE {best:.5f}; weights {a.out}", flush=True)


if __name__ == "__main__":
    main()
