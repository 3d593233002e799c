"""Builds the FP8 inference form of a trained denoiser: calibrates the
per-channel activation scales on fresh city views (seeds apart from the
training and validation ones, mixed sample counts) and writes the E4M3 weights
and epilogues next to the BF16 weights.

    python -m stratusgfx_amd.quantize_kpcn [--weights W.npy] [--out W_fp8.npz]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import kpcn, kpcn_fp8
from .train_kpcn import make_samples

DEFAULT_FP8 = os.path.splitext(kpcn.DEFAULT_WEIGHTS)[0] + "_fp8.npz"


def main(argv=None) -> None:
    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=kpcn.DEFAULT_WEIGHTS)
    ap.add_argument("--out", default=DEFAULT_FP8)
    ap.add_argument("--scenes", type=int, default=8)
    ap.add_argument("--views", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--vpls", type=int, default=4096)
    ap.add_argument("--spp", type=int, nargs="+", default=[4, 8, 16, 32])
    ap.add_argument("--seed", type=int, default=5000, help="first calibration scene seed")
    args = ap.parse_args(argv)

    t0 = time.time()
    params = kpcn.load_params(args.weights)
    wts = kpcn.Weights(params)
    rng = np.random.default_rng(args.seed)
    samples = []
    for k in range(args.scenes):
        samples += make_samples(args.seed + k, args.views, args.width, args.height, args.vpls, args.spp, rng,
                                default_view=False)
        print(f"[calib] scene {k + 1}/{args.scenes}", flush=True)
    frames = len(samples)
    maxima = kpcn_fp8.calibrate(wts, [(smp.feat, smp.rad) for smp in samples])
    del samples
    torch.cuda.empty_cache()
    S, Cs = kpcn_fp8.group_scales(maxima)
    w8, epi = kpcn_fp8.quantize(params, S, Cs)
    kpcn_fp8.save(args.out, S, Cs, w8, epi)
    meta = {"weights": os.path.basename(args.weights), "frames": frames, "scenes": args.scenes,
            "views": args.views, "spp": args.spp, "seed": args.seed, "headroom": kpcn_fp8.HEADROOM,
            "seconds": round(time.time() - t0, 1)}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(f"[done] {args.out} from {frames} frames", flush=True)


if __name__ == "__main__":
    main()
