"""Shared helpers of the GPU parity tests: device inputs, their host copies and
the oracle's outputs on the same inputs."""
import dataclasses

import numpy as np

import oracle as orc
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc


def require_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device; there is no CPU fallback"
    return torch


def small(w: sc.Workload, **kw) -> sc.Workload:
    return dataclasses.replace(w, **kw)


class Case:
    """One workload's inputs on the device and the host."""

    def __init__(self, w: sc.Workload):
        self.torch = require_gpu()
        self.w = w
        self.scene = sc.Scene(w)
        self.dev = sc.generate_device(self.scene)
        self.host = sc.to_host(self.dev)
        self.hg = self.host.gbuffer()
        self.hcubes = self.host.cubes()
        self.hcsm = self.host.csm_struct()
        print(f"[case] {w.name} {w.width}x{w.height} vpls={w.num_vpls} K={w.max_vpls_per_tile}", flush=True)

    def oracle_cull(self):
        upd, vis = orc.cull_global(self.host.frame, self.hcsm, self.host.vpls)
        ent = self.host.entries[vis[1:]] if vis[0] > 0 else np.zeros((0, 2), np.int32)
        return upd, vis, ent


def assert_tiles_equal(got: np.ndarray, want: np.ndarray):
    assert got.shape == want.shape
    for t in range(want.shape[0]):
        n = int(want[t, 0])
        assert int(got[t, 0]) == n, f"tile {t}: count {got[t, 0]} != {n}"
        assert np.array_equal(got[t, 1:1 + n], want[t, 1:1 + n]), f"tile {t}: {got[t, 1:1 + n]} != {want[t, 1:1 + n]}"


def learned_single(inp, shared, n, net, thr, S, tile, K):
    """One view alone through the learned chain: culling, atlas entries, direct light, the sampled
    gather and sgi_denoise_learned."""
    from stratusgfx_amd.gi import GiPass

    torch = require_gpu()
    H, W = inp.height, inp.width
    f32 = dict(dtype=torch.float32, device="cuda")
    gi = GiPass(W, H, tile, K)
    gi.set_frame(inp.frame)
    gi.set_csm(inp.csm_struct())
    upd = torch.zeros((n, 16), **f32)
    vis = torch.zeros((1 + n,), dtype=torch.int32, device="cuda")
    gi.cull_global(shared.vpls, n, upd, vis)
    ent = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    gi.visible_atlas_entries(shared.entries, vis, min(n, _abi.MAX_TOTAL_VPLS_PER_FRAME), ent)
    g = inp.gbuffer()
    out = dict(direct=torch.zeros((H, W, 3), **f32), color=torch.zeros((H, W, 3), **f32),
               reservoir=torch.zeros((H, W, 4), **f32), combined=torch.zeros((H, W, 3), **f32),
               gi=torch.zeros((H, W, 3), **f32), reservoir_out=torch.zeros((H, W, 4), **f32),
               history=torch.zeros((H, W), **f32))
    gi.direct_light(g, out["direct"])
    gi.gather_ex(g, upd, vis, ent, shared.cubes(), None, out["color"], out["reservoir"], mode=_abi.GATHER_SAMPLES,
                 samples=S, rejection_limit=-1)
    ws = gi.kpcn_workspace()
    gi.denoise_learned(g, None, vis, out["direct"], out["color"], out["reservoir"], None, None, net, thr, ws,
                       out["combined"], out["gi"], out["reservoir_out"], out["history"])
    torch.cuda.synchronize()
    return out


def shipped_nets():
    """The shipped network (BF16 and FP8 forms) and its fallback threshold."""
    import json
    import os

    from stratusgfx_amd import kpcn, kpcn_fp8, quantize_kpcn
    wts = kpcn.Weights(kpcn.load_params())
    _, _, w8, epi = kpcn_fp8.load(quantize_kpcn.DEFAULT_FP8)
    f8 = kpcn_fp8.Fp8Net(wts, w8, epi)
    with open(os.path.splitext(kpcn.DEFAULT_WEIGHTS)[0] + ".json") as fh:
        thr = float(json.load(fh)["threshold"])
    return wts, f8, thr
