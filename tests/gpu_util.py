"""Shared helpers of the GPU parity tests: device inputs, their host copies and
the oracle's outputs on the same inputs."""
import dataclasses

import numpy as np

import oracle as orc
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
