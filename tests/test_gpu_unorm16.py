"""The reference's own VPL cube format, GL_DEPTH_COMPONENT16 (UNORM16,
StratusRendererBackend.cpp:296,2412 with StratusTexture.cpp:542), on the GPU.

On the Cornell box generated with UNORM16 distance cubes:
- the device generator writes the host generator's codes, texel for texel;
- every gather mode (the reference's estimator, tile lists, every visible VPL,
  the fixed-count sampler in FP32 and packed-FP16 shading) matches the oracle
  reading the same codes (images within rel-L2 1e-4, the plain form's bar);
- the colouring pass (vpl_light_color.cs) reads the UNORM16 distance texel
  toward -sun into records bit for bit the oracle's.
The decode itself is pinned on every code on the CPU (tests/test_unorm16_cpu.py)."""
import dataclasses

import numpy as np
import pytest

import oracle as orc
from gpu_util import Case, require_gpu
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import GiPass

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4
W = dataclasses.replace(sc.CORNELL_640x360, name="cornell-640x360-64vpl-unorm16", cube_format=_abi.TEXEL_UNORM16)


@pytest.fixture(scope="module")
def case():
    return Case(W)


@pytest.fixture(scope="module")
def gi(case):
    p = GiPass(W.width, W.height, W.tile_size, W.max_vpls_per_tile)
    p.set_frame(case.dev.frame)
    p.set_csm(case.dev.csm_struct())
    return p


def test_device_generator_writes_host_codes(case):
    host = sc.generate_host(sc.Scene(W))
    for d, h in zip(case.dev.atlases, host.atlases):
        got = d.cpu().numpy().view(np.uint16)
        assert np.array_equal(got, h), "UNORM16 codes differ between the device and host generators"
    assert len(np.unique(host.atlases[0])) > 1000


@pytest.mark.parametrize("mode", ["random", "tiles", "all", "samples", "samples_fp16"])
def test_gather_modes(case, gi, mode):
    torch = require_gpu()
    h = case.host
    upd, vis, ent = case.oracle_cull()
    assert vis[0] > 0
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    color = torch.zeros((W.height, W.width, 3), dtype=torch.float32, device="cuda")
    res = torch.zeros((W.height, W.width, 4), dtype=torch.float32, device="cuda")
    tiles = None
    if mode == "tiles":
        _, tiles = orc.cull_tiles(h.frame, case.hg, W.tile_size, W.max_vpls_per_tile, upd, vis)
    if mode in ("random", "tiles"):
        gi.gather(case.dev.gbuffer(), d(upd), d(vis), d(ent), case.dev.cubes(),
                  None if tiles is None else d(tiles), color, res)
        want = orc.gather(h.frame, case.hg, upd, int(vis[0]), ent, case.hcubes, tiles=tiles, tile_size=W.tile_size,
                          K=W.max_vpls_per_tile, mode=orc.GATHER_TILES if tiles is not None else orc.GATHER_RANDOM)
    else:
        fp16 = mode == "samples_fp16"
        gmode = _abi.GATHER_ALL if mode == "all" else _abi.GATHER_SAMPLES
        gi.gather_ex(case.dev.gbuffer(), d(upd), d(vis), d(ent), case.dev.cubes(), None, color, res, mode=gmode,
                     samples=8, rejection_limit=-1,
                     precision=_abi.PRECISION_FP16 if fp16 else _abi.PRECISION_FP32)
        want = orc.gather(h.frame, case.hg, upd, int(vis[0]), ent, case.hcubes,
                          mode=orc.GATHER_CONVERGED if mode == "all" else orc.GATHER_SAMPLES, precision=0,
                          samples=8, rejection_limit=-1)
    torch.cuda.synchronize()
    got_r = res.cpu().numpy()
    e_c = orc.rel_l2(color.cpu().numpy(), want[0])
    e_r = orc.rel_l2(got_r[..., :3], want[1][..., :3])
    e_w = orc.rel_l2(got_r[..., 3], want[1][..., 3])
    print(f"[unorm16 gather {mode}] rel-L2 colour {e_c:.2e} radiance {e_r:.2e} weights {e_w:.2e}", flush=True)
    assert e_c <= REL_L2 and e_w <= REL_L2  # the weights follow the exact decisions in every mode
    bar = 1e-2 if mode == "samples_fp16" else REL_L2  # packed-FP16 shading's stated bar vs the FP32 restatement
    assert e_r <= bar
    assert np.abs(want[1][..., :3]).sum() > 0


def test_colouring_reads_unorm16_distance(case, gi):
    torch = require_gpu()
    n = W.num_vpls
    dif_d, cd_d = sc.diffuse_cubes(case.scene, case.dev)
    upd = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    vis = torch.zeros((1 + n,), dtype=torch.int32, device="cuda")
    ent = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    gi.cull_global(case.dev.vpls, n, upd, vis)
    gi.visible_atlas_entries(case.dev.entries, vis, n, ent)
    gi.color_vpls(vis, ent, cd_d, case.dev.cubes(), n, upd, color_multiplier=_abi.REFERENCE_VPL_COLOR_MULTIPLIER)
    torch.cuda.synchronize()
    dif_h = [a.cpu().numpy() for a in dif_d]
    cd_h = _abi.CubeAtlases()
    cd_h.resolution, cd_h.format, cd_h.layers_per_atlas, cd_h.num_atlases = W.cube_res, _abi.TEXEL_RGBA8, \
        case.host.layers, len(dif_h)
    for i, a in enumerate(dif_h):
        cd_h.atlas[i] = a.ctypes.data
    o_upd, o_vis, o_ent = case.oracle_cull()
    want = orc.color_vpls(case.host.frame, o_upd, o_ent, cd_h, case.hcubes, _abi.REFERENCE_VPL_COLOR_MULTIPLIER)
    c = int(o_vis[0])
    got = upd.cpu().numpy()[:c]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
