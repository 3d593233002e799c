"""sgi_render_vpl_cubes — the VPL shadow pre-pass (StratusRendererBackend.cpp:
1426-1575, shadowVpl.fs:17-37, shadow.vs:22-33) — on the GPU, bit for bit the
oracle's brute-force restatement:

- the Cornell box's 64 VPL cubes (UNORM16, the reference's format, with the
  RGBA8 diffuse cubes) from its 84 triangles;
- two city-block lights (24012 triangles) in every distance format;
- lights split over several internal batches (the last lights of a call
  larger than one batch), checked against the oracle;
- an entry outside the atlases writes nothing; bad arguments are refused;
- a whole frame rendered from the pass's own cubes matches the oracle chain
  on the same cubes (the producer feeding the gather, as in the reference).
"""
import dataclasses

import numpy as np
import pytest

import oracle as orc
from gpu_util import require_gpu
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc
from stratusgfx_amd.gi import FrameBuffers, GiPass

pytestmark = pytest.mark.gpu

NEAR = _abi.REFERENCE_VPL_NEAR_PLANE


def _atlas_pair(torch, layers, res, fmt, num_atlases=1, host=False):
    dt = {_abi.TEXEL_F32: (torch.float32, np.float32), _abi.TEXEL_F16: (torch.int16, np.uint16),
          _abi.TEXEL_UNORM16: (torch.int16, np.uint16)}[fmt]
    if host:
        z = [np.full((layers, 6, res, res), 3, dt[1]) for _ in range(num_atlases)]
        c = [np.full((layers, 6, res, res), 0x0BADF00D, np.uint32) for _ in range(num_atlases)]
        ptr = lambda a: a.ctypes.data  # noqa: E731
    else:
        z = [torch.full((layers, 6, res, res), 3, dtype=dt[0], device="cuda") for _ in range(num_atlases)]
        c = [torch.full((layers, 6, res, res), 0x0BADF00D, dtype=torch.int32, device="cuda")
             for _ in range(num_atlases)]
        ptr = lambda a: a.data_ptr()  # noqa: E731
    cz, cc = _abi.CubeAtlases(), _abi.CubeAtlases()
    cz.resolution, cz.format, cz.layers_per_atlas, cz.num_atlases = res, fmt, layers, num_atlases
    cc.resolution, cc.format, cc.layers_per_atlas, cc.num_atlases = res, _abi.TEXEL_RGBA8, layers, num_atlases
    for i in range(num_atlases):
        cz.atlas[i], cc.atlas[i] = ptr(z[i]), ptr(c[i])
    return z, c, cz, cc


def _run(w, fmt, updates, res, with_colour=True):
    torch = require_gpu()
    scene = sc.Scene(w)
    pos, col = scene.triangles()
    vpls, entries = scene.vpls()
    layers = sc.layers_per_atlas(w.num_vpls)
    na = sc.num_atlases(w.num_vpls)
    gi = GiPass(64, 64, 16, 12)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    z, c, cz, cc = _atlas_pair(torch, layers, res, fmt, na)
    up = np.asarray(updates, np.int32)
    gi.render_vpl_cubes(d(pos), d(col.view(np.int32)), d(vpls), d(up), d(entries), cz,
                        cc if with_colour else None, near=NEAR)
    torch.cuda.synchronize()
    hz, hc, hcz, hcc = _atlas_pair(torch, layers, res, fmt, na, host=True)
    orc.render_vpl_cubes(pos, col, vpls, up, entries, NEAR, hcz, hcc if with_colour else None)
    for a in range(na):
        gz = z[a].cpu().numpy().view(hz[a].dtype)
        assert np.array_equal(gz.view(np.uint16 if gz.itemsize == 2 else np.uint32),
                              hz[a].view(np.uint16 if gz.itemsize == 2 else np.uint32)), f"distance atlas {a}"
        if with_colour:
            assert np.array_equal(c[a].cpu().numpy().view(np.uint32), hc[a]), f"diffuse atlas {a}"
    return hz, hc, entries


def test_cornell_all_cubes_unorm16():
    w = sc.CORNELL_640x360
    hz, hc, entries = _run(w, _abi.TEXEL_UNORM16, np.arange(w.num_vpls), 32)
    written = hz[0][: w.num_vpls]
    print(f"[vpl cubes] cornell: 64 cubes bit-exact; {np.mean(written < 65535):.3f} of texels covered", flush=True)
    assert np.mean(written < 65535) > 0.5


@pytest.mark.parametrize("fmt", [_abi.TEXEL_UNORM16, _abi.TEXEL_F16, _abi.TEXEL_F32])
def test_city_two_lights(fmt):
    w = dataclasses.replace(sc.CITY_1080P, num_vpls=64)
    hz, _, _ = _run(w, fmt, [5, 40], 32)
    print(f"[vpl cubes] city fmt {fmt}: lights 5 and 40 bit-exact", flush=True)


def test_city_many_lights_in_batches():
    """24012 triangles: a batch holds (64 Mi ints) / (6 x 24012 + 6) = 465 lights, so 500 updates run
    in two batches; lights of both batches are checked against the oracle."""
    torch = require_gpu()
    w = dataclasses.replace(sc.CITY_1080P, num_vpls=1024)
    scene = sc.Scene(w)
    pos, col = scene.triangles()
    assert pos.shape[0] == 24012 and (64 << 20) // (6 * 24012 + 6) == 465
    vpls, entries = scene.vpls()
    layers, na, res = sc.layers_per_atlas(w.num_vpls), sc.num_atlases(w.num_vpls), 16
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    z, c, cz, cc = _atlas_pair(torch, layers, res, _abi.TEXEL_UNORM16, na)
    gi = GiPass(64, 64, 16, 12)
    up = np.arange(500, dtype=np.int32)
    gi.render_vpl_cubes(d(pos), d(col.view(np.int32)), d(vpls), d(up), d(entries), cz, cc)
    torch.cuda.synchronize()
    for pick in ([0, 1, 464], [465, 499]):
        hz, hc, hcz, hcc = _atlas_pair(torch, layers, res, _abi.TEXEL_UNORM16, na, host=True)
        orc.render_vpl_cubes(pos, col, vpls, np.array(pick, np.int32), entries, NEAR, hcz, hcc)
        for j in pick:
            a, l = entries[j]
            assert np.array_equal(z[a].cpu().numpy().view(np.uint16)[l], hz[a][l]), j
            assert np.array_equal(c[a].cpu().numpy().view(np.uint32)[l], hc[a][l]), j
    # lights not in the call keep their sentinels
    a, l = entries[700]
    assert np.all(z[a].cpu().numpy()[l] == 3)


def test_out_of_range_entry_and_errors():
    torch = require_gpu()
    w = sc.CORNELL_640x360
    scene = sc.Scene(w)
    pos, col = scene.triangles()
    vpls, entries = scene.vpls()
    entries = entries.copy()
    entries[3] = (5, 0)  # no such atlas
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    z, c, cz, cc = _atlas_pair(torch, 64, 16, _abi.TEXEL_F32)
    gi = GiPass(64, 64, 16, 12)
    gi.render_vpl_cubes(d(pos), d(col.view(np.int32)), d(vpls), d(np.array([3], np.int32)), d(entries), cz, cc)
    torch.cuda.synchronize()
    assert np.all(z[0].cpu().numpy() == 3) and np.all(c[0].cpu().numpy() == 0x0BADF00D)
    up = d(np.array([0], np.int32))
    for kw in (dict(near=0.0), dict(near=float("nan"))):
        with pytest.raises(_abi.SgiError):
            gi.render_vpl_cubes(d(pos), d(col.view(np.int32)), d(vpls), up, d(entries), cz, cc, **kw)
    bad = _abi.CubeAtlases()
    bad.resolution, bad.format, bad.layers_per_atlas, bad.num_atlases = 8, _abi.TEXEL_RGBA8, 64, 1
    bad.atlas[0] = c[0].data_ptr()
    with pytest.raises(_abi.SgiError):  # diffuse geometry differs from the distance atlases
        gi.render_vpl_cubes(d(pos), d(col.view(np.int32)), d(vpls), up, d(entries), cz, bad)
    with pytest.raises(_abi.SgiError):  # diffuse output without colours
        gi.render_vpl_cubes(d(pos), None, d(vpls), up, d(entries), cz, cc)


def test_frame_from_the_pass_cubes():
    """The Cornell frame with its cubes from sgi_render_vpl_cubes (UNORM16): the GPU chain equals the
    oracle chain on the same cubes (lists bit-exact, lit output within 1e-4)."""
    torch = require_gpu()
    w = dataclasses.replace(sc.CORNELL_640x360, cube_format=_abi.TEXEL_UNORM16)
    scene = sc.Scene(w)
    dev = sc.generate_device(scene)
    pos, col = scene.triangles()
    vpls, entries = scene.vpls()
    gi = GiPass(w.width, w.height, w.tile_size, w.max_vpls_per_tile)
    gi.set_frame(dev.frame)
    gi.set_csm(dev.csm_struct())
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    cubes = dev.cubes()
    gi.render_vpl_cubes(d(pos), d(col.view(np.int32)), d(vpls), d(np.arange(w.num_vpls, dtype=np.int32)),
                        d(entries), cubes)
    fb = FrameBuffers.allocate(w.width, w.height, w.num_vpls, gi.num_tiles, w.max_vpls_per_tile)
    gi.render_frame(fb.frame_io(dev.gbuffer(), None, dev.vpls, w.num_vpls, dev.entries, cubes, True))
    torch.cuda.synchronize()
    h = sc.to_host(dev)
    hg, hc, hcsm = h.gbuffer(), h.cubes(), h.csm_struct()
    upd, vis = orc.cull_global(h.frame, hcsm, h.vpls)
    _, tiles = orc.cull_tiles(h.frame, hg, w.tile_size, w.max_vpls_per_tile, upd, vis)
    ent = h.entries[vis[1:]]
    colr, res = orc.gather(h.frame, hg, upd, int(vis[0]), ent, hc, tiles=tiles, tile_size=w.tile_size,
                           K=w.max_vpls_per_tile, mode=orc.GATHER_TILES)
    lit = orc.direct_light(h.frame, hcsm, hg)
    comb, _, _, _ = orc.denoise(h.frame, hg, None, int(vis[0]), lit, colr, res)
    e = orc.rel_l2(fb.combined.cpu().numpy(), comb)
    print(f"[vpl cubes] Cornell frame from the pass's cubes: lit rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-4
    assert np.array_equal(fb.tiles.cpu().numpy()[:, 0], tiles[:, 0])
