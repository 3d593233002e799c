"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU restatement (FP32 and FP64) of
StratusGFX's deferred VPL GI shaders.  Only tests/, __graft_entry__.smoke()
and bench.py's CPU baseline may import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_void_p

import numpy as np

from stratusgfx_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

GATHER_RANDOM, GATHER_TILES, GATHER_CONVERGED, GATHER_SAMPLES = 0, 1, 2, 3

_lib = None


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


_SIGNATURES = {
    "orc_hash1": (c_uint32, [c_uint32]),
    "orc_hash3": (c_uint32, [c_uint32, c_uint32, c_uint32]),
    "orc_random3": (c_float, [c_float, c_float, c_float]),
    "orc_ndf_ggx": (c_double, [c_double, c_double]),
    "orc_ndf_ggx_f": (c_float, [c_float, c_float]),
    "orc_vis_smith": (c_double, [c_double, c_double, c_double]),
    "orc_fresnel": (c_double, [c_double, c_double]),
    "orc_fd_burley": (c_double, [c_double, c_double, c_double, c_double]),
    "orc_vpl_atten": (c_double, [c_double, c_double]),
    "orc_brdf_burley": (None, [c_void_p] * 4 + [c_double] * 4 + [c_int, c_void_p]),
    "orc_vpl_lighting": (None, [c_void_p] * 7 + [c_double] * 5 + [c_int, c_void_p]),
    "orc_cube_fetch": (c_double, [POINTER(_abi.CubeAtlases), c_int, c_int, c_double, c_double, c_double]),
    "orc_shadow1": (c_float, [POINTER(_abi.CubeAtlases), c_int, c_int, c_float, c_void_p, c_void_p,
                              c_float, c_float]),
    "orc_shadow_compare": (c_double, [c_void_p, c_int, c_double, c_double, c_int, c_double]),
    "orc_infinite_shadow": (c_float, [POINTER(_abi.Csm), c_void_p, c_void_p, c_void_p, c_int]),
    "orc_world_pos": (None, [c_float, c_float, c_float, c_void_p, c_void_p]),
    "orc_taa": (None, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "orc_tonemap": (None, [c_void_p, ctypes.c_long, c_float, c_void_p, c_void_p]),
    "orc_gbuffer_unpack": (None, [POINTER(_abi.GbufferPacked)] + [c_void_p] * 5),
    "orc_gbuffer_pack": (None, [POINTER(_abi.Gbuffer)] + [c_void_p] * 5),
    "orc_taa_history": (None, [c_void_p, ctypes.c_long, c_void_p]),
    "orc_fxaa_luminance": (None, [c_void_p, ctypes.c_long, c_void_p]),
    "orc_fxaa_smoothing": (None, [c_void_p, c_int, c_int, c_float, c_float, c_float, c_void_p]),
    "orc_half_rn": (ctypes.c_uint16, [c_float]),
    "orc_render_csm": (None, [c_void_p, c_int, c_int, c_float, POINTER(_abi.Csm), c_void_p]),
    "orc_render_vpl_cubes": (None, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_float,
                                    POINTER(_abi.CubeAtlases), POINTER(_abi.CubeAtlases)]),
    "orc_color_vpls": (None, [POINTER(_abi.Frame), c_int, c_void_p, POINTER(_abi.CubeAtlases),
                              POINTER(_abi.CubeAtlases), c_float, c_void_p]),
    "orc_cull_global": (c_int, [POINTER(_abi.Frame), POINTER(_abi.Csm), c_void_p, c_int, c_void_p,
                                c_void_p]),
    "orc_cull_tiles": (None, [POINTER(_abi.Frame), POINTER(_abi.Gbuffer), c_int, c_int, c_void_p, c_void_p,
                              c_int, c_int, c_int, c_void_p, c_void_p]),
    "orc_gather": (None, [POINTER(_abi.Frame), POINTER(_abi.Gbuffer), c_void_p, c_int, c_void_p,
                          POINTER(_abi.CubeAtlases), c_void_p, c_int, c_int, c_int, c_int, c_int, c_int,
                          c_int, c_int, c_void_p, c_void_p]),
    "orc_direct_light": (None, [POINTER(_abi.Frame), POINTER(_abi.Csm), POINTER(_abi.Gbuffer), c_int, c_int,
                                c_void_p]),
    "orc_denoise": (None, [POINTER(_abi.Frame), POINTER(_abi.Gbuffer), POINTER(_abi.Gbuffer), c_int]
                    + [c_void_p] * 9),
    "orc_set_threads": (None, [c_int]),
    "orc_get_threads": (c_int, []),
}


def load(path: str) -> ctypes.CDLL:
    """Loads a build of the oracle at `path` with the exported signatures set
    (the tests also load deliberately mutated builds to show that a pin fails)."""
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = load(LIB)
    return _lib


def set_threads(n: int) -> int:
    """Worker threads of the row-parallel passes (cull_tiles, gather, direct_light, denoise);
    n <= 0: every hardware thread.  Returns the count in effect.  Results do not depend on it."""
    lib().orc_set_threads(int(n))
    return int(lib().orc_get_threads())


def _p(a) -> int:
    return 0 if a is None else a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def cull_global(frame: _abi.Frame, csm: _abi.Csm, vpls: np.ndarray):
    """viscull_vpls.cs: returns (vplUpdatedData [count,16], visibleIndices [1+count])."""
    vpls = _f32(vpls)
    n = vpls.shape[0]
    cap = min(n, _abi.MAX_TOTAL_VPLS_PER_FRAME)
    out = np.zeros((max(cap, 1), 16), np.float32)
    idx = np.zeros(1 + max(cap, 1), np.int32)
    count = lib().orc_cull_global(ctypes.byref(frame), ctypes.byref(csm), _p(vpls), n, _p(out), _p(idx))
    return out[:count], idx[:1 + count]


def taa(screen, prev, velocity=None):
    """taa.fs: the anti-aliased frame [H][W][3]."""
    screen, prev = _f32(screen), _f32(prev)
    H, W = screen.shape[0], screen.shape[1]
    out = np.zeros((H, W, 3), np.float32)
    vel = None if velocity is None else _f32(velocity)
    lib().orc_taa(_p(screen), _p(prev), _p(vel), W, H, _p(out))
    return out


def tonemap(lit, gamma=2.2):
    """gammaTonemap.fs: returns (RGBA8 texels [H][W] uint32, the values before the store [H][W][3])."""
    lit = _f32(lit)
    H, W = lit.shape[0], lit.shape[1]
    out = np.zeros((H, W), np.uint32)
    film = np.zeros((H, W, 3), np.float32)
    lib().orc_tonemap(_p(lit), H * W, gamma, _p(out), _p(film))
    return out, film


def render_vpl_cubes(positions, colors, vpls, updates, entries, near, dist, diffuse=None, cull=_abi.CULLING_CCW):
    """The VPL shadow pre-pass into host atlases `dist` / `diffuse` (CubeAtlases over numpy arrays)."""
    pos = np.ascontiguousarray(positions, np.float32)
    col = None if colors is None else np.ascontiguousarray(colors, np.uint32)
    vp = _f32(vpls)
    up = np.ascontiguousarray(updates, np.int32)
    en = np.ascontiguousarray(entries, np.int32)
    lib().orc_render_vpl_cubes(_p(pos), _p(col), pos.shape[0], cull, _p(vp), _p(up), up.shape[0], _p(en), near,
                               ctypes.byref(dist), ctypes.byref(diffuse) if diffuse is not None else None)


def render_csm(positions, csm, slope=2.0, cull=_abi.CULLING_CCW):
    """The cascades' depth render: depth [4][res][res] from world-space triangles and csm's matrices."""
    pos = np.ascontiguousarray(positions, np.float32)
    res = csm.resolution
    depth = np.zeros((4, res, res), np.float32)
    lib().orc_render_csm(_p(pos), pos.shape[0], cull, slope, ctypes.byref(csm), _p(depth))
    return depth


def taa_history(taa):
    """The TAA output as the next frame reads it back from the RGB8 history (RGBA16F, clamp, UNORM8)."""
    taa = _f32(taa)
    prev = np.zeros_like(taa)
    lib().orc_taa_history(_p(taa), taa.size // 3, _p(prev))
    return prev


FXAA_CONTRAST_THRESHOLD, FXAA_RELATIVE_THRESHOLD, FXAA_SUBPIXEL_BLENDING = 0.0625, 0.125, 1.0


def fxaa(tonemapped, contrast=FXAA_CONTRAST_THRESHOLD, relative=FXAA_RELATIVE_THRESHOLD,
         subpixel=FXAA_SUBPIXEL_BLENDING):
    """fxaa_luminance.fs then fxaa_smoothing.fs on RGBA8 texels [H][W] (uint32): returns
    (the luminance target, the smoothed target), both RGBA8 texels."""
    t = np.ascontiguousarray(tonemapped, np.uint32)
    H, W = t.shape
    lum = np.zeros_like(t)
    out = np.zeros_like(t)
    lib().orc_fxaa_luminance(_p(t), H * W, _p(lum))
    lib().orc_fxaa_smoothing(_p(lum), W, H, contrast, relative, subpixel, _p(out))
    return lum, out


def gbuffer_unpack(packed: _abi.GbufferPacked, with_velocity=True):
    """The FP32 planes (normal, albedo, reflectivity, rma, velocity) the passes read from packed attachments."""
    H, W = packed.height, packed.width
    out = [np.zeros((H, W, c), np.float32) for c in (3, 4, 2, 3, 2)]
    lib().orc_gbuffer_unpack(ctypes.byref(packed), *[_p(a) for a in out[:4]], _p(out[4]) if with_velocity else None)
    return out


def gbuffer_pack(normal, albedo, refl, rma, velocity=None):
    """The geometry pass's attachment writes of FP32 planes [H][W][c]: (normals uint32, albedo uint32,
    baseReflectivity uint16, roughnessMetallicAmbient uint32, velocity uint16 [H][W][2] or None)."""
    planes = [np.ascontiguousarray(a, np.float32) for a in (normal, albedo, refl, rma)]
    vel = None if velocity is None else np.ascontiguousarray(velocity, np.float32)
    H, W = planes[0].shape[:2]
    g = _abi.Gbuffer(W, H, 0, *[a.ctypes.data for a in planes], _p(vel), 0, 0, None)
    out = [np.zeros((H, W), np.uint32), np.zeros((H, W), np.uint32), np.zeros((H, W), np.uint16),
           np.zeros((H, W), np.uint32), np.zeros((H, W, 2), np.uint16) if vel is not None else None]
    lib().orc_gbuffer_pack(ctypes.byref(g), *[_p(a) for a in out])
    return tuple(out)


def color_vpls(frame, vpls, entries, diffuse, distance, multiplier=_abi.REFERENCE_VPL_COLOR_MULTIPLIER):
    """vpl_light_color.cs over the compacted records vpls [count,16] with their per-slot atlas entries:
    returns a coloured copy."""
    out = _f32(vpls).copy()
    entries = np.ascontiguousarray(entries, np.int32)
    lib().orc_color_vpls(ctypes.byref(frame), out.shape[0], _p(entries), ctypes.byref(diffuse), ctypes.byref(distance),
                         multiplier, _p(out))
    return out


def cull_tiles(frame, gbuf, tile_size, K, vpls, visible, remap=0, t0=0, t1=None, tiles_x=None, tiles_y=None):
    """Stage 1 + stage 2 for tiles [t0, t1): returns (stage1 [T,8] f32, tiles [T,1+K] i32)."""
    W, H = gbuf.width, gbuf.height
    tx = (W + tile_size - 1) // tile_size
    ty = (H + tile_size - 1) // tile_size
    T = tx * ty
    t1 = T if t1 is None else t1
    vpls = _f32(vpls)
    visible = np.ascontiguousarray(visible, np.int32)
    stage1 = np.zeros((T, 8), np.float32)
    tiles = np.zeros((T, 1 + K), np.int32)
    lib().orc_cull_tiles(ctypes.byref(frame), ctypes.byref(gbuf), tile_size, K, _p(vpls), _p(visible), remap, t0, t1,
                         _p(stage1), _p(tiles))
    return stage1, tiles


def gather(frame, gbuf, vpls, num_visible, entries, cubes, tiles=None, tile_size=16, K=12, mode=GATHER_RANDOM,
           precision=0, y0=0, y1=None, samples=8, rejection_limit=5, L=None):
    """vpl_pbr_gi.fs over rows [y0, y1): returns (color [rows,W,3], reservoir [rows,W,4]).
    L: a library from load() in place of the oracle's own build."""
    W, H = gbuf.width, gbuf.height
    y1 = H if y1 is None else y1
    dt = np.float32 if precision == 0 else np.float64
    color = np.zeros((y1 - y0, W, 3), dt)
    res = np.zeros((y1 - y0, W, 4), dt)
    vpls = _f32(vpls)
    entries = np.ascontiguousarray(entries, np.int32)
    t = None if tiles is None else np.ascontiguousarray(tiles, np.int32)
    (L or lib()).orc_gather(ctypes.byref(frame), ctypes.byref(gbuf), _p(vpls), num_visible, _p(entries),
                            ctypes.byref(cubes),
                     _p(t), tile_size, K, mode, samples, rejection_limit, precision, y0, y1, _p(color), _p(res))
    return color, res


def direct_light(frame, csm, gbuf, y0=0, y1=None):
    W, H = gbuf.width, gbuf.height
    y1 = H if y1 is None else y1
    lit = np.zeros((y1 - y0, W, 3), np.float32)
    lib().orc_direct_light(ctypes.byref(frame), ctypes.byref(csm), ctypes.byref(gbuf), y0, y1, _p(lit))
    return lit


def denoise(frame, cur, prev, num_visible, screen, color, reservoir, prev_gi=None, prev_history=None):
    W, H = cur.width, cur.height
    combined = np.zeros((H, W, 3), np.float32)
    gi = np.zeros((H, W, 3), np.float32)
    res = np.zeros((H, W, 4), np.float32)
    hist = np.zeros((H, W), np.float32)
    screen, color, reservoir = _f32(screen), _f32(color), _f32(reservoir)
    pg = None if prev_gi is None else _f32(prev_gi)
    ph = None if prev_history is None else _f32(prev_history)
    lib().orc_denoise(ctypes.byref(frame), ctypes.byref(cur), ctypes.byref(prev) if prev is not None else None,
                      num_visible, _p(screen), _p(color), _p(reservoir), _p(pg), _p(ph), _p(combined), _p(gi),
                      _p(res), _p(hist))
    return combined, gi, res, hist


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
