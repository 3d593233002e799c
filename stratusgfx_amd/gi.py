"""Python handle on the GI pass's C ABI (include/sgi.h) over torch device memory.

torch supplies device memory and the stream; every computation runs in the
HIP kernels of libsgi.so.  A missing library raises — there is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _abi


def _ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def _stream(stream) -> int:
    if stream is not None:
        return stream
    import torch
    return torch.cuda.current_stream().cuda_stream


class GiPass:
    """One viewport's GI pass context (sgi_create / sgi_destroy)."""

    def __init__(self, width: int, height: int, tile_size: int = 16, max_vpls_per_tile: int = 12, lib=None):
        self.lib = lib if lib is not None else _abi.load()
        self.width, self.height = width, height
        self.tile_size, self.K = tile_size, max_vpls_per_tile
        cfg = _abi.Config(width, height, tile_size, max_vpls_per_tile)
        h = ctypes.c_void_p()
        _abi.call(self.lib, "sgi_create", ctypes.byref(cfg), ctypes.byref(h))
        self.ctx = h
        tx, ty = ctypes.c_int32(), ctypes.c_int32()
        _abi.call(self.lib, "sgi_tile_counts", self.ctx, ctypes.byref(tx), ctypes.byref(ty))
        self.tiles_x, self.tiles_y = tx.value, ty.value

    @property
    def num_tiles(self) -> int:
        return self.tiles_x * self.tiles_y

    def close(self) -> None:
        if self.ctx:
            self.lib.sgi_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_frame(self, frame: _abi.Frame) -> None:
        _abi.call(self.lib, "sgi_set_frame", self.ctx, ctypes.byref(frame))

    def set_csm(self, csm: _abi.Csm) -> None:
        _abi.call(self.lib, "sgi_set_csm", self.ctx, ctypes.byref(csm))

    def cull_global(self, vpls, n, out, visible, stream=None) -> None:
        _abi.call(self.lib, "sgi_cull_global", self.ctx, _ptr(vpls), n, _ptr(out), _ptr(visible), _stream(stream))

    def visible_atlas_entries(self, entries, visible, max_count, out, stream=None) -> None:
        _abi.call(self.lib, "sgi_visible_atlas_entries", self.ctx, _ptr(entries), _ptr(visible), max_count, _ptr(out),
                  _stream(stream))

    def taa(self, screen, prev_screen, velocity, out, stream=None) -> None:
        """taa.fs: the current frame blended with the neighbourhood-clamped history."""
        _abi.call(self.lib, "sgi_taa", self.ctx, _ptr(screen), _ptr(prev_screen), _ptr(velocity), _ptr(out),
                  _stream(stream))

    def tonemap(self, lit, out, gamma=2.2, stream=None) -> None:
        """gammaTonemap.fs: ACES film + gamma of the lit buffer into RGBA8 texels."""
        _abi.call(self.lib, "sgi_tonemap", self.ctx, _ptr(lit), gamma, _ptr(out), _stream(stream))

    def gbuffer_unpack(self, packed: _abi.GbufferPacked, normal, albedo, refl, rma, velocity=None,
                       stream=None) -> None:
        """The reference's packed G-buffer attachments into the FP32 planes the passes read."""
        _abi.call(self.lib, "sgi_gbuffer_unpack", self.ctx, ctypes.byref(packed), _ptr(normal), _ptr(albedo),
                  _ptr(refl), _ptr(rma), _ptr(velocity), _stream(stream))

    def color_vpls(self, visible, visible_entries, diffuse: _abi.CubeAtlases, distance: _abi.CubeAtlases,
                   max_count, vpls, color_multiplier=_abi.REFERENCE_VPL_COLOR_MULTIPLIER, stream=None) -> None:
        """vpl_light_color.cs: colour and specular position of the visible VPL records, in place."""
        _abi.call(self.lib, "sgi_color_vpls", self.ctx, _ptr(visible), _ptr(visible_entries), ctypes.byref(diffuse),
                  ctypes.byref(distance), color_multiplier, max_count, _ptr(vpls), _stream(stream))

    def cull_tiles(self, gbuf: _abi.Gbuffer, vpls, visible, tiles, remap=0, stage1=None, stream=None) -> None:
        _abi.call(self.lib, "sgi_cull_tiles", self.ctx, ctypes.byref(gbuf), _ptr(vpls), _ptr(visible), remap,
                  _ptr(stage1), _ptr(tiles), _stream(stream))

    def gather(self, gbuf: _abi.Gbuffer, vpls, visible, entries, cubes: _abi.CubeAtlases, tiles, color, reservoir,
               precision=_abi.PRECISION_FP32, stream=None) -> None:
        _abi.call(self.lib, "sgi_gather", self.ctx, ctypes.byref(gbuf), _ptr(vpls), _ptr(visible), _ptr(entries),
                  ctypes.byref(cubes), _ptr(tiles), precision, _ptr(color), _ptr(reservoir), _stream(stream))

    def gather_ex(self, gbuf: _abi.Gbuffer, vpls, visible, entries, cubes: _abi.CubeAtlases, tiles, color, reservoir,
                  mode=_abi.GATHER_ALL, samples=8, accumulate_fp64=False, precision=_abi.PRECISION_FP32,
                  rejection_limit=5, stream=None) -> None:
        opts = _abi.GatherOpts(mode, samples, precision, 1 if accumulate_fp64 else 0, rejection_limit)
        _abi.call(self.lib, "sgi_gather_ex", self.ctx, ctypes.byref(gbuf), _ptr(vpls), _ptr(visible), _ptr(entries),
                  ctypes.byref(cubes), _ptr(tiles), ctypes.byref(opts), _ptr(color), _ptr(reservoir),
                  _stream(stream))

    def direct_light(self, gbuf: _abi.Gbuffer, lit, stream=None) -> None:
        _abi.call(self.lib, "sgi_direct_light", self.ctx, ctypes.byref(gbuf), _ptr(lit), _stream(stream))

    def denoise(self, cur: _abi.Gbuffer, prev: _abi.Gbuffer | None, visible, screen, color, reservoir, prev_gi,
                prev_history, combined, gi_out, reservoir_out, history_out, stream=None) -> None:
        _abi.call(self.lib, "sgi_denoise", self.ctx, ctypes.byref(cur), ctypes.byref(prev) if prev else None,
                  _ptr(visible), _ptr(screen), _ptr(color), _ptr(reservoir), _ptr(prev_gi), _ptr(prev_history),
                  _ptr(combined), _ptr(gi_out), _ptr(reservoir_out), _ptr(history_out), _stream(stream))

    def kpcn_features(self, gbuf: _abi.Gbuffer, reservoir, features, radiance, stream=None) -> None:
        _abi.call(self.lib, "sgi_kpcn_features", self.ctx, ctypes.byref(gbuf), _ptr(reservoir), _ptr(features),
                  _ptr(radiance), _stream(stream))

    def denoise_learned(self, cur: _abi.Gbuffer, prev: _abi.Gbuffer | None, visible, screen, color, reservoir,
                        prev_gi, prev_history, net: _abi.KpcnNet, threshold: float, workspace, combined, gi_out,
                        reservoir_out, history_out, fallback_tiles=None, stream=None) -> None:
        _abi.call(self.lib, "sgi_denoise_learned", self.ctx, ctypes.byref(cur), ctypes.byref(prev) if prev else None,
                  _ptr(visible), _ptr(screen), _ptr(color), _ptr(reservoir), _ptr(prev_gi), _ptr(prev_history),
                  ctypes.byref(net), ctypes.c_float(threshold), _ptr(workspace),
                  workspace.numel() * workspace.element_size(), _ptr(combined), _ptr(gi_out), _ptr(reservoir_out),
                  _ptr(history_out), _ptr(fallback_tiles), _stream(stream))

    def kpcn_workspace(self, device="cuda"):
        import torch
        n = int(self.lib.sgi_kpcn_workspace_bytes(self.height, self.width))
        return torch.empty(n, dtype=torch.uint8, device=device)

    def render_frame(self, io: _abi.FrameIO, stream=None) -> None:
        _abi.call(self.lib, "sgi_render_frame", self.ctx, ctypes.byref(io), _stream(stream))


@dataclass
class FrameBuffers:
    """Device buffers of one GI frame (scratch and outputs of sgi_render_frame)."""
    vpl_updated: object
    visible: object
    visible_entries: object
    tiles: object
    direct: object
    color: object
    reservoir: object
    combined: object
    gi: object
    reservoir_out: object
    history: object

    @staticmethod
    def allocate(width: int, height: int, num_vpls: int, num_tiles: int, K: int, device="cuda") -> "FrameBuffers":
        import torch
        cap = max(1, min(num_vpls, _abi.MAX_TOTAL_VPLS_PER_FRAME))
        f = dict(dtype=torch.float32, device=device)
        i = dict(dtype=torch.int32, device=device)
        return FrameBuffers(
            vpl_updated=torch.zeros((cap, _abi.VPL_FLOATS), **f),
            visible=torch.zeros((1 + cap,), **i),
            visible_entries=torch.zeros((cap, 2), **i),
            tiles=torch.zeros((num_tiles, 1 + K), **i),
            direct=torch.zeros((height, width, 3), **f),
            color=torch.zeros((height, width, 3), **f),
            reservoir=torch.zeros((height, width, 4), **f),
            combined=torch.zeros((height, width, 3), **f),
            gi=torch.zeros((height, width, 3), **f),
            reservoir_out=torch.zeros((height, width, 4), **f),
            history=torch.zeros((height, width), **f),
        )

    def frame_io(self, gbuf_ref, prev_gbuf_ref, vpls, num_vpls, entries, cubes_ref, use_tiles: bool,
                 precision=_abi.PRECISION_FP32, prev_gi=None, prev_history=None, diffuse_ref=None,
                 color_multiplier=_abi.REFERENCE_VPL_COLOR_MULTIPLIER) -> _abi.FrameIO:
        io = _abi.FrameIO()
        io.diffuseCubes = ctypes.pointer(diffuse_ref) if diffuse_ref is not None else None
        io.vplColorMultiplier = color_multiplier
        io.gbuffer = ctypes.pointer(gbuf_ref)
        io.prevGbuffer = ctypes.pointer(prev_gbuf_ref) if prev_gbuf_ref is not None else None
        io.vplData = _ptr(vpls)
        io.numVpls = num_vpls
        io.shadowIndices = _ptr(entries)
        io.cubes = ctypes.pointer(cubes_ref)
        io.useTiles = 1 if use_tiles else 0
        io.precision = precision
        io.vplUpdatedData = _ptr(self.vpl_updated)
        io.visibleIndices = _ptr(self.visible)
        io.visibleShadowIndices = _ptr(self.visible_entries)
        io.tiles = _ptr(self.tiles)
        io.direct = _ptr(self.direct)
        io.color = _ptr(self.color)
        io.reservoir = _ptr(self.reservoir)
        io.prevGi = _ptr(prev_gi)
        io.prevHistory = _ptr(prev_history)
        io.combined = _ptr(self.combined)
        io.giOut = _ptr(self.gi)
        io.reservoirOut = _ptr(self.reservoir_out)
        io.historyOut = _ptr(self.history)
        return io


def view_io(inp, tiles, direct, color, reservoir, combined, gi, reservoir_out, history, prev_gi=None,
            prev_history=None) -> _abi.ViewIO:
    """One view of a batch (sgi_view_io) from a scene view's device inputs and its buffers."""
    io = _abi.ViewIO()
    io.gbuffer = inp.gbuffer()
    ctypes.memmove(ctypes.byref(io.frame), ctypes.byref(inp.frame), ctypes.sizeof(_abi.Frame))
    io.csm = inp.csm_struct()
    io.tiles, io.direct, io.color, io.reservoir = _ptr(tiles), _ptr(direct), _ptr(color), _ptr(reservoir)
    io.prevGi, io.prevHistory = _ptr(prev_gi), _ptr(prev_history)
    io.combined, io.giOut, io.reservoirOut, io.historyOut = _ptr(combined), _ptr(gi), _ptr(reservoir_out), _ptr(history)
    return io


class Views:
    """A batch of views rendered with one launch per stage (sgi_views_*)."""

    def __init__(self, max_views: int, max_pixels: int, tile_size: int = 16, max_vpls_per_tile: int = 64, lib=None):
        self.lib = lib if lib is not None else _abi.load()
        cfg = _abi.ViewsConfig(max_views, max_pixels, tile_size, max_vpls_per_tile)
        h = ctypes.c_void_p()
        _abi.call(self.lib, "sgi_views_create", ctypes.byref(cfg), ctypes.byref(h))
        self.handle = h
        self.tile_size, self.K = tile_size, max_vpls_per_tile

    def close(self) -> None:
        if self.handle:
            self.lib.sgi_views_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, ios: list, vpls, num_vpls: int, entries, cubes: _abi.CubeAtlases, stream=None) -> None:
        arr = (_abi.ViewIO * len(ios))(*ios)
        _abi.call(self.lib, "sgi_views_set", self.handle, len(ios), arr, _ptr(vpls), num_vpls, _ptr(entries),
                  ctypes.byref(cubes), _stream(stream))

    def run(self, stream=None) -> None:
        _abi.call(self.lib, "sgi_views_run", self.handle, _stream(stream))

    def capture(self, stream) -> None:
        _abi.call(self.lib, "sgi_views_capture", self.handle, stream)

    def replay(self, stream=None) -> None:
        _abi.call(self.lib, "sgi_views_replay", self.handle, _stream(stream))
