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


_F32, _I32, _BF16, _U8 = "float32", "int32", "bfloat16", "uint8"


def _dev(t, name: str, numel: int, dtypes=(_F32,), optional: bool = False) -> int:
    """Device pointer of `t` after checking what the C ABI cannot: a contiguous
    GPU tensor of an accepted dtype holding at least `numel` elements.  A wrong
    buffer here would otherwise become an out-of-bounds device access."""
    if t is None:
        if optional:
            return 0
        raise _abi.SgiArgumentError(f"{name}: a device tensor is required")
    if not getattr(t, "is_cuda", False):
        raise _abi.SgiArgumentError(f"{name}: expected a GPU tensor, got {getattr(t, 'device', type(t))}")
    dt = str(t.dtype).replace("torch.", "")
    if dt not in dtypes:
        raise _abi.SgiArgumentError(f"{name}: dtype {dt}, expected one of {dtypes}")
    if not t.is_contiguous():
        raise _abi.SgiArgumentError(f"{name}: tensor must be contiguous")
    if t.numel() < numel:
        raise _abi.SgiArgumentError(f"{name}: {t.numel()} elements, at least {numel} needed")
    return t.data_ptr()


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

    def _check_gbuf(self, gbuf) -> None:
        if (gbuf.width, gbuf.height) != (self.width, self.height):
            raise _abi.SgiArgumentError(f"G-buffer is {gbuf.width}x{gbuf.height}, the context's viewport "
                             f"{self.width}x{self.height}")

    def set_frame(self, frame: _abi.Frame) -> None:
        _abi.call(self.lib, "sgi_set_frame", self.ctx, ctypes.byref(frame))

    def set_csm(self, csm: _abi.Csm) -> None:
        _abi.call(self.lib, "sgi_set_csm", self.ctx, ctypes.byref(csm))

    def cull_global(self, vpls, n, out, visible, stream=None) -> None:
        cap = min(max(n, 0), _abi.MAX_TOTAL_VPLS_PER_FRAME)
        vf = _abi.VPL_FLOATS
        _abi.call(self.lib, "sgi_cull_global", self.ctx, _dev(vpls, "vpls", n * vf), n,
                  _dev(out, "out", cap * vf), _dev(visible, "visible", 1 + cap, (_I32,)), _stream(stream))

    def visible_atlas_entries(self, entries, visible, max_count, out, stream=None) -> None:
        _abi.call(self.lib, "sgi_visible_atlas_entries", self.ctx, _dev(entries, "entries", 0, (_I32,)),
                  _dev(visible, "visible", 1, (_I32,)), max_count, _dev(out, "out", 2 * max_count, (_I32,)),
                  _stream(stream))

    def taa(self, screen, prev_screen, velocity, out, stream=None) -> None:
        """taa.fs: the current frame blended with the neighbourhood-clamped history."""
        P = self.width * self.height
        _abi.call(self.lib, "sgi_taa", self.ctx, _dev(screen, "screen", 3 * P), _dev(prev_screen, "prev_screen", 3 * P),
                  _dev(velocity, "velocity", 2 * P, optional=True), _dev(out, "out", 3 * P), _stream(stream))

    def tonemap(self, lit, out, gamma=2.2, stream=None) -> None:
        """gammaTonemap.fs: ACES film + gamma of the lit buffer into RGBA8 texels."""
        P = self.width * self.height
        _abi.call(self.lib, "sgi_tonemap", self.ctx, _dev(lit, "lit", 3 * P), gamma,
                  _dev(out, "out", P * 4 // out.element_size() if out is not None else 0, (_I32, "uint32", _U8)),
                  _stream(stream))

    def render_vpl_cubes(self, positions, colors, vpls, updates, entries, distance: _abi.CubeAtlases,
                         diffuse: _abi.CubeAtlases | None = None, near=_abi.REFERENCE_VPL_NEAR_PLANE,
                         cull=_abi.CULLING_CCW, stream=None) -> None:
        """The VPL shadow pre-pass: the cubes of vpls[updates] from world-space triangles."""
        n = 0 if updates is None else updates.numel()
        count = 0 if positions is None else positions.numel() // 9
        if n > 0:  # the kernels index vpls and entries by these values
            lo, hi = int(updates.min()), int(updates.max())
            if lo < 0 or hi >= vpls.numel() // _abi.VPL_FLOATS or hi >= entries.numel() // 2:
                raise _abi.SgiArgumentError(f"updates must index vpls and entries (got {lo}..{hi})")
        tri = _abi.Triangles(count, _dev(positions, "positions", 9 * count, optional=count == 0),
                             _dev(colors, "colors", count, (_I32, "uint32"), optional=True), cull)
        _abi.call(self.lib, "sgi_render_vpl_cubes", self.ctx, ctypes.byref(tri),
                  _dev(vpls, "vpls", 0, optional=n == 0), _dev(updates, "updates", n, (_I32,), optional=n == 0), n,
                  _dev(entries, "entries", 0, (_I32,), optional=n == 0), near, ctypes.byref(distance),
                  ctypes.byref(diffuse) if diffuse is not None else None, _stream(stream))

    def render_csm(self, positions, csm: _abi.Csm, depth, slope=_abi.REFERENCE_CSM_SLOPE_FACTOR,
                   cull=_abi.CULLING_CCW, stream=None) -> None:
        """The sun cascades' depth render into depth [4][res][res] (csm's matrices and resolution)."""
        count = 0 if positions is None else positions.numel() // 9
        tri = _abi.Triangles(count, _dev(positions, "positions", 9 * count, optional=count == 0), 0, cull)
        r = csm.resolution
        _abi.call(self.lib, "sgi_render_csm", self.ctx, ctypes.byref(tri), ctypes.byref(csm), slope,
                  _dev(depth, "depth", 4 * r * r), _stream(stream))

    def taa_history(self, taa_out, prev_screen, stream=None) -> None:
        """The TAA output as the next frame's history (RGBA16F rounding, clamp, RGB8 store, k / 255)."""
        P = self.width * self.height
        _abi.call(self.lib, "sgi_taa_history", self.ctx, _dev(taa_out, "taa_out", 3 * P),
                  _dev(prev_screen, "prev_screen", 3 * P), _stream(stream))

    def fxaa(self, tonemapped, out, params: _abi.FxaaParams | None = None, stream=None) -> None:
        """fxaa_luminance.fs + fxaa_smoothing.fs of the tonemapped RGBA8 target."""
        P = self.width * self.height
        n = lambda t: P * 4 // t.element_size() if t is not None else 0  # noqa: E731
        _abi.call(self.lib, "sgi_fxaa", self.ctx, _dev(tonemapped, "tonemapped", n(tonemapped), (_I32, "uint32", _U8)),
                  ctypes.byref(params) if params is not None else None,
                  _dev(out, "out", n(out), (_I32, "uint32", _U8)), _stream(stream))

    def gbuffer_pack(self, planes: _abi.Gbuffer, out: _abi.GbufferPacked, stream=None) -> None:
        """The geometry pass's attachment writes: planes' FP32 normal / albedo / reflectivity /
        roughness / velocity into the packed attachments `out` names."""
        _abi.call(self.lib, "sgi_gbuffer_pack", self.ctx, ctypes.byref(planes), ctypes.byref(out), _stream(stream))

    def gbuffer_unpack(self, packed: _abi.GbufferPacked, normal, albedo, refl, rma, velocity=None,
                       stream=None) -> None:
        """The reference's packed G-buffer attachments into the FP32 planes the passes read."""
        P = self.width * self.height
        _abi.call(self.lib, "sgi_gbuffer_unpack", self.ctx, ctypes.byref(packed), _dev(normal, "normal", 3 * P),
                  _dev(albedo, "albedo", 4 * P), _dev(refl, "refl", 2 * P), _dev(rma, "rma", 3 * P),
                  _dev(velocity, "velocity", 2 * P, optional=True), _stream(stream))

    def color_vpls(self, visible, visible_entries, diffuse: _abi.CubeAtlases, distance: _abi.CubeAtlases,
                   max_count, vpls, color_multiplier=_abi.REFERENCE_VPL_COLOR_MULTIPLIER, stream=None) -> None:
        """vpl_light_color.cs: colour and specular position of the visible VPL records, in place."""
        n = min(max(max_count, 0), _abi.MAX_TOTAL_VPLS_PER_FRAME)  # the C ABI refuses counts outside 0..4480
        _abi.call(self.lib, "sgi_color_vpls", self.ctx, _dev(visible, "visible", 1, (_I32,)),
                  _dev(visible_entries, "visible_entries", 2 * n, (_I32,), optional=n == 0),
                  ctypes.byref(diffuse), ctypes.byref(distance), color_multiplier, max_count,
                  _dev(vpls, "vpls", _abi.VPL_FLOATS * n, optional=n == 0), _stream(stream))

    def cull_tiles(self, gbuf: _abi.Gbuffer, vpls, visible, tiles, remap=0, stage1=None, stream=None) -> None:
        self._check_gbuf(gbuf)
        _abi.call(self.lib, "sgi_cull_tiles", self.ctx, ctypes.byref(gbuf), _dev(vpls, "vpls", 0),
                  _dev(visible, "visible", 1, (_I32,)), remap, _dev(stage1, "stage1", 8 * self.num_tiles, optional=True),
                  _dev(tiles, "tiles", self.num_tiles * (1 + self.K), (_I32,)), _stream(stream))

    def gather(self, gbuf: _abi.Gbuffer, vpls, visible, entries, cubes: _abi.CubeAtlases, tiles, color, reservoir,
               precision=_abi.PRECISION_FP32, stream=None) -> None:
        self._check_gbuf(gbuf)
        P = self.width * self.height
        _abi.call(self.lib, "sgi_gather", self.ctx, ctypes.byref(gbuf), _dev(vpls, "vpls", 0),
                  _dev(visible, "visible", 1, (_I32,)), _dev(entries, "entries", 0, (_I32,)), ctypes.byref(cubes),
                  _dev(tiles, "tiles", self.num_tiles * (1 + self.K), (_I32,), optional=True), precision,
                  _dev(color, "color", 3 * P), _dev(reservoir, "reservoir", 4 * P), _stream(stream))

    def gather_ex(self, gbuf: _abi.Gbuffer, vpls, visible, entries, cubes: _abi.CubeAtlases, tiles, color, reservoir,
                  mode=_abi.GATHER_ALL, samples=8, accumulate_fp64=False, precision=_abi.PRECISION_FP32,
                  rejection_limit=5, stream=None) -> None:
        opts = _abi.GatherOpts(mode, samples, precision, 1 if accumulate_fp64 else 0, rejection_limit)
        self._check_gbuf(gbuf)
        P = self.width * self.height
        _abi.call(self.lib, "sgi_gather_ex", self.ctx, ctypes.byref(gbuf), _dev(vpls, "vpls", 0),
                  _dev(visible, "visible", 1, (_I32,)), _dev(entries, "entries", 0, (_I32,)), ctypes.byref(cubes),
                  _dev(tiles, "tiles", self.num_tiles * (1 + self.K), (_I32,), optional=True), ctypes.byref(opts),
                  _dev(color, "color", 3 * P), _dev(reservoir, "reservoir", 4 * P), _stream(stream))

    def direct_light(self, gbuf: _abi.Gbuffer, lit, stream=None) -> None:
        self._check_gbuf(gbuf)
        _abi.call(self.lib, "sgi_direct_light", self.ctx, ctypes.byref(gbuf),
                  _dev(lit, "lit", 3 * self.width * self.height), _stream(stream))

    def set_screen_event(self, event) -> None:
        """sgi_set_screen_event: a torch.cuda.Event (recorded at least once), a raw hipEvent_t, or None.
        The denoisers then wait for it just before their temporal pass, the only one that reads screen."""
        handle = getattr(event, "cuda_event", event) if event is not None else None
        if event is not None and not handle:
            raise ValueError("set_screen_event: the event has not been recorded yet")
        self._screen_event = event  # kept alive while it is set
        _abi.call(self.lib, "sgi_set_screen_event", self.ctx, ctypes.c_void_p(handle) if handle else None)

    def denoise(self, cur: _abi.Gbuffer, prev: _abi.Gbuffer | None, visible, screen, color, reservoir, prev_gi,
                prev_history, combined, gi_out, reservoir_out, history_out, stream=None) -> None:
        self._check_gbuf(cur)
        if prev is not None:
            self._check_gbuf(prev)
        P = self.width * self.height
        _abi.call(self.lib, "sgi_denoise", self.ctx, ctypes.byref(cur), ctypes.byref(prev) if prev else None,
                  _dev(visible, "visible", 1, (_I32,)), _dev(screen, "screen", 3 * P), _dev(color, "color", 3 * P),
                  _dev(reservoir, "reservoir", 4 * P), _dev(prev_gi, "prev_gi", 3 * P, optional=True),
                  _dev(prev_history, "prev_history", P, optional=True), _dev(combined, "combined", 3 * P),
                  _dev(gi_out, "gi_out", 3 * P), _dev(reservoir_out, "reservoir_out", 4 * P),
                  _dev(history_out, "history_out", P), _stream(stream))

    def kpcn_features(self, gbuf: _abi.Gbuffer, reservoir, features, radiance, stream=None) -> None:
        self._check_gbuf(gbuf)
        P = self.width * self.height
        _abi.call(self.lib, "sgi_kpcn_features", self.ctx, ctypes.byref(gbuf), _dev(reservoir, "reservoir", 4 * P),
                  _dev(features, "features", 32 * P, (_BF16,)), _dev(radiance, "radiance", 3 * P), _stream(stream))

    def denoise_learned(self, cur: _abi.Gbuffer, prev: _abi.Gbuffer | None, visible, screen, color, reservoir,
                        prev_gi, prev_history, net: _abi.KpcnNet, threshold: float, workspace, combined, gi_out,
                        reservoir_out, history_out, fallback_tiles=None, stream=None) -> None:
        self._check_gbuf(cur)
        if prev is not None:
            self._check_gbuf(prev)
        P = self.width * self.height
        _abi.call(self.lib, "sgi_denoise_learned", self.ctx, ctypes.byref(cur), ctypes.byref(prev) if prev else None,
                  _dev(visible, "visible", 1, (_I32,)), _dev(screen, "screen", 3 * P), _dev(color, "color", 3 * P),
                  _dev(reservoir, "reservoir", 4 * P), _dev(prev_gi, "prev_gi", 3 * P, optional=True),
                  _dev(prev_history, "prev_history", P, optional=True), ctypes.byref(net), ctypes.c_float(threshold),
                  _dev(workspace, "workspace", int(self.lib.sgi_kpcn_workspace_bytes(self.height, self.width)), (_U8,)),
                  workspace.numel() * workspace.element_size(), _dev(combined, "combined", 3 * P),
                  _dev(gi_out, "gi_out", 3 * P), _dev(reservoir_out, "reservoir_out", 4 * P),
                  _dev(history_out, "history_out", P),
                  _dev(fallback_tiles, "fallback_tiles", 1 + (-(-self.width // 16)) * (-(-self.height // 16)), (_I32,),
                       optional=True), _stream(stream))

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

    def set_learned(self, samples: int, net: _abi.KpcnNet | None = None, threshold: float = float("inf"),
                    rejection_limit: int = -1) -> None:
        """The batch's learned form (sgi_views_set_learned): per view the sampled gather at `samples`
        draws per pixel and the learned denoiser; samples = 0 restores the tile-list form.  The
        tensors behind `net` must outlive the batch."""
        _abi.call(self.lib, "sgi_views_set_learned", self.handle, samples, rejection_limit,
                  ctypes.byref(net) if net is not None else None, threshold)

    def run(self, stream=None) -> None:
        _abi.call(self.lib, "sgi_views_run", self.handle, _stream(stream))

    def capture(self, stream) -> None:
        _abi.call(self.lib, "sgi_views_capture", self.handle, stream)

    def replay(self, stream=None) -> None:
        _abi.call(self.lib, "sgi_views_replay", self.handle, _stream(stream))
