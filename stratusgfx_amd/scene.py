"""Seeded synthetic inputs for the GI pass (wraps include/sgi_scene.h).

Two scenes, as the benchmark configurations describe them:
  * the Cornell box (640x360, 64 VPLs, 6x64x64 FP32 cubes, 1024^2 cascades);
  * the procedural city block (1920x1080, ground + 2000 boxes, 1024 VPLs,
    6x256x256 FP16 cubes, up to 64 VPLs per 16x16 tile).
Inputs are produced either on the host (numpy; the host ray caster) or on the
device (torch tensors; the HIP ray caster) from the same per-texel code.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi

# VPL colour = albedo * irradiance * this; the reference scales its VPL colours
# by 50000 (vpl_light_color.cs:44), which with these scenes makes the indirect
# term ~40x the direct one, so the generator uses 1/10 of it.
VPL_COLOR_MULTIPLIER = 5000.0


@dataclass
class Workload:
    name: str
    kind: int
    seed: int
    width: int
    height: int
    num_vpls: int
    cube_res: int
    cube_format: int
    csm_res: int
    tile_size: int
    max_vpls_per_tile: int
    num_city_boxes: int = 2000
    city_half_extent: float = 100.0


# The benchmark configurations this pass is measured and tested on.
CORNELL_640x360 = Workload("cornell-640x360-64vpl", _abi.SCENE_CORNELL, 0, 640, 360, 64, 64, _abi.TEXEL_F32,
                           1024, 16, 12)
CITY_1080P = Workload("city-1920x1080-1024vpl", _abi.SCENE_CITY, 1, 1920, 1080, 1024, 256, _abi.TEXEL_F16,
                      1024, 16, 64)


def layers_per_atlas(num_vpls: int) -> int:
    """Cube layers per atlas as CreateShadowMap3DCache_ lays them out (<= 320 per atlas)."""
    return max(1, min(_abi.MAX_TOTAL_SHADOWS_PER_ATLAS, num_vpls))


def num_atlases(num_vpls: int) -> int:
    lpa = layers_per_atlas(num_vpls)
    return max(1, (num_vpls + lpa - 1) // lpa)


class Scene:
    """A generated scene; `lib` is libsgi.so (device) or libsgi_host.so (host only)."""

    def __init__(self, w: Workload, lib: ctypes.CDLL | None = None):
        self.w = w
        self.lib = lib if lib is not None else _abi.load()
        desc = _abi.SceneDesc(w.kind, w.seed, w.num_vpls, w.num_city_boxes, w.city_half_extent,
                              VPL_COLOR_MULTIPLIER)
        h = ctypes.c_void_p()
        _abi.call(self.lib, "sgi_scene_create", ctypes.byref(desc), ctypes.byref(h))
        self.handle = h
        self.view = _abi.View()
        _abi.call(self.lib, "sgi_scene_default_view", self.handle, w.width, w.height, ctypes.byref(self.view))
        self.view.csm_resolution = w.csm_res
        self.frame = _abi.Frame()
        self.csm = _abi.Csm()
        self.set_view(self.view)

    def set_view(self, view: _abi.View, prev: _abi.View | None = None) -> None:
        self.view = view
        _abi.call(self.lib, "sgi_scene_set_view", self.handle, ctypes.byref(view),
                  ctypes.byref(prev) if prev is not None else None, ctypes.byref(self.frame),
                  ctypes.byref(self.csm))

    def close(self) -> None:
        if self.handle:
            self.lib.sgi_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def boxes(self) -> np.ndarray:
        """The scene's boxes, [n][16] float32 (sgi_scene_boxes' records)."""
        n = int(self.lib.sgi_scene_num_boxes(self.handle))
        b = np.zeros((n, 16), np.float32)
        _abi.call(self.lib, "sgi_scene_boxes", self.handle, b.ctypes.data)
        return b

    def triangles(self) -> tuple[np.ndarray, np.ndarray]:
        """The boxes as a world-space triangle list (12 per box, in box order) with RGBA8 base colours:
        the static meshes the VPL cube pass (sgi_render_vpl_cubes) draws."""
        return box_triangles(self.boxes())

    def vpls(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.w.num_vpls
        v = np.zeros((n, _abi.VPL_FLOATS), np.float32)
        e = np.zeros((n, 2), np.int32)
        _abi.call(self.lib, "sgi_scene_vpls", self.handle, layers_per_atlas(n), v.ctypes.data, e.ctypes.data)
        return v, e


@dataclass
class Inputs:
    """One frame's pass inputs; arrays are numpy (host) or torch tensors (device)."""
    width: int
    height: int
    frame: _abi.Frame
    csm: _abi.Csm
    depth: object
    normal: object
    albedo: object
    refl: object
    rma: object
    velocity: object
    ids: object
    vpls: object
    entries: object
    atlases: list = field(default_factory=list)
    cube_res: int = 0
    cube_format: int = 0
    layers: int = 0
    csm_depth: object = None

    def ptr(self, a) -> int:
        if a is None:
            return 0
        return a.ctypes.data if isinstance(a, np.ndarray) else a.data_ptr()

    def gbuffer(self, ssao=None) -> _abi.Gbuffer:
        p = self.ptr
        return _abi.Gbuffer(self.width, self.height, p(self.depth), p(self.normal), p(self.albedo),
                            p(self.refl), p(self.rma), p(self.velocity), p(self.ids), p(ssao), None)

    def gbuffer_packed(self, packed: "PackedAttachments", ssao=None) -> _abi.Gbuffer:
        """This view's G-buffer read from packed attachments (the FP32 planes of the normals,
        albedo, reflectivity, roughness and velocity left out): depth, ids and SSAO stay FP32."""
        p = self.ptr
        return _abi.Gbuffer(self.width, self.height, p(self.depth), None, None, None, None, None, p(self.ids),
                            p(ssao), ctypes.addressof(packed.struct))

    def cubes(self) -> _abi.CubeAtlases:
        c = _abi.CubeAtlases()
        c.resolution = self.cube_res
        c.format = self.cube_format
        c.layers_per_atlas = self.layers
        c.num_atlases = len(self.atlases)
        for i, a in enumerate(self.atlases):
            c.atlas[i] = self.ptr(a)
        return c

    def csm_struct(self) -> _abi.Csm:
        c = _abi.Csm()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self.csm), ctypes.sizeof(_abi.Csm))
        c.depth = self.ptr(self.csm_depth)
        return c


# corners of a box's faces (local signs), two triangles per face
_FACES = [((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)), ((-1, -1, 1), (-1, 1, 1), (-1, 1, -1), (-1, -1, -1)),
          ((-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)), ((-1, -1, 1), (-1, -1, -1), (1, -1, -1), (1, -1, 1)),
          ((-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)), ((1, -1, -1), (-1, -1, -1), (-1, 1, -1), (1, 1, -1))]


def box_triangles(boxes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """World-space triangles [12 n][3][3] float32 of boxes rotated about +Y (records of sgi_scene_boxes:
    centre, cos yaw, half extents, sin yaw, albedo ...), and their base colours as RGBA8 (UNORM8 albedo,
    alpha 255)."""
    b = np.asarray(boxes, np.float32)
    n = b.shape[0]
    c, cy, h, sy = b[:, 0:3], b[:, 3], b[:, 4:7], b[:, 7]
    corners = np.zeros((n, 6, 4, 3), np.float32)
    for f, quad in enumerate(_FACES):
        for k, sgn in enumerate(quad):
            lx, ly, lz = (np.float32(sgn[i]) * h[:, i] for i in range(3))
            corners[:, f, k, 0] = c[:, 0] + (cy * lx + sy * lz)
            corners[:, f, k, 1] = c[:, 1] + ly
            corners[:, f, k, 2] = c[:, 2] + (-sy * lx + cy * lz)
    tri = np.stack([corners[:, :, [0, 1, 2]], corners[:, :, [0, 2, 3]]], axis=2)  # [n][6][2][3][3]
    pos = np.ascontiguousarray(tri.reshape(n * 12, 3, 3), np.float32)
    alb = np.clip(b[:, 8:11], 0, 1)
    code = np.floor(alb * np.float32(255) + np.float32(0.5)).astype(np.uint32)
    rgba = code[:, 0] | (code[:, 1] << 8) | (code[:, 2] << 16) | np.uint32(0xFF000000)
    return pos, np.repeat(rgba, 12).astype(np.uint32)


def _cube_dtype(fmt: int):
    return {_abi.TEXEL_F32: np.float32, _abi.TEXEL_F16: np.float16, _abi.TEXEL_UNORM16: np.uint16}[fmt]


def _cube_dtype_torch(fmt: int):
    import torch
    # torch has no general uint16 tensor: the UNORM16 codes are held as int16 bits
    return {_abi.TEXEL_F32: torch.float32, _abi.TEXEL_F16: torch.float16, _abi.TEXEL_UNORM16: torch.int16}[fmt]


def generate_host(scene: Scene) -> Inputs:
    """Ray-casts every input on the host (small workloads)."""
    w, lib, h = scene.w, scene.lib, scene.handle
    W, H = w.width, w.height
    depth = np.zeros((H, W), np.float32)
    normal = np.zeros((H, W, 3), np.float32)
    albedo = np.zeros((H, W, 4), np.float32)
    refl = np.zeros((H, W, 2), np.float32)
    rma = np.zeros((H, W, 3), np.float32)
    vel = np.zeros((H, W, 2), np.float32)
    ids = np.zeros((H, W), np.float32)
    _abi.call(lib, "sgi_scene_gbuffer_host", h, depth.ctypes.data, normal.ctypes.data, albedo.ctypes.data,
              refl.ctypes.data, rma.ctypes.data, vel.ctypes.data, ids.ctypes.data)
    vpls, entries = scene.vpls()
    lpa, na = layers_per_atlas(w.num_vpls), num_atlases(w.num_vpls)
    atlases = [np.zeros((lpa, 6, w.cube_res, w.cube_res), _cube_dtype(w.cube_format)) for _ in range(na)]
    inp = Inputs(W, H, scene.frame, scene.csm, depth, normal, albedo, refl, rma, vel, ids, vpls, entries,
                 atlases, w.cube_res, w.cube_format, lpa)
    cubes = inp.cubes()
    _abi.call(lib, "sgi_scene_cubes_host", h, ctypes.byref(cubes), 0, w.num_vpls)
    inp.csm_depth = np.zeros((4, w.csm_res, w.csm_res), np.float32)
    _abi.call(lib, "sgi_scene_csm_host", h, inp.csm_depth.ctypes.data)
    return inp


def generate_device(scene: Scene, device="cuda", stream=None) -> Inputs:
    """Ray-casts every input on the GPU into torch tensors (any workload size)."""
    import torch

    w, lib, h = scene.w, scene.lib, scene.handle
    W, H = w.width, w.height
    f32 = dict(dtype=torch.float32, device=device)
    depth = torch.empty((H, W), **f32)
    normal = torch.empty((H, W, 3), **f32)
    albedo = torch.empty((H, W, 4), **f32)
    refl = torch.empty((H, W, 2), **f32)
    rma = torch.empty((H, W, 3), **f32)
    vel = torch.empty((H, W, 2), **f32)
    ids = torch.empty((H, W), **f32)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _abi.call(lib, "sgi_scene_gbuffer_device", h, depth.data_ptr(), normal.data_ptr(), albedo.data_ptr(),
              refl.data_ptr(), rma.data_ptr(), vel.data_ptr(), ids.data_ptr(), s)
    vpls_np, entries_np = scene.vpls()
    vpls = torch.from_numpy(vpls_np).to(device)
    entries = torch.from_numpy(entries_np).to(device)
    lpa, na = layers_per_atlas(w.num_vpls), num_atlases(w.num_vpls)
    cdt = _cube_dtype_torch(w.cube_format)
    atlases = [torch.empty((lpa, 6, w.cube_res, w.cube_res), dtype=cdt, device=device) for _ in range(na)]
    inp = Inputs(W, H, scene.frame, scene.csm, depth, normal, albedo, refl, rma, vel, ids, vpls, entries,
                 atlases, w.cube_res, w.cube_format, lpa)
    cubes = inp.cubes()
    _abi.call(lib, "sgi_scene_cubes_device", h, ctypes.byref(cubes), 0, w.num_vpls, s)
    inp.csm_depth = torch.empty((4, w.csm_res, w.csm_res), **f32)
    _abi.call(lib, "sgi_scene_csm_device", h, inp.csm_depth.data_ptr(), s)
    torch.cuda.synchronize()
    return inp


def diffuse_cubes(scene: Scene, inp: Inputs, stream=None) -> tuple[list, _abi.CubeAtlases]:
    """The VPLs' diffuse (colour) cube atlases, RGBA8 [layers][6][res][res][4], with the geometry of
    inp's distance atlases: numpy arrays ray-cast on the host when inp holds host data, torch uint8
    tensors ray-cast on the device otherwise."""
    shape = (inp.layers, 6, inp.cube_res, inp.cube_res, 4)
    host = isinstance(inp.atlases[0], np.ndarray)
    if host:
        atlases = [np.zeros(shape, np.uint8) for _ in inp.atlases]
    else:
        import torch
        atlases = [torch.empty(shape, dtype=torch.uint8, device=inp.atlases[0].device) for _ in inp.atlases]
    c = _abi.CubeAtlases()
    c.resolution, c.format, c.layers_per_atlas, c.num_atlases = inp.cube_res, _abi.TEXEL_RGBA8, inp.layers, len(atlases)
    for i, a in enumerate(atlases):
        c.atlas[i] = inp.ptr(a)
    if host:
        _abi.call(scene.lib, "sgi_scene_cubes_host", scene.handle, ctypes.byref(c), 0, scene.w.num_vpls)
    else:
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _abi.call(scene.lib, "sgi_scene_cubes_device", scene.handle, ctypes.byref(c), 0, scene.w.num_vpls, s)
        torch.cuda.synchronize()
    return atlases, c


def to_host(inp: Inputs) -> Inputs:
    """Copies device inputs to numpy (for the oracle)."""
    def h(a):
        if a is None or isinstance(a, np.ndarray):
            return a
        return a.detach().cpu().numpy()
    return Inputs(inp.width, inp.height, inp.frame, inp.csm, h(inp.depth), h(inp.normal), h(inp.albedo),
                  h(inp.refl), h(inp.rma), h(inp.velocity), h(inp.ids), h(inp.vpls), h(inp.entries),
                  [h(a) for a in inp.atlases], inp.cube_res, inp.cube_format, inp.layers, h(inp.csm_depth))


def generate_view_device(scene: Scene, shared: Inputs, device="cuda", stream=None) -> Inputs:
    """The G-buffer and cascades of the scene's current view (sized by the view), with the VPL
    records and cube atlases of `shared` (another view of the same scene: they do not depend on
    the camera)."""
    import torch

    def clone(st):
        c = type(st)()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(st), ctypes.sizeof(st))
        return c

    lib, h = scene.lib, scene.handle
    W, H = scene.view.width, scene.view.height
    f32 = dict(dtype=torch.float32, device=device)
    planes = [torch.empty((H, W), **f32), torch.empty((H, W, 3), **f32), torch.empty((H, W, 4), **f32),
              torch.empty((H, W, 2), **f32), torch.empty((H, W, 3), **f32), torch.empty((H, W, 2), **f32),
              torch.empty((H, W), **f32)]
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _abi.call(lib, "sgi_scene_gbuffer_device", h, *[p.data_ptr() for p in planes], s)
    inp = Inputs(W, H, clone(scene.frame), clone(scene.csm), *planes, shared.vpls, shared.entries,
                 shared.atlases, shared.cube_res, shared.cube_format, shared.layers)
    inp.csm_depth = torch.empty((4, scene.view.csm_resolution, scene.view.csm_resolution), **f32)
    _abi.call(lib, "sgi_scene_csm_device", h, inp.csm_depth.data_ptr(), s)
    torch.cuda.synchronize()
    return inp


class PackedAttachments:
    """The reference's G-buffer attachments on the device (sgi_gbuffer_packed: RGB8 normals, RGBA8
    albedo, RG8 reflectivity, RGB8 roughness / metallic / ambient, RG16F velocity), written from a
    view's FP32 planes by sgi_gbuffer_pack (the geometry pass's attachment writes)."""

    def __init__(self, width: int, height: int, velocity: bool = True, device="cuda"):
        import torch

        i32 = dict(dtype=torch.int32, device=device)
        self.width, self.height = width, height
        self.normals = torch.zeros((height, width), **i32)
        self.albedo = torch.zeros((height, width), **i32)
        self.refl = torch.zeros((height, width), dtype=torch.int16, device=device)
        self.rma = torch.zeros((height, width), **i32)
        self.velocity = torch.zeros((height, width), **i32) if velocity else None
        self.struct = _abi.GbufferPacked(width, height, self.normals.data_ptr(), self.albedo.data_ptr(),
                                         self.refl.data_ptr(), self.rma.data_ptr(),
                                         self.velocity.data_ptr() if velocity else None)

    @classmethod
    def pack(cls, gi, inp: Inputs, stream=None) -> "PackedAttachments":
        """Attachments holding inp's FP32 planes as the geometry pass stores them (gi: a GiPass of
        the view's size)."""
        a = cls(inp.width, inp.height, inp.velocity is not None)
        gi.gbuffer_pack(inp.gbuffer(), a.struct, stream=stream)
        return a
