"""ctypes mirror of include/sgi.h, include/sgi_kpcn.h and include/sgi_scene.h.

The structures below are byte-for-byte the C structures (natural alignment),
so a caller can fill them from torch device pointers or numpy host pointers.
"""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import POINTER, c_float, c_int32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(_HERE)
INCLUDE_DIR = os.path.join(REPO, "include")
LIB_PATH = os.path.join(_HERE, "libsgi.so")
HOST_LIB_PATH = os.path.join(_HERE, "libsgi_host.so")

SGI_OK = 0
SGI_E_INVALID = -1
SGI_E_HIP = -2
SGI_E_STATE = -3
MAX_TOTAL_VPLS_BEFORE_CULLING = 10000
MAX_TOTAL_SHADOW_ATLASES = 14
MAX_TOTAL_SHADOWS_PER_ATLAS = 320
MAX_TOTAL_VPLS_PER_FRAME = MAX_TOTAL_SHADOW_ATLASES * MAX_TOTAL_SHADOWS_PER_ATLAS
MAX_VPLS_PER_TILE_LIMIT = 64
NUM_CASCADES = 4
TEXEL_F32 = 0
TEXEL_F16 = 1
TEXEL_RGBA8 = 2
TEXEL_UNORM16 = 3  # GL_DEPTH_COMPONENT16, the reference's VPL cube format
REFERENCE_VPL_COLOR_MULTIPLIER = 50000.0  # vpl_light_color.cs:44
PRECISION_FP32 = 0
PRECISION_FP16 = 1
SCENE_CORNELL = 0
SCENE_CITY = 1

VPL_FLOATS = 16  # sgi_vpl is 16 floats (64 B)
KPCN_CHANNELS = 32
KPCN_LEVELS = 5
KPCN_LAYERS = 17
KPCN_LAYER_WEIGHTS = 9 * 32 * 32
KPCN_LAYER_PARAMS = KPCN_LAYER_WEIGHTS + 32
KPCN_PARAMS = KPCN_LAYERS * KPCN_LAYER_PARAMS
KPCN_TAPS = 25
KPCN_ERR_CHANNEL = 25
KPCN_BLEND_CHANNEL = 26
KPCN_FEATURES = 10


class Gbuffer(ctypes.Structure):
    _fields_ = [
        ("width", c_int32), ("height", c_int32),
        ("depth", c_void_p), ("normal", c_void_p), ("albedo", c_void_p),
        ("baseReflectivity", c_void_p), ("roughnessMetallicAmbient", c_void_p),
        ("velocity", c_void_p), ("id", c_void_p), ("ssao", c_void_p),
        ("packed", c_void_p),  # const sgi_gbuffer_packed* (NULL: the FP32 planes)
    ]


class CubeAtlases(ctypes.Structure):
    _fields_ = [
        ("resolution", c_int32), ("format", c_int32), ("layers_per_atlas", c_int32),
        ("num_atlases", c_int32), ("atlas", c_void_p * MAX_TOTAL_SHADOW_ATLASES),
    ]


class Csm(ctypes.Structure):
    _fields_ = [
        ("resolution", c_int32), ("depth", c_void_p),
        ("cascadeProjViews", (c_float * 16) * NUM_CASCADES),
        ("cascadePlanes", (c_float * 4) * (NUM_CASCADES - 1)),
        ("znear", c_float), ("zfar", c_float), ("depthBias", c_float),
    ]


class Frame(ctypes.Structure):
    _fields_ = [
        ("invProjectionView", c_float * 16), ("prevInvProjectionView", c_float * 16),
        ("viewPosition", c_float * 3), ("infiniteLightDirection", c_float * 3),
        ("infiniteLightColor", c_float * 3), ("time", c_float), ("framesPerSecond", c_float),
        ("frameCount", c_int32), ("minRoughness", c_float), ("usePerceptualRoughness", c_int32),
        ("minGiOcclusionFactor", c_float), ("fogDensity", c_float), ("fogColor", c_float * 3),
        ("emissionStrength", c_float),
    ]


class Config(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("tile_size", c_int32),
                ("max_vpls_per_tile", c_int32)]


class GbufferPacked(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("normals", c_void_p), ("albedo", c_void_p),
                ("baseReflectivity", c_void_p), ("roughnessMetallicAmbient", c_void_p), ("velocity", c_void_p)]


class FrameIO(ctypes.Structure):
    _fields_ = [
        ("gbuffer", POINTER(Gbuffer)), ("prevGbuffer", POINTER(Gbuffer)),
        ("vplData", c_void_p), ("numVpls", c_int32), ("shadowIndices", c_void_p),
        ("cubes", POINTER(CubeAtlases)), ("useTiles", c_int32), ("precision", c_int32),
        ("vplUpdatedData", c_void_p), ("visibleIndices", c_void_p), ("visibleShadowIndices", c_void_p),
        ("tiles", c_void_p), ("direct", c_void_p), ("color", c_void_p), ("reservoir", c_void_p),
        ("prevGi", c_void_p), ("prevHistory", c_void_p), ("combined", c_void_p), ("giOut", c_void_p),
        ("reservoirOut", c_void_p), ("historyOut", c_void_p),
        ("diffuseCubes", POINTER(CubeAtlases)), ("vplColorMultiplier", ctypes.c_float),
    ]


class ViewsConfig(ctypes.Structure):
    _fields_ = [("max_views", c_int32), ("max_pixels", ctypes.c_int64), ("tile_size", c_int32),
                ("max_vpls_per_tile", c_int32)]


class ViewIO(ctypes.Structure):
    _fields_ = [("gbuffer", Gbuffer), ("frame", Frame), ("csm", Csm), ("tiles", c_void_p), ("direct", c_void_p),
                ("color", c_void_p), ("reservoir", c_void_p), ("prevGi", c_void_p), ("prevHistory", c_void_p),
                ("combined", c_void_p), ("giOut", c_void_p), ("reservoirOut", c_void_p), ("historyOut", c_void_p)]


class GatherOpts(ctypes.Structure):
    _fields_ = [("mode", c_int32), ("samples", c_int32), ("precision", c_int32), ("accumulate_fp64", c_int32),
                ("rejection_limit", c_int32)]


GATHER_REFERENCE, GATHER_TILES, GATHER_ALL, GATHER_SAMPLES = 0, 1, 2, 3


class KpcnNet(ctypes.Structure):
    _fields_ = [("weights", c_void_p), ("biases", c_void_p), ("weights_fp8", c_void_p), ("epilogue", c_void_p)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("seed", c_uint64), ("num_vpls", c_int32),
                ("num_city_boxes", c_int32), ("city_half_extent", c_float),
                ("vpl_color_multiplier", c_float)]


class Triangles(ctypes.Structure):
    _fields_ = [("count", c_int32), ("positions", c_void_p), ("colors", c_void_p), ("cull", c_int32)]


CULLING_NONE, CULLING_CW, CULLING_CCW = 0, 1, 2  # RenderFaceCulling; meshes default to CCW


REFERENCE_VPL_NEAR_PLANE = 0.1  # StratusLight.h:291
REFERENCE_CSM_SLOPE_FACTOR = 2.0  # glPolygonOffset(2, 0), StratusRendererBackend.cpp:1097


class FxaaParams(ctypes.Structure):
    _fields_ = [("contrastThreshold", c_float), ("relativeThreshold", c_float), ("subpixelBlending", c_float)]


class View(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("eye", c_float * 3), ("yaw_deg", c_float),
                ("pitch_deg", c_float), ("fovy_deg", c_float), ("znear", c_float), ("zfar", c_float),
                ("csm_resolution", c_int32)]


_P = c_void_p
_I = c_int32

# name -> (restype, argtypes) for every entry point of include/sgi.h and include/sgi_scene.h
SGI_SIGNATURES = {
    "sgi_create": (c_int32, [POINTER(Config), POINTER(_P)]),
    "sgi_destroy": (None, [_P]),
    "sgi_last_error": (ctypes.c_char_p, []),
    "sgi_tile_counts": (c_int32, [_P, POINTER(c_int32), POINTER(c_int32)]),
    "sgi_set_frame": (c_int32, [_P, POINTER(Frame)]),
    "sgi_set_csm": (c_int32, [_P, POINTER(Csm)]),
    "sgi_taa": (c_int32, [_P, _P, _P, _P, _P, _P]),
    "sgi_tonemap": (c_int32, [_P, _P, c_float, _P, _P]),
    "sgi_taa_history": (c_int32, [_P, _P, _P, _P]),
    "sgi_gather_samples_noalloc": (c_int32, [POINTER(Frame), POINTER(Gbuffer), _P, _P, _P, POINTER(CubeAtlases), _I,
                                             _I, _P, _P, _P]),
    "sgi_render_csm": (c_int32, [_P, POINTER(Triangles), POINTER(Csm), c_float, _P, _P]),
    "sgi_render_vpl_cubes": (c_int32, [_P, POINTER(Triangles), _P, _P, _I, _P, c_float, POINTER(CubeAtlases),
                                       POINTER(CubeAtlases), _P]),
    "sgi_fxaa": (c_int32, [_P, _P, _P, _P, _P]),
    "sgi_gbuffer_unpack": (c_int32, [_P, POINTER(GbufferPacked), _P, _P, _P, _P, _P, _P]),
    "sgi_gbuffer_pack": (c_int32, [_P, POINTER(Gbuffer), POINTER(GbufferPacked), _P]),
    "sgi_cull_global": (c_int32, [_P, _P, _I, _P, _P, _P]),
    "sgi_visible_atlas_entries": (c_int32, [_P, _P, _P, _I, _P, _P]),
    "sgi_color_vpls": (c_int32, [_P, _P, _P, POINTER(CubeAtlases), POINTER(CubeAtlases), c_float, _I, _P, _P]),
    "sgi_cull_tiles": (c_int32, [_P, POINTER(Gbuffer), _P, _P, _I, _P, _P, _P]),
    "sgi_gather": (c_int32, [_P, POINTER(Gbuffer), _P, _P, _P, POINTER(CubeAtlases), _P, _I, _P, _P, _P]),
    "sgi_gather_ex": (c_int32, [_P, POINTER(Gbuffer), _P, _P, _P, POINTER(CubeAtlases), _P, POINTER(GatherOpts), _P,
                                _P, _P]),
    "sgi_direct_light": (c_int32, [_P, POINTER(Gbuffer), _P, _P]),
    "sgi_denoise": (c_int32, [_P, POINTER(Gbuffer), POINTER(Gbuffer), _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                              _P]),
    "sgi_set_screen_event": (c_int32, [_P, _P]),
    "sgi_render_frame": (c_int32, [_P, POINTER(FrameIO), _P]),
    "sgi_views_create": (c_int32, [POINTER(ViewsConfig), POINTER(_P)]),
    "sgi_views_destroy": (None, [_P]),
    "sgi_views_set": (c_int32, [_P, _I, POINTER(ViewIO), _P, _I, _P, POINTER(CubeAtlases), _P]),
    "sgi_views_run": (c_int32, [_P, _P]),
    "sgi_views_capture": (c_int32, [_P, _P]),
    "sgi_views_replay": (c_int32, [_P, _P]),
    "sgi_views_set_learned": (c_int32, [_P, _I, _I, POINTER(KpcnNet), c_float]),
    # include/sgi_kpcn.h
    "sgi_kpcn_conv3x3": (c_int32, [_P, _P, _P, _P, _I, _I, _I, _P, _I, _P]),
    "sgi_kpcn_conv3x3_fp8": (c_int32, [_P, _P, _P, _I, _I, _P, _P]),
    "sgi_kpcn_bf16_encode": (c_int32, [_P, ctypes.c_int64, _P, _P]),
    "sgi_kpcn_fp8_encode": (c_int32, [_P, ctypes.c_int64, _P, _P]),
    "sgi_kpcn_wgrad_partials": (ctypes.c_size_t, [_I, _I]),
    "sgi_kpcn_conv3x3_wgrad": (c_int32, [_P, _P, _P, _I, _I, _P, _P, _P]),
    "sgi_kpcn_pool": (c_int32, [_P, _I, _I, _P, _P]),
    "sgi_kpcn_pool_backward": (c_int32, [_P, _I, _I, _P, _I, _P]),
    "sgi_kpcn_upsample_add": (c_int32, [_P, _P, _I, _I, _P, _P]),
    "sgi_kpcn_upsample_backward": (c_int32, [_P, _I, _I, _P, _I, _P]),
    "sgi_kpcn_pyramid_floats": (ctypes.c_size_t, [_I, _I]),
    "sgi_kpcn_radiance_pyramid": (c_int32, [_P, _I, _I, _P, _P]),
    "sgi_kpcn_apply_backward": (c_int32, [_P, _P, _P, _P, _P, _I, _I, _P, _P, _P]),
    "sgi_kpcn_upsample_bilinear_backward": (c_int32, [_P, _I, _I, _P, _P]),
    "sgi_kpcn_features": (c_int32, [_P, POINTER(Gbuffer), _P, _P, _P, _P]),
    "sgi_kpcn_apply": (c_int32, [_P, _P, _P, _I, _I, _P, _P, _P]),
    "sgi_kpcn_loss": (c_int32, [_P, _P, _P, _I, _I, c_float, c_float, _P, _P, _P, _P]),
    "sgi_kpcn_loss_ratio": (c_int32, [_P, _P, _P, _P, _I, _I, c_float, c_float, _P, _P, _P, _P]),
    "sgi_kpcn_adam": (c_int32, [_P, _P, _P, _P, c_float, c_float, c_float, c_float, _I, _P, _P, _P, _P]),
    "sgi_kpcn_pack_weights": (c_int32, [_P, _P, _P, _P, _P]),
    "sgi_kpcn_workspace_bytes": (ctypes.c_size_t, [_I, _I]),
    "sgi_denoise_learned": (c_int32, [_P, POINTER(Gbuffer), POINTER(Gbuffer), _P, _P, _P, _P, _P, _P,
                                      POINTER(KpcnNet), c_float, _P, ctypes.c_size_t, _P, _P, _P, _P, _P, _P]),
}
SCENE_SIGNATURES = {
    "sgi_scene_create": (c_int32, [POINTER(SceneDesc), POINTER(_P)]),
    "sgi_scene_destroy": (None, [_P]),
    "sgi_scene_num_boxes": (c_int32, [_P]),
    "sgi_scene_boxes": (c_int32, [_P, _P]),
    "sgi_scene_default_view": (c_int32, [_P, _I, _I, POINTER(View)]),
    "sgi_scene_set_view": (c_int32, [_P, POINTER(View), POINTER(View), POINTER(Frame), POINTER(Csm)]),
    "sgi_scene_vpls": (c_int32, [_P, _I, _P, _P]),
    "sgi_scene_gbuffer_host": (c_int32, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "sgi_scene_gbuffer_device": (c_int32, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sgi_scene_cubes_host": (c_int32, [_P, POINTER(CubeAtlases), _I, _I]),
    "sgi_scene_cubes_device": (c_int32, [_P, POINTER(CubeAtlases), _I, _I, _P]),
    "sgi_scene_csm_host": (c_int32, [_P, _P]),
    "sgi_scene_csm_device": (c_int32, [_P, _P, _P]),
}
HOST_LIB_SYMBOLS = [n for n in SCENE_SIGNATURES if not n.endswith("_device")] + ["sgi_last_error"]


def header_functions(path: str) -> list[str]:
    """Names of the functions a public header declares (`int name(`, `void name(`, `size_t name(`,
    `const char* name(`)."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:int|void|size_t|const char\*)\s+(sgi_\w+)\s*\(", text, flags=re.M)


class SgiError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with {code}: {msg}")
        self.code = code


class SgiArgumentError(SgiError, ValueError):
    """A buffer the Python layer refuses before the call (SGI_E_INVALID): the C
    ABI takes plain pointers and cannot check their sizes itself."""

    def __init__(self, msg: str):
        super().__init__("argument check", SGI_E_INVALID, msg)


_libs: dict[str, ctypes.CDLL] = {}


def _bind(lib: ctypes.CDLL, sigs: dict) -> None:
    for name, (res, args) in sigs.items():
        if not hasattr(lib, name):
            continue
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the HIP library.  Raises if it has not been built: there is no fallback."""
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    if path != HOST_LIB_PATH:
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as the one this
        # library links: loading torch first makes the process bind that one runtime for both, so
        # the device memory and streams torch hands over belong to the runtime the kernels run on
        import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    _bind(lib, SGI_SIGNATURES)
    _bind(lib, SCENE_SIGNATURES)
    _libs[path] = lib
    return lib


def load_host() -> ctypes.CDLL:
    """Loads the host-only scene generator library (no HIP runtime needed)."""
    return load(HOST_LIB_PATH)


def check(lib: ctypes.CDLL, fn: str, code: int) -> None:
    if code != SGI_OK:
        msg = lib.sgi_last_error()
        raise SgiError(fn, code, msg.decode() if msg else "")


def call(lib: ctypes.CDLL, fn: str, *args) -> None:
    check(lib, fn, getattr(lib, fn)(*args))
