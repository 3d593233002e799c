"""The C ABI without a GPU: every declared symbol is exported, the ctypes
mirror has the C layouts, and argument validation fails loudly."""
import ctypes
import os
import subprocess
import tempfile

import pytest

from stratusgfx_amd import _abi


def _headers():
    return [os.path.join(_abi.INCLUDE_DIR, h) for h in ("sgi.h", "sgi_kpcn.h", "sgi_scene.h")]


def test_library_exports_every_declared_function():
    lib = _abi.load()
    declared = set()
    for h in _headers():
        names = _abi.header_functions(h)
        assert names, h
        declared.update(names)
    for n in sorted(declared):
        assert hasattr(lib, n), f"libsgi.so does not export {n}"
    assert declared == set(_abi.SGI_SIGNATURES) | set(_abi.SCENE_SIGNATURES), "ctypes signatures out of date"


def test_host_library_exports_host_generator():
    lib = _abi.load_host()
    for n in _abi.HOST_LIB_SYMBOLS:
        assert hasattr(lib, n), n


STRUCTS = {"sgi_vpl": None, "sgi_atlas_entry": None, "sgi_tile_stage1": None, "sgi_gbuffer": _abi.Gbuffer,
           "sgi_cube_atlases": _abi.CubeAtlases, "sgi_csm": _abi.Csm, "sgi_frame": _abi.Frame,
           "sgi_config": _abi.Config, "sgi_frame_io": _abi.FrameIO, "sgi_scene_desc": _abi.SceneDesc,
           "sgi_view": _abi.View, "sgi_kpcn_net": _abi.KpcnNet, "sgi_gather_opts": _abi.GatherOpts,
           "sgi_views_config": _abi.ViewsConfig, "sgi_view_io": _abi.ViewIO,
           "sgi_gbuffer_packed": _abi.GbufferPacked, "sgi_fxaa_params": _abi.FxaaParams}


CONSTANTS = {"SGI_KPCN_CHANNELS": _abi.KPCN_CHANNELS, "SGI_KPCN_LEVELS": _abi.KPCN_LEVELS,
             "SGI_KPCN_LAYERS": _abi.KPCN_LAYERS, "SGI_KPCN_LAYER_PARAMS": _abi.KPCN_LAYER_PARAMS,
             "SGI_KPCN_PARAMS": _abi.KPCN_PARAMS, "SGI_KPCN_TAPS": _abi.KPCN_TAPS,
             "SGI_KPCN_ERR_CHANNEL": _abi.KPCN_ERR_CHANNEL, "SGI_KPCN_BLEND_CHANNEL": _abi.KPCN_BLEND_CHANNEL,
             "SGI_KPCN_FEATURES": _abi.KPCN_FEATURES, "SGI_TEXEL_RGBA8": _abi.TEXEL_RGBA8,
             "SGI_TEXEL_UNORM16": _abi.TEXEL_UNORM16, "SGI_TEXEL_F16": _abi.TEXEL_F16,
             "SGI_TEXEL_F32": _abi.TEXEL_F32}


def test_struct_layouts_match_c():
    """sizeof of every public struct as gcc lays it out vs the ctypes mirror and the reference sizes;
    the learned denoiser's constants vs the mirror's."""
    src = '#include <stdio.h>\n#include "sgi_scene.h"\n#include "sgi_kpcn.h"\nint main(void){\n'
    for name in STRUCTS:
        src += f'printf("{name} %zu\\n", sizeof({name}));\n'
    for name in CONSTANTS:
        src += f'printf("{name} %d\\n", (int)({name}));\n'
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", _abi.INCLUDE_DIR, c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    sizes = dict(line.split() for line in out.strip().splitlines())
    # StratusGpuCommon.h:297-307 static_asserts
    assert int(sizes["sgi_vpl"]) == 64
    assert int(sizes["sgi_atlas_entry"]) == 8
    assert int(sizes["sgi_tile_stage1"]) == 32
    for name, cls in STRUCTS.items():
        if cls is not None:
            assert ctypes.sizeof(cls) == int(sizes[name]), name
    for name, value in CONSTANTS.items():
        assert int(sizes[name]) == value, name


def test_create_validates_config_before_touching_the_device():
    lib = _abi.load()
    h = ctypes.c_void_p()
    for cfg in [(640, 360, 12, 12), (640, 360, 16, 0), (640, 360, 16, 65), (0, 360, 16, 12), (20000, 10, 8, 1)]:
        c = _abi.Config(*cfg)
        rc = lib.sgi_create(ctypes.byref(c), ctypes.byref(h))
        assert rc == _abi.SGI_E_INVALID, cfg
        assert lib.sgi_last_error().decode().startswith("sgi_create")
    with pytest.raises(_abi.SgiError):
        _abi.call(lib, "sgi_create", None, ctypes.byref(h))


def test_entry_points_reject_null_context():
    lib = _abi.load()
    assert lib.sgi_cull_global(None, None, 0, None, None, None) == _abi.SGI_E_INVALID
    assert lib.sgi_set_frame(None, None) == _abi.SGI_E_INVALID
    assert lib.sgi_render_frame(None, None, None) == _abi.SGI_E_INVALID
    assert lib.sgi_gather(None, None, None, None, None, None, None, 0, None, None, None) == _abi.SGI_E_INVALID
    assert "NULL" in lib.sgi_last_error().decode()


def test_scene_generator_rejects_bad_descriptions():
    lib = _abi.load_host()
    h = ctypes.c_void_p()
    d = _abi.SceneDesc(7, 0, 10, 0, 100.0, 1.0)
    assert lib.sgi_scene_create(ctypes.byref(d), ctypes.byref(h)) == _abi.SGI_E_INVALID
    d = _abi.SceneDesc(_abi.SCENE_CORNELL, 0, 10001, 0, 100.0, 1.0)
    assert lib.sgi_scene_create(ctypes.byref(d), ctypes.byref(h)) == _abi.SGI_E_INVALID


def test_missing_library_raises(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _abi.load(str(tmp_path / "libsgi.so"))
