"""The reference's VPL distance-cube format, GL_DEPTH_COMPONENT16 (UNORM16).

CreateShadowMap3DCache_ allocates the VPL cubes' depth attachment as DEPTH,
BITS_16 (StratusRendererBackend.cpp:296,2412), which StratusTexture.cpp:542
maps to GL_DEPTH_COMPONENT16: code k reads back as k / 65535.  shadowVpl.fs
writes saturate(distance / farPlane) into it.  Pinned here:
- the device decode y + y (2^-16 + 2^-32) with y = k 2^-16, rounded once by an
  FMA (the C library's fmaf rounds exactly as the GPU's v_fma_f32), equals the
  correctly rounded k / 65535 for every code;
- the oracle's cube fetch reads every code as that value;
- the generator writes rint(saturate(v) 65535) of the FP32 value it writes in
  the F32 format, texel for texel;
- the oracle's gather on UNORM16 cubes equals its gather on F32 cubes holding
  the decoded values, bit for bit (the format enters only through the fetch).
"""
import ctypes
import dataclasses

import numpy as np

import oracle as orc
from stratusgfx_amd import _abi
from stratusgfx_amd import scene as sc

L = orc.lib()
CODES = np.arange(65536, dtype=np.uint32)
EXACT = (CODES.astype(np.float64) / 65535.0).astype(np.float32)  # no double-rounding case: k / 65535 is never
#                                                                    within 2^-53 of a float32 midpoint


def test_exact_values_are_correctly_rounded_quotients():
    # neighbours in float32 bracket the rational k / 65535 with EXACT the nearer one
    from fractions import Fraction
    for k in list(range(0, 65536, 257)) + [1, 2, 3, 32767, 32768, 65534, 65535]:
        q = Fraction(k, 65535)
        v = float(EXACT[k])
        up, dn = float(np.nextafter(EXACT[k], np.float32(2))), float(np.nextafter(EXACT[k], np.float32(-1)))
        assert abs(Fraction(v) - q) <= abs(Fraction(up) - q) and abs(Fraction(v) - q) <= abs(Fraction(dn) - q), k


def test_device_decode_formula_is_correctly_rounded_for_every_code():
    libm = ctypes.CDLL("libm.so.6")
    libm.fmaf.restype = ctypes.c_float
    libm.fmaf.argtypes = [ctypes.c_float] * 3
    b = np.float32(2.0 ** -16 + 2.0 ** -32)
    assert float(b) == 2.0 ** -16 + 2.0 ** -32  # representable: 17 significant bits
    y = (CODES.astype(np.float32) * np.float32(2.0 ** -16))
    got = np.array([libm.fmaf(float(yy), float(b), float(yy)) for yy in y], np.float32)
    assert np.array_equal(got.view(np.uint32), EXACT.view(np.uint32))
    # the one-multiply shortcut is not exact, which is why the FMA form is used
    assert np.any(CODES.astype(np.float32) * np.float32(1.0 / 65535.0) != EXACT)


def _codes_atlas(res=256):
    a = np.zeros((1, 6, res, res), np.uint16)
    a[0, 0] = CODES.reshape(res, res)  # face +X holds every code once, k = j * res + i
    c = _abi.CubeAtlases()
    c.resolution, c.format, c.layers_per_atlas, c.num_atlases = res, _abi.TEXEL_UNORM16, 1, 1
    c.atlas[0] = a.ctypes.data
    return a, c


def test_oracle_fetch_decodes_every_code():
    res = 256
    a, c = _codes_atlas(res)
    i = np.arange(res)
    sc_ = 2 * (i + 0.5) / res - 1
    got = np.empty(65536, np.float32)
    for j in range(res):
        tc = 2 * (j + 0.5) / res - 1
        for ii in range(res):
            # face +X: s = (-z / x + 1) / 2, t = (-y / x + 1) / 2
            got[j * res + ii] = L.orc_cube_fetch(ctypes.byref(c), 0, 0, 1.0, -tc, -sc_[ii])
    assert np.array_equal(got.view(np.uint32), EXACT.view(np.uint32))


def _cornell(fmt):
    w = dataclasses.replace(sc.CORNELL_640x360, name=f"cornell-96x64-fmt{fmt}", width=96, height=64, num_vpls=12,
                            cube_res=32, csm_res=128, cube_format=fmt)
    return sc.generate_host(sc.Scene(w))


def test_generator_writes_rounded_codes_of_the_f32_values():
    f32, u16 = _cornell(_abi.TEXEL_F32), _cornell(_abi.TEXEL_UNORM16)
    for af, au in zip(f32.atlases, u16.atlases):
        assert au.dtype == np.uint16
        want = np.rint(np.clip(af, 0, 1).astype(np.float32) * np.float32(65535)).astype(np.uint16)
        assert np.array_equal(au, want)
        assert len(np.unique(au)) > 100  # real distances, not only the clear value


def test_oracle_gather_on_unorm16_equals_f32_cubes_of_the_decoded_values():
    u16 = _cornell(_abi.TEXEL_UNORM16)
    dec = dataclasses.replace(u16, atlases=[EXACT[a.astype(np.int64)] for a in u16.atlases],
                              cube_format=_abi.TEXEL_F32)
    hg, hcsm = u16.gbuffer(), u16.csm_struct()
    upd, vis = orc.cull_global(u16.frame, hcsm, u16.vpls)
    assert vis[0] > 0
    ent = u16.entries[vis[1:]]
    for mode in (orc.GATHER_RANDOM, orc.GATHER_CONVERGED):
        a = orc.gather(u16.frame, hg, upd, int(vis[0]), ent, u16.cubes(), mode=mode)
        b = orc.gather(u16.frame, hg, upd, int(vis[0]), ent, dec.cubes(), mode=mode)
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
