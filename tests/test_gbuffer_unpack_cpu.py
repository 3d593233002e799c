"""The oracle's reads of the reference's packed G-buffer attachments
(StratusRendererBackend.cpp:384-424), pinned against independent library
routines: every UNORM8 code k of every channel is numpy's correctly rounded
float32 k / 255 (0 -> 0, 255 -> 1, strictly increasing, and re-encoding with
round(x * 255) gives k back); every binary16 bit pattern of the velocity plane
is numpy's float16 -> float32 widening (subnormals, infinities and signed
zeros included); bytes 0..2 of an RGB8 texel are r, g, b and byte 3 is
ignored, byte 3 of an RGBA8 texel is alpha; a NULL velocity plane writes
nothing."""
import numpy as np

import oracle as orc
from stratusgfx_amd import _abi


def _packed(H, W, normals, albedo, refl, rma, vel):
    p = _abi.GbufferPacked()
    p.width, p.height = W, H
    p.normals, p.albedo, p.baseReflectivity = normals.ctypes.data, albedo.ctypes.data, refl.ctypes.data
    p.roughnessMetallicAmbient = rma.ctypes.data
    p.velocity = vel.ctypes.data if vel is not None else None
    return p


def _pack_bytes(b):
    """[..., 4] uint8 -> [...] uint32 texels (byte 0 lowest)."""
    b = b.astype(np.uint32)
    return b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16) | (b[..., 3] << 24)


def test_every_code_and_channel():
    H, W = 4, 256
    k = np.arange(256, dtype=np.uint32)
    bytes4 = np.zeros((H, W, 4), np.uint8)
    for c in range(4):                      # row c carries code k in byte c, the other bytes 77
        bytes4[c] = 77
        bytes4[c, :, c] = k
    tex = np.ascontiguousarray(_pack_bytes(bytes4))
    refl = np.ascontiguousarray((bytes4[..., 0].astype(np.uint16) | (bytes4[..., 1].astype(np.uint16) << 8)))
    vel = np.zeros((H, W, 2), np.float16)
    n, a, r, m, v = orc.gbuffer_unpack(_packed(H, W, tex, tex, refl, tex, vel))
    want = np.float32(k) / np.float32(255.0)
    other = np.float32(77) / np.float32(255.0)
    for c in range(4):
        np.testing.assert_array_equal(a[c, :, c], want)
        assert np.all(np.delete(a[c], c, axis=1) == other)
        if c < 3:
            np.testing.assert_array_equal(n[c, :, c], want)
            np.testing.assert_array_equal(m[c, :, c], want)
            assert np.all(np.delete(n[c], c, axis=1) == other)
        else:                               # byte 3 of an RGB8 texel is not read
            assert np.all(n[3] == other) and np.all(m[3] == other)
        if c < 2:
            np.testing.assert_array_equal(r[c, :, c], want)
    assert want[0] == 0.0 and want[255] == 1.0 and np.all(np.diff(want) > 0)
    np.testing.assert_array_equal(np.round(want.astype(np.float64) * 255.0), k)


def test_every_binary16_velocity():
    H, W = 256, 256
    bits = np.arange(65536, dtype=np.uint32).reshape(H, W)
    vel = np.stack([bits, bits[::-1, ::-1]], -1).astype(np.uint16)
    z32 = np.zeros((H, W), np.uint32)
    z16 = np.zeros((H, W), np.uint16)
    *_, v = orc.gbuffer_unpack(_packed(H, W, z32, z32, z16, z32, vel))
    want = vel.view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(v), nan)
    np.testing.assert_array_equal(v[~nan].view(np.uint32), want[~nan].view(np.uint32))  # signed zeros too


def test_null_velocity_writes_nothing():
    H, W = 2, 3
    z32 = np.zeros((H, W), np.uint32)
    z16 = np.zeros((H, W), np.uint16)
    n, a, r, m, v = orc.gbuffer_unpack(_packed(H, W, z32, z32, z16, z32, None), with_velocity=False)
    assert np.all(v == 0.0) and np.all(n == 0.0)
