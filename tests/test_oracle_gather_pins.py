"""Independent pins of the oracle's per-pixel gather (vpl_pbr_gi.fs:81-207,
oracle/gi_oracle.cpp gatherPixel), CPU only.

Every check builds a G-buffer whose pixels all reconstruct to one fragment P
on the plane z = P.z (normal +z), puts the VPLs and the camera on the z axis
above or below P, and gives each VPL a cube layer of one constant texel, so
that the shadow test, the side / radius test and the BRDF are decided by
construction:
  - at N = L = V the VPL lighting has a closed form: GGX D = 1 / (pi a^2),
    Smith Vis = 1/4, Fresnel F = f0, Lambert 1 / pi, with the albedo
    demodulation of calculateVirtualPointLighting2 (pbr2.glsl:613-657) and
    the attenuation of vplDiffuseAttenuation (pbr.glsl:260-267);
  - a texel of 0 occludes (shadow 1), a texel of 1 with farPlane = radius
    does not (shadow 0) for any VPL nearer than its radius;
  - the estimator's draw sequence follows from the hash (pinned by known
    answers in test_oracle_pins.py) and the reference's loop rules.
The expected values are computed here in float64 from those closed forms,
not by calling the oracle's own formulas.

Each check is a function of the library handle, so that
test_oracle_gather_mutations.py can run the same checks on builds of the
oracle with one token changed and show that each change is caught.  The
mutations each check is meant to catch are named in its docstring.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

import oracle as orc
from stratusgfx_amd import _abi



def f32(x):
    """The value a float32 G-buffer texel, uniform or record field holds for x."""
    return float(np.float32(x))


PI = 3.14159265359          # the shaders' PI (pbr2.glsl), a double in the FP64 twin
EPS = 1e-5                  # PREVENT_DIV_BY_ZERO
MIN_GI_OCC = f32(0.95)      # minGiOcclusionFactor, a float uniform
MIN_ROUGHNESS = f32(0.08)


# ---- scene construction -----------------------------------------------------------------
class Scene:
    """W x H pixels that all reconstruct to fragment (0, 0, 0) with normal +z;
    the camera sits at (0, 0, cam)."""

    def __init__(self, W=1, H=1, cam=3.0, albedo=(0.5, 0.25, 0.125), refl=0.5, rough=0.6, metal=0.0,
                 ambient=0.3, ssao=None, time=0.0):
        self.W, self.H = W, H
        P = W * H
        self.depth = np.full((H, W), 0.5, np.float32)  # ndc z = 0
        self.normal = np.tile(np.array([0.5, 0.5, 1.0], np.float32), (H, W, 1))
        self.albedo = np.tile(np.array([*albedo, 0.0], np.float32), (H, W, 1))
        self.refl = np.tile(np.array([refl, 0.0], np.float32), (H, W, 1))
        self.rma = np.tile(np.array([rough, metal, ambient], np.float32), (H, W, 1))
        self.ssao = None if ssao is None else np.full((H, W), ssao, np.float32)
        assert self.depth.size == P
        self.gbuf = _abi.Gbuffer(W, H, self.depth.ctypes.data, self.normal.ctypes.data, self.albedo.ctypes.data,
                                 self.refl.ctypes.data, self.rma.ctypes.data, 0, 0,
                                 0 if self.ssao is None else self.ssao.ctypes.data)
        f = _abi.Frame()
        # column-major: x = 0 * ndc.x + 0, y = 0 * ndc.y + 0, z = ndc.z, w = 1
        inv = [0.0] * 16
        inv[10] = 1.0
        inv[15] = 1.0
        for i, v in enumerate(inv):
            f.invProjectionView[i] = v
        f.viewPosition[:] = [0.0, 0.0, cam]
        f.infiniteLightDirection[:] = [0.0, 0.0, 1.0]
        f.time = time
        f.framesPerSecond = 60.0
        f.minRoughness = 0.08
        f.usePerceptualRoughness = 1
        f.minGiOcclusionFactor = MIN_GI_OCC
        f.fogDensity = 0.0
        self.frame = f


def vpl_records(vpls):
    """vpls: list of dicts(z, radius, color=(r,g,b), spec_z=None) -> [n,16] records on the z axis."""
    out = np.zeros((len(vpls), 16), np.float32)
    for i, v in enumerate(vpls):
        out[i, 0:4] = [0.0, 0.0, v["z"], 1.0]
        out[i, 4:8] = [*v.get("color", (1.0, 1.0, 1.0)), 1.0]
        out[i, 8:12] = [0.0, 0.0, v.get("spec_z", v["z"]), 1.0]
        out[i, 12] = v["radius"]
        out[i, 13] = v["radius"]  # farPlane = radius (StratusLight.h)
        out[i, 14] = 1.0
    return out


def cubes_for(texels):
    """One F32 atlas, one 2x2 cube layer per VPL holding a constant distance / farPlane."""
    a = np.zeros((len(texels), 6, 2, 2), np.float32)
    for l, t in enumerate(texels):
        a[l] = t
    c = _abi.CubeAtlases()
    c.resolution, c.format, c.layers_per_atlas, c.num_atlases = 2, _abi.TEXEL_F32, len(texels), 1
    c.atlas[0] = a.ctypes.data
    entries = np.array([[0, l] for l in range(len(texels))], np.int32)
    return a, c, entries


def run(L, sc, vpls, texels, mode, tiles=None, precision=1, samples=8, rejection_limit=5):
    recs = vpl_records(vpls)
    keep, cubes, entries = cubes_for(texels)
    tile_arr = None
    ts = max(sc.W, sc.H)
    if tiles is not None:
        K = max(1, len(tiles) - 1)
        tile_arr = np.zeros((1, 1 + K), np.int32)
        tile_arr[0, :len(tiles)] = tiles
    else:
        K = 1
    col, res = orc.gather(sc.frame, sc.gbuf, recs, len(vpls), entries, cubes, tiles=tile_arr, tile_size=ts, K=K,
                          mode=mode, precision=precision, samples=samples, rejection_limit=rejection_limit, L=L)
    del keep
    return col, res


# ---- closed forms (float64) ---------------------------------------------------------------
def atten(d, R):
    d, R = f32(d), f32(R)
    r = min(d / R, 1.0)
    e = math.exp(-3.0 * r)
    return 1.0 / ((10.0 * (1.0 - e) + 1.0 * e) * R + d * d)


def tmp_closed_form(h, R, rough=0.6, albedo=(0.5, 0.25, 0.125), refl=0.5, metal=0.0, ao=1.0,
                    color=(1.0, 1.0, 1.0), spec_h=None):
    """ambientOcclusion * calculateVirtualPointLighting2 at N = L = V for a VPL at height h
    (specular position at spec_h) above the fragment."""
    spec_h = h if spec_h is None else spec_h
    r = max(0.5, f32(rough))
    a = max(MIN_ROUGHNESS, r * r)
    refl, metal, ao = f32(refl), f32(metal), f32(ao)
    out = []
    for k in range(3):
        alb = max(f32(albedo[k]), 0.01)
        f0 = (0.16 * refl * refl * (1.0 - metal) + alb * metal) / (alb + EPS)
        spec = f0 / (4.0 * PI * a * a)
        diff = (1.0 - metal) * alb / (alb + EPS) / PI
        out.append(ao * (spec * atten(spec_h, R) + diff * atten(h, R)) * f32(color[k]))
    return np.array(out)


def occlusion_floor(d, R):
    """min(1, mix(minGiOcclusionFactor, 1, clamp(2 d / R, 0, 1))) for an occluded VPL."""
    ratio = min(max(2.0 * f32(d) / f32(R), 0.0), 1.0)
    return min(1.0, MIN_GI_OCC * (1.0 - ratio) + ratio)


def close(a, b, rel=1e-9, abs_=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


OCCLUDED, OPEN = 0.0, 1.0


# ---- checks -----------------------------------------------------------------------------
def check_unoccluded_light_is_the_closed_form(L):
    """An unoccluded VPL straight above the fragment: rgb = the N = L = V closed form, w = 1.
    Catches: the roughness floor max(0.5, r) changed, the albedo floor 0.01 changed, the AO factor
    replaced by the G-buffer's ambient channel, diffuse / specular positions swapped, roughness and
    metallic swapped, (1 - shadow) turned into (1 + shadow)."""
    cases = [dict(h=2.0, R=10.0, rough=0.6, spec_h=2.0), dict(h=1.0, R=4.0, rough=0.3, spec_h=2.5),
             dict(h=3.0, R=30.0, rough=0.9, spec_h=0.75), dict(h=0.5, R=2.0, rough=0.8, spec_h=0.5, albedo=(0.005, 0.9, 0.3))]
    for c in cases:
        albedo = c.get("albedo", (0.5, 0.25, 0.125))
        sc = Scene(rough=c["rough"], albedo=albedo, cam=5.0)
        vp = [dict(z=c["h"], radius=c["R"], spec_z=c["spec_h"], color=(2.0, 1.0, 0.5))]
        want = tmp_closed_form(c["h"], c["R"], rough=c["rough"], albedo=albedo, color=(2.0, 1.0, 0.5),
                               spec_h=c["spec_h"])
        for prec, rel in ((1, 1e-9), (0, 2e-6)):
            col, res = run(L, sc, vp, [OPEN], orc.GATHER_TILES, tiles=[1, 0], precision=prec)
            assert close(res[0, 0, :3], want, rel=rel), (c, prec, res[0, 0, :3], want)
            assert res[0, 0, 3] == 1.0, (c, res[0, 0, 3])
            # color output: max(albedo, 0.01) + 1e-5
            assert close(col[0, 0], [max(f32(x), 0.01) + EPS for x in albedo], rel=1e-6), col[0, 0]


def check_occlusion_floor_and_reweighting(L):
    """An occluded VPL at distance d: shadow = min(1, mix(0.95, 1, clamp(2 d / R))),
    rgb = (1 - shadow) * light, w = (1 - ratio) 0.95 + ratio.  The VPL is listed twice, so that
    sum w = 2 w > 1 is not hidden by the w >= 1 floor.
    Catches: 2 d / R turned into d / R, mix's ends swapped, (1 - shadow) * light turned into
    (1 + shadow) * light."""
    R = 10.0
    for d in (0.5, 1.0, 2.0, 2.4, 4.0, 7.5):
        sc = Scene(cam=5.0)
        sf = occlusion_floor(d, R)
        want = 2.0 * (1.0 - sf) * tmp_closed_form(d, R)
        for prec, rel, ab in ((1, 1e-9, 1e-15), (0, 2e-6, 1e-9)):
            _, res = run(L, sc, [dict(z=d, radius=R)], [OCCLUDED], orc.GATHER_TILES, tiles=[2, 0, 0],
                         precision=prec)
            assert close(res[0, 0, :3], want, rel=rel, abs_=ab * max(1.0, float(np.max(want)))), (d, prec)
            assert close(res[0, 0, 3], 2.0 * sf, rel=rel * 10), (d, prec, res[0, 0, 3], sf)
    # the reweighting is taken only when the shadow is > 0: an unoccluded VPL at ratio 0.2 weighs 1, not 0.96
    sc = Scene(cam=5.0)
    _, res = run(L, sc, [dict(z=1.0, radius=R), dict(z=1.0, radius=R)], [OPEN, OPEN], orc.GATHER_TILES,
                 tiles=[2, 0, 1])
    assert res[0, 0, 3] == 2.0, res[0, 0, 3]


def check_shadow_bias_is_a_tenth_of_the_distance(L):
    """calculateShadowValue1Sample with minBias 0.1 (vpl_pbr_gi.fs:184): the slope term
    saturate(dot(L - P, N)) * 0.03 <= 0.03 never wins, so the VPL is occluded exactly when
    0.9 |P - L| > texel * farPlane.  Catches: minBias 0.1 turned into 0.01."""
    R, h = 10.0, 2.0
    sc = Scene(cam=5.0)
    for texel, occluded in ((0.19, False), (0.185, False), (0.17, True), (0.0, True), (1.0, False)):
        _, res = run(L, sc, [dict(z=h, radius=R)], [texel], orc.GATHER_TILES, tiles=[2, 0, 0])
        sf = occlusion_floor(h, R) if occluded else 0.0
        assert close(res[0, 0, :3], 2.0 * (1.0 - sf) * tmp_closed_form(h, R), rel=1e-9, abs_=1e-15), (texel, res[0, 0])
        assert close(res[0, 0, 3], 2.0 * sf if occluded else 2.0), (texel, res[0, 0, 3])


def check_shadow_only_nearer_than_500(L):
    """distToCamera < 500 ? shadow test : 0 (vpl_pbr_gi.fs:181-187): an occluded VPL is shadowed for a
    camera at 499 and not at 501.  Catches: the 500 cutoff changed."""
    R, h = 10.0, 1.0
    for cam, shadowed in ((499.0, True), (501.0, False), (4000.0, False)):
        sc = Scene(cam=cam)
        _, res = run(L, sc, [dict(z=h, radius=R)], [OCCLUDED], orc.GATHER_TILES, tiles=[2, 0, 0])
        sf = occlusion_floor(h, R) if shadowed else 0.0
        assert close(res[0, 0, :3], 2.0 * (1.0 - sf) * tmp_closed_form(h, R), rel=1e-9, abs_=1e-15), (cam, res[0, 0])
        assert close(res[0, 0, 3], 2.0 * sf if shadowed else 2.0), (cam, res[0, 0, 3])


def check_ssao_is_clamped_to_035_1(L):
    """ambientOcclusion = clamp(ssao, 0.35, 1) multiplies the light (vpl_pbr_gi.fs:114); no SSAO
    plane reads 1.  Catches: the 0.35 floor changed."""
    R, h = 10.0, 2.0
    base = tmp_closed_form(h, R)
    for ssao, ao in ((None, 1.0), (0.1, 0.35), (0.6, f32(0.6)), (1.7, 1.0), (0.35, 0.35)):
        sc = Scene(cam=5.0, ssao=ssao)
        _, res = run(L, sc, [dict(z=h, radius=R)], [OPEN], orc.GATHER_TILES, tiles=[1, 0])
        assert close(res[0, 0, :3], ao * base, rel=1e-9), (ssao, res[0, 0, :3] / base)


def check_tile_list_evaluates_each_entry_once(L):
    """TILES: each listed VPL once, skipping back-facing ones and ones beyond their radius; a
    duplicated entry counts twice; unlisted VPLs are not read; an empty list gives w = 1, no light.
    Catches: the list read from slot 0 instead of 1, the side test's sign flipped, side || radius
    turned into side && radius, the w >= 1 floor removed."""
    R = 10.0
    vp = [dict(z=2.0, radius=R),           # 0 front, in range, open
          dict(z=-2.0, radius=R),          # 1 behind the surface
          dict(z=15.0, radius=R),          # 2 front, beyond its radius
          dict(z=3.0, radius=R),           # 3 front, in range, occluded
          dict(z=1.0, radius=R, color=(9.0, 9.0, 9.0))]  # 4 never listed
    tex = [OPEN, OPEN, OPEN, OCCLUDED, OPEN]
    sc = Scene(cam=5.0)
    sf3 = occlusion_floor(3.0, R)
    want_rgb = tmp_closed_form(2.0, R) + (1.0 - sf3) * tmp_closed_form(3.0, R)
    _, res = run(L, sc, vp, tex, orc.GATHER_TILES, tiles=[4, 0, 1, 2, 3])
    assert close(res[0, 0, :3], want_rgb, rel=1e-9), res[0, 0]
    assert close(res[0, 0, 3], 1.0 + sf3), res[0, 0, 3]
    _, res = run(L, sc, vp, tex, orc.GATHER_TILES, tiles=[2, 0, 0])
    assert close(res[0, 0, :3], 2.0 * tmp_closed_form(2.0, R), rel=1e-9) and res[0, 0, 3] == 2.0, res[0, 0]
    _, res = run(L, sc, vp, tex, orc.GATHER_TILES, tiles=[1, 3])  # w = sf3 < 1 is floored to 1
    assert res[0, 0, 3] == 1.0 and close(res[0, 0, :3], (1.0 - sf3) * tmp_closed_form(3.0, R)), res[0, 0]
    _, res = run(L, sc, vp, tex, orc.GATHER_TILES, tiles=[0])
    assert list(res[0, 0]) == [0.0, 0.0, 0.0, 1.0], res[0, 0]
    # converged mode: every visible VPL once, the same skips
    _, res = run(L, sc, vp[:4], tex[:4], orc.GATHER_CONVERGED)
    assert close(res[0, 0, :3], want_rgb, rel=1e-9) and close(res[0, 0, 3], 1.0 + sf3), res[0, 0]


def check_radiance_is_bounded_to_65504(L):
    """boundHDR (common.glsl:89-92): the summed light is clamped to 65504 per channel.
    Catches: the bound changed."""
    sc = Scene(cam=5.0)
    _, res = run(L, sc, [dict(z=2.0, radius=10.0, color=(1e12, 1e12, 1e-3))], [OPEN], orc.GATHER_TILES,
                 tiles=[1, 0])
    assert res[0, 0, 0] == 65504.0 and res[0, 0, 1] == 65504.0 and 0.0 < res[0, 0, 2] < 1.0, res[0, 0]


def reference_sample_count(rough, dcam):
    """vpl_pbr_gi.fs:139-143: int(mix(2, 10, 1 - max(0.5, r))) scaled by min(1 - d / 1000, 1),
    truncated, at least 1 (cases chosen away from integer boundaries)."""
    rw = 1.0 - max(0.5, rough)
    m = int(2.0 * (1.0 - rw) + 10.0 * rw)
    return max(1, int(m * min(1.0 - dcam / 1000.0, 1.0)))


SAMPLE_COUNT_CASES = [(0.3, 1.0, 5), (0.62, 250.0, 3), (0.9, 100.0, 1), (0.55, 550.0, 2), (1.0, 10.0, 1),
                      (0.5, 2000.0, 1), (0.05, 120.0, 5), (0.7, 60.0, 3)]


def check_reference_estimator_sample_count(L):
    """RANDOM mode with every draw accepted and unoccluded: sum w = the sample count, and light is
    present.  Catches: mix(2, 10) ends swapped, 1000 turned into 100, min(1 - d / 1000, 1) turned
    into max, max(1, n) turned into max(0, n), the 0.5 roughness floor changed."""
    vp = [dict(z=1.0, radius=10.0), dict(z=2.0, radius=10.0), dict(z=3.0, radius=10.0)]
    for rough, dcam, n in SAMPLE_COUNT_CASES:
        assert reference_sample_count(rough, dcam) == n
        sc = Scene(W=4, H=4, rough=rough, cam=dcam)
        _, res = run(L, sc, vp, [OPEN] * 3, orc.GATHER_RANDOM)
        assert np.all(res[..., 3] == n), (rough, dcam, n, np.unique(res[..., 3]))
        assert np.all(res[..., :3] > 0), (rough, dcam)


def _random(L, x, y, z):
    return np.float32(L.orc_random3(np.float32(x), np.float32(y), np.float32(z)))


def reference_draws(L, px, py, time, n_visible, count, accepts, limit=5, slot_of=None, fixed=False):
    """The draw sequence of the reference's loop (vpl_pbr_gi.fs:145-167): seed (x + 0.5, y + 0.5, time),
    seed.z += 10000 per draw, slot = int((N - 1) rand); a draw the side / radius test fails is drawn
    again while fewer than `limit` draws were rejected.  fixed: exactly `count` draws, failing ones
    skipped.  Returns the accepted slots in order."""
    z = np.float32(time)
    out, i, rejected = [], 0, 0
    while i < count:
        r = _random(L, px + 0.5, py + 0.5, z)
        z = np.float32(z + np.float32(10000.0))
        s = slot_of(r)
        if fixed:
            if accepts[s]:
                out.append(s)
            i += 1
            continue
        if rejected < limit and not accepts[s]:
            rejected += 1
            continue
        out.append(s)
        i += 1
    return out


def check_reference_estimator_never_draws_the_last_slot(L):
    """slot = int((N - 1) * rand) (vpl_pbr_gi.fs:147): with the only front-facing VPL in the last slot,
    no pixel receives light; every pixel's draws fall through the 5 rejections onto the back-facing
    slots, which weigh 1 each.  Catches: N - 1 turned into N."""
    for n in (2, 3):
        vp = [dict(z=-1.0 - k, radius=10.0) for k in range(n - 1)] + [dict(z=2.0, radius=10.0)]
        sc = Scene(W=16, H=16, rough=0.6, cam=50.0)
        count = reference_sample_count(0.6, 50.0)
        _, res = run(L, sc, vp, [OPEN] * n, orc.GATHER_RANDOM)
        assert np.all(res[..., :3] == 0.0), (n, int(np.count_nonzero(res[..., 0])))
        assert np.all(res[..., 3] == count), (n, np.unique(res[..., 3]))


def check_reference_estimator_rejection_rule(L):
    """At most 5 rejected draws per pixel, then draws are taken unconditionally, with seed.z += 10000
    per draw.  Slot 0 faces the surface and is occluded (w = 0.96), slots 1-3 are behind it (w = 1
    when taken), and slot 4, a bright unoccluded VPL facing the surface, is never drawn; the
    expected sum w and light of every pixel follow from the hash.
    Catches: the rejection limit 5 turned into 4, the seed step 10000 changed, N - 1 turned into N."""
    R = 10.0
    vp = [dict(z=1.0, radius=R), dict(z=-1.0, radius=R), dict(z=-2.0, radius=R), dict(z=-3.0, radius=R),
          dict(z=2.0, radius=R, color=(100.0, 100.0, 100.0))]
    tex = [OCCLUDED, OPEN, OPEN, OPEN, OPEN]
    accepts = [True, False, False, False, True]
    sf = occlusion_floor(1.0, R)  # 0.96
    light = (1.0 - sf) * tmp_closed_form(1.0, R, rough=0.3)
    forced_pixels = 0
    for time in (0.0, 12345.0):
        sc = Scene(W=16, H=16, rough=0.3, cam=5.0, time=time)
        count = reference_sample_count(0.3, 5.0)
        for prec in (1, 0):
            _, res = run(L, sc, vp, tex, orc.GATHER_RANDOM, precision=prec)
            if prec == 1:
                slot_of = lambda r: int(4.0 * float(r))
            else:
                slot_of = lambda r: int(np.float32(4.0) * r)
            mismatches = 0
            for y in range(sc.H):
                for x in range(sc.W):
                    draws = reference_draws(L, x, y, time, 5, count, accepts, slot_of=slot_of)
                    front = sum(1 for s in draws if accepts[s])
                    forced_pixels += front < len(draws)
                    want_w = max(1.0, front * sf + (len(draws) - front) * 1.0)
                    if not (close(res[y, x, 3], want_w, rel=1e-6)
                            and close(res[y, x, :3], front * light, rel=1e-5, abs_=1e-12)):
                        mismatches += 1
            assert mismatches == 0, (time, prec, mismatches)
    assert forced_pixels > 100, forced_pixels  # the fall-through is exercised


def check_fixed_count_sampler(L):
    """The SAMPLES estimator (no reference counterpart; the same hash and seed step): slot
    min(int(N rand), N - 1); rejection_limit < 0 takes exactly `samples` draws and skips failing
    ones; rejection_limit >= 0 draws again as the reference does.
    Catches: the slot formula turned into int((N - 1) rand)."""
    R = 10.0
    vp = [dict(z=1.0, radius=R), dict(z=-1.0, radius=R), dict(z=2.0, radius=R)]
    tex = [OCCLUDED, OPEN, OPEN]
    accepts = [True, False, True]
    w_of = [occlusion_floor(1.0, R), 1.0, 1.0]
    sc = Scene(W=12, H=12, cam=5.0, time=7.0)
    slot_of = lambda r: min(int(np.float32(3.0) * r), 2)
    for limit, fixed in ((-1, True), (5, False), (0, False)):
        _, res = run(L, sc, vp, tex, orc.GATHER_SAMPLES, samples=6, rejection_limit=limit)
        for y in range(sc.H):
            for x in range(sc.W):
                draws = reference_draws(L, x, y, 7.0, 3, 6, accepts, limit=max(limit, 0), slot_of=slot_of,
                                        fixed=fixed)
                want_w = max(1.0, sum(w_of[s] for s in draws))
                assert close(res[y, x, 3], want_w, rel=1e-9), (limit, x, y, res[y, x, 3], want_w)


CHECKS = [check_unoccluded_light_is_the_closed_form, check_occlusion_floor_and_reweighting,
          check_shadow_bias_is_a_tenth_of_the_distance, check_shadow_only_nearer_than_500,
          check_ssao_is_clamped_to_035_1, check_tile_list_evaluates_each_entry_once,
          check_radiance_is_bounded_to_65504, check_reference_estimator_sample_count,
          check_reference_estimator_never_draws_the_last_slot, check_reference_estimator_rejection_rule,
          check_fixed_count_sampler]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__ for c in CHECKS])
def test_gather_pin(check):
    check(orc.lib())
