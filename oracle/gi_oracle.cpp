// ORACLE — test infrastructure only.
//
// A CPU restatement of StratusGFX's deferred VPL GI pass (the GLSL under
// Source/Shaders), written line by line after the reference shaders and
// templated on the real type: R = float is the FP32 restatement that the HIP
// kernels must match, R = double is the FP64 twin used to state the error of
// reduced-precision shading.  Only tests/, __graft_entry__.smoke() and
// bench.py's CPU baseline may load this library, and only as the checker.
//
// Arithmetic follows GLSL as written: left-to-right sums (dot, mat4 * vec4),
// mix(x, y, a) = x * (1 - a) + y * a, length = sqrt(dot(v, v)),
// normalize(v) = v / length(v), pow(x, 5) and pow(x, 128) by exact products,
// correctly rounded division and sqrt, no FMA contraction (built with
// -ffp-contract=off).  Texture reads restate the reference's GL sampler state:
// CLAMP_TO_EDGE everywhere, texel-centre taps, NEAREST cube lookups through
// the GL major-axis table, LINEAR + LEQUAL compare for the cascade array.
//
// Which reference lines each function follows is given on the function.
// Parity pins: tests/test_oracle_*.py (closed forms, brute force, invariants).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../include/sgi.h"

namespace {

// The worker threads of the row-parallel passes (orc_set_threads; 1 by
// default).  Every pixel of those passes depends only on the pass's inputs,
// so the split of rows over threads changes no result.
std::atomic<int> g_threads{1};

template <class F>
void forRows(int y0, int y1, F&& body) {
    const int n = std::max(1, std::min(g_threads.load(), y1 - y0));
    if (n == 1) {
        for (int y = y0; y < y1; ++y) body(y);
        return;
    }
    std::atomic<int> next{y0};
    auto work = [&] {
        for (int y = next.fetch_add(1); y < y1; y = next.fetch_add(1)) body(y);
    };
    std::vector<std::thread> th;
    for (int i = 1; i < n; ++i) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

template <class R>
struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> v3s(R s) { return V3<R>{s, s, s}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> operator+(V3<R> a, R s) { return {a.x + s, a.y + s, a.z + s}; }
template <class R> inline V3<R> operator-(V3<R> a, R s) { return {a.x - s, a.y - s, a.z - s}; }
template <class R> inline V3<R> operator-(R s, V3<R> a) { return {s - a.x, s - a.y, s - a.z}; }
template <class R> inline V3<R> operator/(R s, V3<R> a) { return {s / a.x, s / a.y, s / a.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a / length(a); }
template <class R> inline R gmax(R a, R b) { return a > b ? a : b; }   // GLSL max (NaN: returns b)
template <class R> inline R gmin(R a, R b) { return a < b ? a : b; }
template <class R> inline R clampv(R x, R lo, R hi) { return gmin(gmax(x, lo), hi); }
template <class R> inline R saturate(R x) { return clampv(x, R(0), R(1)); }
template <class R> inline R mixv(R x, R y, R a) { return x * (R(1) - a) + y * a; }
template <class R> inline V3<R> mixv(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <class R> inline V3<R> vmin(V3<R> a, R s) { return {gmin(a.x, s), gmin(a.y, s), gmin(a.z, s)}; }
template <class R> inline R pow5(R x) { const R x2 = x * x; return x2 * x2 * x; }
template <class R> inline R pow128(R x) { for (int i = 0; i < 7; ++i) x = x * x; return x; }
template <class R> inline R rexp(R x) { return std::exp(x); }

// common.glsl:3-4
template <class R> inline R PI_() { return R(3.14159265359); }
template <class R> inline R EPS_() { return R(0.00001); }

// ---- RNG: common.glsl:183-214 ------------------------------------------------
inline uint32_t hash1(uint32_t x) {
    x += (x << 10u);
    x ^= (x >> 6u);
    x += (x << 3u);
    x ^= (x >> 11u);
    x += (x << 15u);
    return x;
}
inline uint32_t hash3(uint32_t x, uint32_t y, uint32_t z) { return hash1(x ^ hash1(y) ^ hash1(z)); }
inline float floatConstruct(uint32_t m) {
    m &= 0x007FFFFFu;
    m |= 0x3F800000u;
    float f;
    std::memcpy(&f, &m, 4);
    return f - 1.0f;
}
inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float random3(float x, float y, float z) { return floatConstruct(hash3(fbits(x), fbits(y), fbits(z))); }

// ---- textures ---------------------------------------------------------------
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// floor(coordinate) -> texel index clamped to [0, n-1]; the value is clamped
// before the int conversion so NaN / huge coordinates stay defined (NaN -> 0).
template <class R> inline int texelIndex(R f, int n) {
    const R c = f > R(-1) ? (f < R(n) ? f : R(n)) : R(-1);
    return clampi((int)c, 0, n - 1);
}

struct Plane {
    const float* p;
    int W, H, C;
    float at(int x, int y, int c) const {
        x = clampi(x, 0, W - 1);
        y = clampi(y, 0, H - 1);
        return p[((size_t)y * W + x) * C + c];
    }
};

// GL LINEAR filtering with CLAMP_TO_EDGE at arbitrary uv (GL 4.6 spec 8.14.2).
template <class R>
R bilinear(const Plane& t, R u, R v, int c) {
    const R fu = u * R(t.W) - R(0.5), fv = v * R(t.H) - R(0.5);
    const R i0f = std::floor(fu), j0f = std::floor(fv);
    const R a = fu - i0f, b = fv - j0f;
    const int ia = texelIndex(i0f, t.W), ib = texelIndex(i0f + R(1), t.W);
    const int ja = texelIndex(j0f, t.H), jb = texelIndex(j0f + R(1), t.H);
    const R t00 = R(t.at(ia, ja, c)), t10 = R(t.at(ib, ja, c));
    const R t01 = R(t.at(ia, jb, c)), t11 = R(t.at(ib, jb, c));
    return (R(1) - a) * (R(1) - b) * t00 + a * (R(1) - b) * t10 + (R(1) - a) * b * t01 + a * b * t11;
}
// GL NEAREST with CLAMP_TO_EDGE.
template <class R>
R nearest(const Plane& t, R u, R v, int c) {
    const int i = texelIndex(std::floor(u * R(t.W)), t.W), j = texelIndex(std::floor(v * R(t.H)), t.H);
    return R(t.at(i, j, c));
}

inline float halfToFloat(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu, bits;
    if (exp == 0) {
        if (mant == 0) bits = sign;
        else {
            int e = -1;
            do { ++e; mant <<= 1; } while (!(mant & 0x400u));
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3FFu) << 13);
        }
    } else if (exp == 31) bits = sign | 0x7F800000u | (mant << 13);
    else bits = sign | ((exp + 127 - 15) << 23) | (mant << 13);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// FP32 -> IEEE binary16, round to nearest, ties to even (the conversion a
// 16F render target applies to a fragment's output); overflow to infinity.
inline uint16_t floatToHalfRN(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) return sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u);  // NaN / infinity
    const int e = (int)(ax >> 23) - 127 + 15;                                    // biased binary16 exponent
    if (e >= 31) return sign | 0x7C00u;
    if (e <= 0) {  // subnormal or zero: value = ax's significand >> (14 - e) in units of 2^-24
        if (e < -10) return sign;
        const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
        const int shift = 14 - e;  // 14 .. 24
        uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return sign | (uint16_t)q;
    }
    uint32_t q = ((uint32_t)e << 10) | ((ax >> 13) & 0x3FFu);
    const uint32_t rem = ax & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;  // a carry into the exponent is still correct
    return sign | (uint16_t)q;
}

// samplerCubeArray NEAREST lookup (GL 4.6 spec table 8.19, major axis ties
// resolved x before y before z), CLAMP_TO_EDGE within the face: the texel
// offset inside the clamped atlas, returned with that atlas.
template <class R>
size_t cubeTexel(const sgi_cube_atlases& cubes, int& atlas, R layerCoord, V3<R> d) {
    atlas = clampi(atlas, 0, cubes.num_atlases - 1);  // out-of-range entries read atlas 0
    const R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int face;
    R sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= R(0)) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= R(0)) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= R(0)) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    }
    const R s = R(0.5) * (sc / ma + R(1)), t = R(0.5) * (tc / ma + R(1));
    const int res = cubes.resolution;
    const int i = texelIndex(std::floor(s * R(res)), res);
    const int j = texelIndex(std::floor(t * R(res)), res);
    const int layer = clampi((int)std::floor(layerCoord + R(0.5)), 0, cubes.layers_per_atlas - 1);
    return (((size_t)layer * 6 + face) * res + j) * res + i;
}

template <class R>
R cubeFetch(const sgi_cube_atlases& cubes, int atlas, R layerCoord, V3<R> d) {
    const size_t off = cubeTexel(cubes, atlas, layerCoord, d);
    if (cubes.format == SGI_TEXEL_F16) return R(halfToFloat(((const uint16_t*)cubes.atlas[atlas])[off]));
    // GL_DEPTH_COMPONENT16: code k reads as k / (2^16 - 1), rounded to FP32
    if (cubes.format == SGI_TEXEL_UNORM16) return R(float(((const uint16_t*)cubes.atlas[atlas])[off]) / 65535.0f);
    return R(((const float*)cubes.atlas[atlas])[off]);
}

// sampler2DArrayShadow, LINEAR + COMPARE_REF_TO_TEXTURE / LEQUAL, CLAMP_TO_EDGE
// (StratusRendererBackend.cpp:348-352): bilinear blend of the 4 comparisons.
template <class R>
R shadowCompare(const float* csm, int res, R s, R t, R layerCoord, R ref) {
    const int layer = clampi((int)std::floor(layerCoord + R(0.5)), 0, SGI_NUM_CASCADES - 1);
    const float* base = csm + (size_t)layer * res * res;
    const R fu = s * R(res) - R(0.5), fv = t * R(res) - R(0.5);
    const R i0f = std::floor(fu), j0f = std::floor(fv);
    const R a = fu - i0f, b = fv - j0f;
    const int ia = texelIndex(i0f, res), ib = texelIndex(i0f + R(1), res);
    const int ja = texelIndex(j0f, res), jb = texelIndex(j0f + R(1), res);
    auto cmp = [&](int i, int j) -> R {
        const R d = R(base[(size_t)j * res + i]);
        return ref <= d ? R(1) : R(0);
    };
    return (R(1) - a) * (R(1) - b) * cmp(ia, ja) + a * (R(1) - b) * cmp(ib, ja) +
           (R(1) - a) * b * cmp(ia, jb) + a * b * cmp(ib, jb);
}

// mat4 * vec4 as GLSL evaluates it: m[0]*v.x + m[1]*v.y + m[2]*v.z + m[3]*v.w.
template <class R>
void mulMat4(const float* m, R x, R y, R z, R w, R out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = R(m[r]) * x + R(m[4 + r]) * y + R(m[8 + r]) * z + R(m[12 + r]) * w;
}

// common.glsl:143-155
template <class R>
V3<R> worldPositionFromDepth(R u, R v, R depth, const float* invPV) {
    const R z = depth * R(2) - R(1);
    R w[4];
    mulMat4<R>(invPV, u * R(2) - R(1), v * R(2) - R(1), z, R(1), w);
    return v3(w[0] / w[3], w[1] / w[3], w[2] / w[3]);
}

template <class R> inline V3<R> ld3(const float* p) { return v3(R(p[0]), R(p[1]), R(p[2])); }

// ---- BRDF: pbr2.glsl ---------------------------------------------------------
template <class R>
struct Settings {
    R minRoughness;
    bool perceptual;
    R fogDensity;
    V3<R> fogColor;
};

// pbr2.glsl:51-55
template <class R> R NDF_GGX(R NdotH, R roughness) {
    const R r2 = roughness * roughness;
    const R f = (NdotH * r2 - NdotH) * NdotH + R(1);
    return r2 / gmax(PI_<R>() * f * f, EPS_<R>());
}
// pbr2.glsl:59-64
template <class R> R Visibility_G_SmithGGX(R NdotV, R NdotL, R roughness) {
    const R a2 = roughness * roughness;
    const R ggxv = NdotL * std::sqrt((NdotV - a2 * NdotV) * NdotV + a2);
    const R ggxl = NdotV * std::sqrt((NdotL - a2 * NdotL) * NdotL + a2);
    return R(0.5) / gmax(ggxv + ggxl, EPS_<R>());
}
// pbr2.glsl:73-76
template <class R> V3<R> Fr_Fresnel_Schlick(R u, V3<R> f0) {
    const R f = pow5(R(1) - u);
    return v3s(f) + f0 * (R(1) - f);
}
// pbr2.glsl:85-88
template <class R> R Fd_Fresnel_Schlick(R u, R f0, R f90) {
    const R f = pow5(R(1) - u);
    return f0 + (f90 - f0) * f;
}
// pbr2.glsl:91-96
template <class R> R Fd_Burley(R NdotV, R NdotL, R LdotH, R roughness) {
    const R f90 = R(0.5) + R(2) * roughness * LdotH * LdotH;
    const R lightScatter = Fd_Fresnel_Schlick(NdotL, R(1), f90);
    const R viewScatter = Fd_Fresnel_Schlick(NdotV, R(1), f90);
    return lightScatter * viewScatter * (R(1) / PI_<R>());
}
// pbr2.glsl:139-144
template <class R> R RemapRoughness(R roughness, const Settings<R>& st) {
    if (st.perceptual) roughness = roughness * roughness;
    return gmax(gmax(R(0), st.minRoughness), roughness);
}
// pbr2.glsl:146-168
template <class R>
void BRDF_Base(V3<R> baseColor, V3<R> baseReflectance, R roughness, R metallic, V3<R> diffuseDivisor,
               V3<R> specularMultiplier, const Settings<R>& st, R& remapped, V3<R>& diffuseColor, V3<R>& f0) {
    remapped = RemapRoughness(roughness, st);
    diffuseColor = (R(1) - metallic) * baseColor;
    diffuseColor = diffuseColor / (diffuseDivisor + EPS_<R>());
    f0 = R(0.16) * baseReflectance * baseReflectance * (R(1) - metallic) + baseColor * metallic;
    f0 = f0 * specularMultiplier;
}
// pbr2.glsl:170-188
template <class R>
void BRDF_DotProducts(V3<R> viewDir, V3<R> lightDir, V3<R> normal, R& NdotV, R& NdotL, R& NdotH, R& LdotH) {
    const V3<R> V = viewDir;
    const V3<R> L = normalize(lightDir);
    const V3<R> H = normalize(V + L);
    const V3<R> N = normal;
    NdotV = saturate(dot(N, V));
    NdotL = saturate(dot(N, L));
    NdotH = saturate(dot(N, H));
    LdotH = saturate(dot(L, H));
}
// pbr2.glsl:102-115
template <class R>
V3<R> singleScatteringBRDF_Specular(R NdotV, R NdotL, R NdotH, R LdotH, R a, V3<R> f0) {
    const R NDF = NDF_GGX(NdotH, a);
    const V3<R> F = Fr_Fresnel_Schlick(LdotH, f0);
    const R Vis = Visibility_G_SmithGGX(NdotV, NdotL, a);
    return (NDF * Vis) * F;
}
// pbr2.glsl:194-250
template <class R>
V3<R> BRDF_Burley(V3<R> lightDir, V3<R> viewDir, V3<R> normal, V3<R> baseColor, V3<R> baseReflectance, R roughness,
                  R metallic, V3<R> diffuseDivisor, V3<R> specularMultiplier, const Settings<R>& st) {
    R a;
    V3<R> diffuseColor, f0;
    BRDF_Base(baseColor, baseReflectance, roughness, metallic, diffuseDivisor, specularMultiplier, st, a, diffuseColor, f0);
    R NdotV, NdotL, NdotH, LdotH;
    BRDF_DotProducts(viewDir, lightDir, normal, NdotV, NdotL, NdotH, LdotH);
    const V3<R> Fr = singleScatteringBRDF_Specular(NdotV, NdotL, NdotH, LdotH, a, f0);
    const V3<R> Fd = diffuseColor * Fd_Burley(NdotV, NdotL, LdotH, a);
    return (Fd + Fr) * NdotL;
}
// pbr2.glsl:382-452
template <class R>
void BRDF_Lambert_SeparateDiffuseSpecular(V3<R> specDir, V3<R> diffDir, V3<R> viewDir, V3<R> normal, V3<R> baseColor,
                                          V3<R> baseReflectance, R roughness, R metallic, V3<R> diffuseDivisor,
                                          V3<R> specularMultiplier, const Settings<R>& st, V3<R>& specular,
                                          V3<R>& diffuse) {
    R a;
    V3<R> diffuseColor, f0;
    BRDF_Base(baseColor, baseReflectance, roughness, metallic, diffuseDivisor, specularMultiplier, st, a, diffuseColor, f0);
    R NdotV, NdotL, NdotH, LdotH;
    BRDF_DotProducts(viewDir, specDir, normal, NdotV, NdotL, NdotH, LdotH);
    const V3<R> Fr = singleScatteringBRDF_Specular(NdotV, NdotL, NdotH, LdotH, a, f0) * NdotL;
    BRDF_DotProducts(viewDir, diffDir, normal, NdotV, NdotL, NdotH, LdotH);
    const V3<R> Fd = diffuseColor * (R(1) / PI_<R>()) * NdotL;  // singleScatteringBRDF_Diffuse_Lambert
    specular = Fr;
    diffuse = Fd;
}
// fog.glsl:38-42
template <class R> V3<R> applyFog(V3<R> color, R dist, R intensity, const Settings<R>& st) {
    const R f = rexp(-st.fogDensity * dist);
    const V3<R> fc = st.fogColor * intensity;
    return color * f + fc * (R(1) - f);
}
// pbr2.glsl:454-478
template <class R>
V3<R> calculateLighting_Burley(V3<R> lightColor, V3<R> lightDir, V3<R> viewDir, V3<R> normal, V3<R> baseColor,
                               R viewDist, R fogIntensity, R roughness, R metallic, R ambientOcclusion,
                               R shadowFactor, V3<R> baseReflectance, R attenuationFactor, R ambientIntensity,
                               V3<R> diffuseDivisor, V3<R> specularMultiplier, const Settings<R>& st) {
    const V3<R> brdf = BRDF_Burley(lightDir, viewDir, normal, baseColor, baseReflectance, roughness, metallic,
                                   diffuseDivisor, specularMultiplier, st);
    const V3<R> ambient = brdf * ambientOcclusion * lightColor * ambientIntensity;
    const V3<R> finalBrightness = brdf * lightColor;
    return attenuationFactor * (ambient + shadowFactor * applyFog(finalBrightness, viewDist, fogIntensity, st));
}
// pbr2.glsl:541-574
template <class R>
V3<R> calculateLighting_Lambert_SeparateDiffuseSpecular(V3<R> lightColor, V3<R> specDir, V3<R> diffDir,
                                                        V3<R> viewDir, V3<R> normal, V3<R> baseColor, R viewDist,
                                                        R fogIntensity, R roughness, R metallic,
                                                        R ambientOcclusion, R shadowFactor, V3<R> baseReflectance,
                                                        R specAtt, R diffAtt, R ambientIntensity,
                                                        V3<R> diffuseDivisor, V3<R> specularMultiplier,
                                                        const Settings<R>& st) {
    V3<R> s, d;
    BRDF_Lambert_SeparateDiffuseSpecular(specDir, diffDir, viewDir, normal, baseColor, baseReflectance, roughness,
                                         metallic, diffuseDivisor, specularMultiplier, st, s, d);
    s = specAtt * s;
    d = diffAtt * d;
    const V3<R> brdf = s + d;
    const V3<R> ambient = brdf * ambientOcclusion * lightColor * ambientIntensity;
    const V3<R> finalBrightness = brdf * lightColor;
    return ambient + shadowFactor * applyFog(finalBrightness, viewDist, fogIntensity, st);
}
// pbr.glsl:260-267
template <class R> R vplDiffuseAttenuation(V3<R> lightDir, R lightRadius) {
    const R lightDist = length(lightDir);
    const R ratio = gmin(lightDist / lightRadius, R(1));
    const R minDist = mixv(R(10), R(1), rexp(R(-3) * ratio)) * lightRadius;
    return R(1) / (minDist + R(1) * lightDist * lightDist);
}
// pbr2.glsl:613-657
template <class R>
V3<R> calculateVirtualPointLighting2(V3<R> fragPos, V3<R> baseColor, V3<R> normal, V3<R> viewDir, V3<R> specPos,
                                     V3<R> diffPos, V3<R> lightColor, R viewDist, R lightRadius, R roughness,
                                     R metallic, R ambientOcclusion, R shadowFactor, V3<R> baseReflectance,
                                     const Settings<R>& st) {
    const V3<R> specDir = specPos - fragPos;
    const V3<R> diffDir = diffPos - fragPos;
    const R adjusted = R(1) - shadowFactor;
    return calculateLighting_Lambert_SeparateDiffuseSpecular(
        lightColor, specDir, diffDir, viewDir, normal, baseColor, viewDist, R(0), roughness, metallic,
        ambientOcclusion, adjusted, baseReflectance, vplDiffuseAttenuation(specDir, lightRadius),
        vplDiffuseAttenuation(diffDir, lightRadius), R(0), baseColor, R(1) / (baseColor + EPS_<R>()), st);
}

// pbr.glsl:116-136
template <class R>
R calculateShadowValue1Sample(const sgi_cube_atlases& cubes, int atlas, int layer, R farPlane, V3<R> fragPos,
                              V3<R> lightPos, R lightNormalDotProduct, R minBias) {
    const V3<R> fragDir = fragPos - lightPos;
    const R currentDepth = length(fragDir);
    const R bias = currentDepth * gmax(minBias, saturate(lightNormalDotProduct) * R(0.03));
    R depth = cubeFetch(cubes, atlas, R(layer), fragDir);
    depth = depth * farPlane;
    return (currentDepth - bias) > depth ? R(1) : R(0);
}

// pbr.glsl:158-227
template <class R>
R calculateInfiniteShadowValue(const sgi_csm& csm, V3<R> sunDir, V3<R> fragPos, V3<R> cascadeBlends, V3<R> normal,
                               bool useDepthBias) {
    R bias = R(csm.depthBias) / (R(csm.zfar) - R(csm.znear));
    if (!useDepthBias) bias = R(0);
    V3<R> position = fragPos + normal * (R(1) - saturate(dot(normal, sunDir))) * R(1);
    R cc[4][3];
    for (int i = 0; i < 4; ++i) {
        R c[4];
        mulMat4<R>(csm.cascadeProjViews[i], position.x, position.y, position.z, R(1), c);
        for (int k = 0; k < 3; ++k) cc[i][k] = (c[k] / c[3]) * R(0.5) + R(0.5);
    }
    const bool beyond2 = cascadeBlends.y >= R(0);
    const bool beyond3 = cascadeBlends.z >= R(0);
    const int index1 = beyond2 ? 2 : 0;
    const int index2 = beyond3 ? 3 : 1;
    const R depth1 = cc[index1][2];
    R depth2 = cc[index2][2];
    depth2 = beyond3 ? saturate(depth2) : depth2;
    const R weight = beyond2 ? saturate(cascadeBlends.y) - saturate(cascadeBlends.z) : R(1) - saturate(cascadeBlends.x);
    const R wh = R(1) / R(csm.resolution);
    R light1 = R(0), light2 = R(0), samples = R(0);
    for (R y = R(-1); y <= R(1); y += R(1)) {
        for (R x = R(-1); x <= R(1); x += R(1)) {
            // sampleShadowTexture (pbr.glsl:139-147): coords.w = depth - bias; coords.xy += offset
            light1 += shadowCompare<R>(csm.depth, csm.resolution, cc[index1][0] + x * wh, cc[index1][1] + y * wh,
                                       R(index1), depth1 - bias);
            light2 += shadowCompare<R>(csm.depth, csm.resolution, cc[index2][0] + x * wh, cc[index2][1] + y * wh,
                                       R(index2), depth2 - bias);
            samples += R(1);
        }
    }
    return mixv(light2, light1, weight) * (R(1) / samples);
}

template <class R> V3<R> cascadeBlendsAt(const sgi_csm& csm, V3<R> p) {
    R b[3];
    for (int i = 0; i < 3; ++i) {
        const float* pl = csm.cascadePlanes[i];
        b[i] = R(pl[0]) * p.x + R(pl[1]) * p.y + R(pl[2]) * p.z + R(pl[3]) * R(1);
    }
    return v3(b[0], b[1], b[2]);
}

template <class R> Settings<R> settingsOf(const sgi_frame& f) {
    Settings<R> s;
    s.minRoughness = R(f.minRoughness);
    s.perceptual = f.usePerceptualRoughness != 0;
    s.fogDensity = R(f.fogDensity);
    s.fogColor = ld3<R>(f.fogColor);
    return s;
}

struct GB {
    Plane depth, normal, albedo, refl, rma, velocity, id, ssao;
    bool hasVelocity, hasId, hasSsao;
};
GB wrap(const sgi_gbuffer& g) {
    GB b;
    const int W = g.width, H = g.height;
    b.depth = {g.depth, W, H, 1};
    b.normal = {g.normal, W, H, 3};
    b.albedo = {g.albedo, W, H, 4};
    b.refl = {g.baseReflectivity, W, H, 2};
    b.rma = {g.roughnessMetallicAmbient, W, H, 3};
    b.velocity = {g.velocity, W, H, 2};
    b.id = {g.id, W, H, 1};
    b.ssao = {g.ssao, W, H, 1};
    b.hasVelocity = g.velocity != nullptr;
    b.hasId = g.id != nullptr;
    b.hasSsao = g.ssao != nullptr;
    return b;
}
template <class R> V3<R> texNormal(const Plane& n, int x, int y) {
    return normalize(v3(R(n.at(x, y, 0)), R(n.at(x, y, 1)), R(n.at(x, y, 2))) * R(2) - v3s(R(1)));
}

// ---- gather: vpl_pbr_gi.fs:81-207 ---------------------------------------------
enum GatherMode { RANDOM = 0, TILES = 1, CONVERGED = 2, SAMPLES = 3 };

template <class R>
void gatherPixel(const sgi_frame& fr, const GB& g, int x, int y, const sgi_vpl* vpls, int numVisible,
                 const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases& cubes, const int32_t* tileList,
                 int mode, int samples, int rejectionLimit, R outColor[3], R outRes[4]) {
    const Settings<R> st = settingsOf<R>(fr);
    const int W = g.depth.W, H = g.depth.H;
    const R u = R((float)((float)x + 0.5f) / (float)W), v = R((float)((float)y + 0.5f) / (float)H);
    const R depth = R(g.depth.at(x, y, 0));
    const V3<R> fragPos = worldPositionFromDepth<R>(u, v, depth, fr.invProjectionView);
    const V3<R> viewPos = ld3<R>(fr.viewPosition);
    const V3<R> viewDir = normalize(viewPos - fragPos);
    const R distToCamera = length(viewPos - fragPos);
    V3<R> baseColor = v3(R(g.albedo.at(x, y, 0)), R(g.albedo.at(x, y, 1)), R(g.albedo.at(x, y, 2)));
    baseColor = v3(gmax(baseColor.x, R(0.01)), gmax(baseColor.y, R(0.01)), gmax(baseColor.z, R(0.01)));
    const V3<R> normal = texNormal<R>(g.normal, x, y);
    R roughness = R(g.rma.at(x, y, 0));
    roughness = gmax(R(0.5), roughness);
    const R metallic = R(g.rma.at(x, y, 1));
    const R ambientOcclusion = clampv(g.hasSsao ? R(g.ssao.at(x, y, 0)) : R(1), R(0.35), R(1));
    const R ambient = R(g.rma.at(x, y, 2));
    const V3<R> baseReflectivity = v3s(R(g.refl.at(x, y, 0)));
    const R roughnessWeight = R(1) - roughness;
    const R minGiOcc = R(fr.minGiOcclusionFactor);

    V3<R> vplColor = v3s(R(0));
    R validSamples = R(0);

    // per-VPL body shared by the three modes; returns false when rejected
    auto evalVpl = [&](int lightIndex, bool allowReject) -> bool {
        const sgi_atlas_entry entry = shadowIndices[lightIndex];
        const sgi_vpl& L = vpls[lightIndex];
        const V3<R> lightPosition = ld3<R>(L.position);
        const V3<R> specularLightPosition = ld3<R>(L.specularPosition);
        const V3<R> lightMinusFrag = lightPosition - fragPos;
        const R lightRadius = R(L.radius);
        const R distance = length(lightMinusFrag);
        if (allowReject) {
            const R sideCheck = dot(normal, normalize(lightMinusFrag));
            if (sideCheck < R(0) || distance > lightRadius) return false;
        }
        const R distanceRatio = clampv((R(2) * distance) / lightRadius, R(0), R(1));
        const R distAttenuation = distanceRatio;
        const V3<R> lightColor = ld3<R>(L.color);
        R shadowFactor = distToCamera < R(500)
                             ? calculateShadowValue1Sample<R>(cubes, entry.index, entry.layer, R(L.farPlane), fragPos,
                                                              lightPosition, dot(lightPosition - fragPos, normal),
                                                              R(0.1))
                             : R(0);
        shadowFactor = gmin(shadowFactor, mixv(minGiOcc, R(1), distanceRatio));
        R reweightingFactor = R(1);
        if (shadowFactor > R(0)) reweightingFactor = (R(1) - distAttenuation) * minGiOcc + distAttenuation;
        validSamples += reweightingFactor;
        const V3<R> tmpColor =
            ambientOcclusion * calculateVirtualPointLighting2<R>(fragPos, baseColor, normal, viewDir,
                                                                 specularLightPosition, lightPosition, lightColor,
                                                                 distToCamera, lightRadius, roughness, metallic,
                                                                 ambient, R(0), baseReflectivity, st);
        vplColor = vplColor + (R(1) - shadowFactor) * tmpColor;
        return true;
    };

    if (mode == RANDOM) {
        float seedZ = fr.time;
        const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
        const R distRatioToCamera = gmin(R(1) - distToCamera / R(1000), R(1));
        const int maxSamplesPerPixel = (int)mixv(R(2), R(10), roughnessWeight);
        int samplesMax = maxSamplesPerPixel;
        samplesMax = std::max(1, (int)(R(samplesMax) * distRatioToCamera));
        const int sampleCount = samplesMax;
        const int maxRandomIndex = numVisible - 1;
        for (int i = 0, resamples = 0; i < sampleCount; i += 1) {
            const float rnd = random3(seedX, seedY, seedZ);
            seedZ += 10000.0f;
            const int lightIndex = (int)(R(maxRandomIndex) * R(rnd));
            if (!evalVpl(lightIndex, resamples < 5)) {
                ++resamples;
                --i;
                continue;
            }
        }
    } else if (mode == TILES) {
        const int n = tileList[0];
        for (int k = 0; k < n; ++k) evalVpl(tileList[1 + k], true);
    } else if (mode == SAMPLES) {
        // the fixed-count estimator: slot int(count * rand) over all visible slots,
        // the reference's hash and seed step.  rejectionLimit >= 0: a failing draw
        // is drawn again, and after that many rejections draws are taken
        // unconditionally (the reference's rule is 5); rejectionLimit < 0:
        // exactly `samples` draws, failing ones skipped (no light, no weight)
        float seedZ = fr.time;
        const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
        for (int i = 0, resamples = 0; numVisible > 0 && i < samples; i += 1) {
            const float rnd = random3(seedX, seedY, seedZ);
            seedZ += 10000.0f;
            int lightIndex = (int)((float)numVisible * rnd);
            if (lightIndex > numVisible - 1) lightIndex = numVisible - 1;
            if (rejectionLimit < 0) {
                evalVpl(lightIndex, true);
                continue;
            }
            if (!evalVpl(lightIndex, resamples < rejectionLimit)) {
                ++resamples;
                --i;
                continue;
            }
        }
    } else {
        for (int s = 0; s < numVisible; ++s) evalVpl(s, true);
    }
    validSamples = gmax(validSamples, R(1));
    outColor[0] = baseColor.x + EPS_<R>();
    outColor[1] = baseColor.y + EPS_<R>();
    outColor[2] = baseColor.z + EPS_<R>();
    const V3<R> bounded = vmin(vplColor, R(65504));  // boundHDR, common.glsl:89-92
    outRes[0] = bounded.x;
    outRes[1] = bounded.y;
    outRes[2] = bounded.z;
    outRes[3] = validSamples;
}

// ---- direct light: pbr.fs:82-145 (INFINITE_LIGHTING_ENABLED, no point lights) ----
template <class R>
void directPixel(const sgi_frame& fr, const sgi_csm& csm, const GB& g, int x, int y, R out[3]) {
    const Settings<R> st = settingsOf<R>(fr);
    const int W = g.depth.W, H = g.depth.H;
    const R u = R((float)((float)x + 0.5f) / (float)W), v = R((float)((float)y + 0.5f) / (float)H);
    const R depth = R(g.depth.at(x, y, 0));
    const V3<R> fragPos = worldPositionFromDepth<R>(u, v, depth, fr.invProjectionView);
    const V3<R> viewMinusFrag = ld3<R>(fr.viewPosition) - fragPos;
    const V3<R> viewDir = normalize(viewMinusFrag);
    const R viewDist = length(viewMinusFrag);
    const V3<R> baseColor = v3(R(g.albedo.at(x, y, 0)), R(g.albedo.at(x, y, 1)), R(g.albedo.at(x, y, 2)));
    const R albedoA = R(g.albedo.at(x, y, 3));
    const V3<R> normal = texNormal<R>(g.normal, x, y);
    const R roughness = R(g.rma.at(x, y, 0));
    const R metallic = R(g.rma.at(x, y, 1));
    const R emissiveB = R(g.rma.at(x, y, 2));
    const R ambient = g.hasSsao ? R(g.ssao.at(x, y, 0)) : R(1);
    const R reflR = R(g.refl.at(x, y, 0)), reflG = R(g.refl.at(x, y, 1));
    const V3<R> emissive = v3(albedoA, reflG, emissiveB);
    V3<R> color = v3s(R(0));
    const V3<R> lightDir = ld3<R>(fr.infiniteLightDirection);
    const V3<R> blends = cascadeBlendsAt<R>(csm, fragPos);
    const R shadowFactor = calculateInfiniteShadowValue<R>(csm, lightDir, fragPos, blends, normal, true);
    // calculateDirectionalLighting (pbr2.glsl:576-591) with shadowFactor = 1 - s
    const R sf = R(1) - shadowFactor;
    color = color + calculateLighting_Burley<R>(ld3<R>(fr.infiniteLightColor), lightDir, viewDir, normal, baseColor,
                                                viewDist, R(0), roughness, metallic, ambient, R(1) - sf,
                                                v3s(reflR), R(1), R(0), v3s(R(1)), v3s(R(1)), st);
    const V3<R> res = vmin(color + emissive * R(fr.emissionStrength), R(65504));
    out[0] = res.x;
    out[1] = res.y;
    out[2] = res.z;
}

// ---- global cull: viscull_vpls.cs:47-120 -------------------------------------
template <class R>
int cullGlobal(const sgi_frame& fr, const sgi_csm& csm, const sgi_vpl* in, int n, sgi_vpl* out, int32_t* idx) {
    const V3<R> sun = ld3<R>(fr.infiniteLightDirection);
    int count = 0;
    for (int i = 0; i < n; ++i) {
        const V3<R> p = ld3<R>(in[i].position);
        const V3<R> blends = cascadeBlendsAt<R>(csm, p);
        const R shadowFactor = R(1) - calculateInfiniteShadowValue<R>(csm, sun, p, blends, sun, false);
        if (shadowFactor < R(1)) {
            if (count < SGI_MAX_TOTAL_VPLS_PER_FRAME) {
                if (out) out[count] = in[i];
                if (idx) idx[1 + count] = i;
            }
            ++count;
        }
    }
    if (count > SGI_MAX_TOTAL_VPLS_PER_FRAME) count = SGI_MAX_TOTAL_VPLS_PER_FRAME;
    if (idx) idx[0] = count;
    return count;
}

// ---- tile culling: vpl_tiled_deferred_culling_stage1.cs:50-90 / _stage2.cs:61-129 ----
template <class R>
void tileStage1(const sgi_frame& fr, const GB& g, int ts, int tx, int ty, R pos[3], R nrm[3]) {
    const int W = g.depth.W, H = g.depth.H;
    V3<R> sp = v3s(R(0)), sn = v3s(R(0));
    int cnt = 0;
    for (int ly = 0; ly < ts; ++ly)
        for (int lx = 0; lx < ts; ++lx) {
            const int x = tx * ts + lx, y = ty * ts + ly;
            if (x >= W || y >= H) continue;
            const R u = R((float)((float)x + 0.5f) / (float)W), v = R((float)((float)y + 0.5f) / (float)H);
            const V3<R> fp = worldPositionFromDepth<R>(u, v, R(g.depth.at(x, y, 0)), fr.invProjectionView);
            const V3<R> nn = texNormal<R>(g.normal, x, y);
            sp = sp + fp;
            sn = sn + nn;
            ++cnt;
        }
    sp = sp / R(cnt);
    sn = sn / R(cnt);
    sn = normalize(sn);
    pos[0] = sp.x; pos[1] = sp.y; pos[2] = sp.z;
    nrm[0] = sn.x; nrm[1] = sn.y; nrm[2] = sn.z;
}

template <class R>
int tileStage2(const R pos[3], const R nrm[3], const sgi_vpl* vpls, const int32_t* visible, int remap, int K,
               int32_t* outIdx) {
    const V3<R> fragPos = v3(pos[0], pos[1], pos[2]), normal = v3(nrm[0], nrm[1], nrm[2]);
    const int numVisible = visible[0];
    std::vector<int> idx(K);
    std::vector<R> dist(K);
    for (int i = 0; i < K; ++i) { idx[i] = i; dist[i] = R(3.402823466e+38); }
    int numVis = 0;
    for (int s = 0; s < numVisible; ++s) {
        const int lightIndex = remap ? visible[1 + s] : s;
        const V3<R> lightPosition = ld3<R>(vpls[s].position);
        const V3<R> lightMinusFrag = lightPosition - fragPos;
        const R sideCheck = dot(normal, lightMinusFrag);
        if (sideCheck < R(0)) continue;
        const R radius = R(vpls[s].radius);
        const R distance = length(lightMinusFrag);
        const R ratio = distance / radius;
        if (ratio > R(1)) continue;
        for (int ii = 0; ii < K; ++ii) {
            if (ratio < dist[ii]) {
                if (dist[ii] != R(3.402823466e+38)) {
                    for (int k = K - 2; k >= ii; --k) { idx[k + 1] = idx[k]; dist[k + 1] = dist[k]; }
                    idx[ii] = 0;
                    dist[ii] = R(0);
                }
                idx[ii] = lightIndex;
                dist[ii] = ratio;
                if (numVis < K) ++numVis;
                break;
            }
        }
    }
    outIdx[0] = numVis;
    for (int i = 0; i < numVis; ++i) outIdx[1 + i] = idx[i];
    return numVis;
}

// ---- denoiser: vpl_pbr_gi_denoise.fs:142-435 -----------------------------------
template <class R>
struct Img {
    std::vector<R> d;
    int W, H, C;
    Img(int w, int h, int c) : d((size_t)w * h * c, R(0)), W(w), H(h), C(c) {}
    R& at(int x, int y, int c) { return d[((size_t)y * W + x) * C + c]; }
    R get(int x, int y, int c) const {
        x = clampi(x, 0, W - 1);
        y = clampi(y, 0, H - 1);
        return d[((size_t)y * W + x) * C + c];
    }
};

template <class R>
void denoise(const sgi_frame& fr, const GB& cur, const GB& prev, int numVisible, const float* screen,
             const float* colorIn, const float* reservoirIn, const float* prevGi, const float* prevHistory,
             R* combined, R* giOut, R* resOut, R* histOut) {
    const int W = cur.depth.W, H = cur.depth.H;
    Img<R> res(W, H, 4), res2(W, H, 4);
    for (size_t i = 0; i < res.d.size(); ++i) res.d[i] = R(reservoirIn[i]);
    const Plane illum{colorIn, W, H, 3};
    const Plane scr{screen, W, H, 3};
    const Plane pGi{prevGi, W, H, 3};
    const Plane pHist{prevHistory, W, H, 1};
    for (int pass = 0; pass < 4; ++pass) {
        const int multiplier = (1 << pass) - 1;
        const bool merge = pass == 0, fin = pass == 3;
        forRows(0, H, [&](int y) {
            for (int x = 0; x < W; ++x) {
                const R u = R((float)((float)x + 0.5f) / (float)W), v = R((float)((float)y + 0.5f) / (float)H);
                const V3<R> centerIllum = v3(R(illum.at(x, y, 0)), R(illum.at(x, y, 1)), R(illum.at(x, y, 2)));
                const V3<R> centerNormal = texNormal<R>(cur.normal, x, y);
                const R centerDepth = R(cur.depth.at(x, y, 0));
                const R historyAccumIn = prevHistory ? R(pHist.at(x, y, 0)) : R(0);
                R historyAccum = historyAccumIn;
                R rf[4] = {R(0), R(0), R(0), R(0)};
                if (merge) {
                    // computeMergedReservoir (:192-273) with minmaxNearest = 0
                    const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
                    float seedZ = fr.time;
                    R cr[4] = {R(0), R(0), R(0), R(0)};
                    const R depthCutoff = R(0.1) * centerDepth;
                    R runningSum = R(0);
                    const R probabilisticWeight = R(1) / R(numVisible);
                    const V3<R> currNormal = texNormal<R>(cur.normal, x, y);
                    bool skip = dot(centerNormal, currNormal) < R(0.906);
                    if (!skip) {
                        const R currDepth = R(cur.depth.at(x, y, 0));
                        skip = std::fabs(currDepth - centerDepth) > depthCutoff;
                    }
                    if (!skip) {
                        const R rr[4] = {res.get(x, y, 0), res.get(x, y, 1), res.get(x, y, 2), res.get(x, y, 3)};
                        const float randUpdate = random3(seedX, seedY, seedZ);
                        seedZ += 10000.0f;
                        const R probability = rr[3] * probabilisticWeight;
                        cr[3] += probability;
                        if (R(randUpdate) < (probability / cr[3])) {
                            cr[0] += rr[0];
                            cr[1] += rr[1];
                            cr[2] += rr[2];
                            runningSum += rr[3];
                        }
                    }
                    cr[3] = runningSum;
                    for (int k = 0; k < 4; ++k) rf[k] = cr[k];
                } else {
                    for (int dx = -2; dx <= 2; ++dx)
                        for (int dy = -2; dy <= 2; ++dy) {
                            const int ox = dx + dx * multiplier, oy = dy + dy * multiplier;
                            const R rr[4] = {res.get(x + ox, y + oy, 0), res.get(x + ox, y + oy, 1),
                                             res.get(x + ox, y + oy, 2), res.get(x + ox, y + oy, 3)};
                            R filtered;
                            if (dx == 0 && dy == 0) filtered = R(1);
                            else {  // filterInput (:142-190)
                                const R currDepth = R(cur.depth.at(x + ox, y + oy, 0));
                                const R wz = rexp(-std::fabs(centerDepth - currDepth)) / R(1);
                                const V3<R> currNormal = texNormal<R>(cur.normal, x + ox, y + oy);
                                R wn = gmax(R(0), dot(centerNormal, currNormal));
                                wn = pow128(wn);
                                filtered = wn * wz * R(1);
                            }
                            for (int k = 0; k < 4; ++k) rf[k] += filtered * rr[k];
                        }
                }
                const R den = gmax(EPS_<R>(), rf[3]);
                const V3<R> shadowFactor = v3(rf[0] / den, rf[1] / den, rf[2] / den);
                const V3<R> gi = centerIllum;
                V3<R> illumAvg = gi;
                if (fin) {
                    R accumMultiplier = R(1);
                    const V3<R> currGi = shadowFactor;
                    const R vx = cur.hasVelocity ? R(cur.velocity.at(x, y, 0)) : R(0);
                    const R vy = cur.hasVelocity ? R(cur.velocity.at(x, y, 1)) : R(0);
                    const R pu = u - vx, pv = v - vy;
                    const R prevCenterDepth = bilinear<R>(prev.depth, pu, pv, 0);
                    const V3<R> prevCenterNormal =
                        normalize(v3(bilinear<R>(prev.normal, pu, pv, 0), bilinear<R>(prev.normal, pu, pv, 1),
                                     bilinear<R>(prev.normal, pu, pv, 2)) * R(2) - v3s(R(1)));
                    V3<R> prevGiV = v3s(R(0));
                    if (prevGi)
                        prevGiV = v3(bilinear<R>(pGi, pu, pv, 0), bilinear<R>(pGi, pu, pv, 1), bilinear<R>(pGi, pu, pv, 2));
                    const R currId = cur.hasId ? R(cur.id.at(x, y, 0)) : R(0);
                    const R prevId = prev.hasId ? nearest<R>(prev.id, pu, pv, 0) : R(0);
                    const R wn = gmax(R(0), dot(centerNormal, prevCenterNormal));
                    const R wz = rexp(-std::fabs(centerDepth - prevCenterDepth));
                    const R wid = currId != prevId ? R(0) : R(1);
                    const R similarity = wn * wz * wid;
                    if (similarity < R(0.99)) accumMultiplier = R(0);
                    historyAccum = gmin(R(1) + historyAccum * accumMultiplier, R(fr.framesPerSecond));
                    const R maxAccumulationFactor = R(1) / historyAccum;
                    illumAvg = mixv(prevGiV, currGi, maxAccumulationFactor);
                }
                const size_t p = (size_t)y * W + x;
                if (fin) {
                    const V3<R> sc = v3(R(scr.at(x, y, 0)), R(scr.at(x, y, 1)), R(scr.at(x, y, 2)));
                    const V3<R> comb = sc + gi * illumAvg;
                    combined[p * 3 + 0] = comb.x;
                    combined[p * 3 + 1] = comb.y;
                    combined[p * 3 + 2] = comb.z;
                    giOut[p * 3 + 0] = illumAvg.x;
                    giOut[p * 3 + 1] = illumAvg.y;
                    giOut[p * 3 + 2] = illumAvg.z;
                    resOut[p * 4 + 0] = shadowFactor.x;
                    resOut[p * 4 + 1] = shadowFactor.y;
                    resOut[p * 4 + 2] = shadowFactor.z;
                    resOut[p * 4 + 3] = R(1);
                    histOut[p] = historyAccum;
                }
                res2.at(x, y, 0) = shadowFactor.x;
                res2.at(x, y, 1) = shadowFactor.y;
                res2.at(x, y, 2) = shadowFactor.z;
                res2.at(x, y, 3) = R(1);
            }
        });
        std::swap(res.d, res2.d);
    }
}

}  // namespace

// ============================== C ABI ==========================================
// vpl_light_color.cs:41-76 (dispatched by PerformVirtualPointLightCullingStage2_,
// StratusRendererBackend.cpp:1703-1731), FP32, one VPL slot at a time:
//   color    = textureLod(diffuseCubeMaps[e.index], vec4(-L, e.layer), 0).rgb * sunColor   (:52)
//   magnitude = radius * textureLod(shadowCubeMaps[e.index], vec4(-L, e.layer), 0).r   (:53)
//   specularPos = vec4(position.xyz - 1.1 * magnitude * L, 1)                          (:71)
//   lightData[i].color = vec4(color / totalColors * intensity * colorMultiplier, 1)    (:73)
// with totalColors = 1 (:56; the 8-sample loop at :57-66 is commented out).  The
// RGB8 attachment reads as k / 255 (GL UNORM conversion).
void colorVpls(const sgi_frame& fr, int count, const sgi_atlas_entry* entries, const sgi_cube_atlases& diffuse,
               const sgi_cube_atlases& dist, float colorMultiplier, sgi_vpl* vpls) {
    const V3<float> L = ld3<float>(fr.infiniteLightDirection);
    const V3<float> d = v3(-L.x, -L.y, -L.z);
    const float sun[3] = {fr.infiniteLightColor[0], fr.infiniteLightColor[1], fr.infiniteLightColor[2]};
    const float Lv[3] = {L.x, L.y, L.z};
    for (int i = 0; i < count; ++i) {
        const sgi_atlas_entry e = entries[i];
        int a = e.index;
        const size_t off = cubeTexel<float>(diffuse, a, float(e.layer), d);
        const uint8_t* rgb = (const uint8_t*)diffuse.atlas[a] + off * 4;
        sgi_vpl& v = vpls[i];
        const float magnitude = v.radius * cubeFetch<float>(dist, e.index, float(e.layer), d);
        const float totalColors = 1.0f;
        for (int k = 0; k < 3; ++k) {
            const float color = (float)rgb[k] / 255.0f * sun[k];
            v.color[k] = color / totalColors * v.intensity * colorMultiplier;
            v.specularPosition[k] = v.position[k] - 1.1f * magnitude * Lv[k];
        }
        v.color[3] = 1.0f;
        v.specularPosition[3] = 1.0f;
    }
}

extern "C" {

// Worker threads of the row-parallel passes (cull_tiles, gather, direct_light,
// denoise); n <= 0 means every hardware thread.  Results do not depend on it.
void orc_set_threads(int n) {
    if (n <= 0) n = (int)std::max(1u, std::thread::hardware_concurrency());
    g_threads.store(n);
}
int orc_get_threads() { return g_threads.load(); }

// -- unit functions (for closed-form pins) --
uint32_t orc_hash1(uint32_t x) { return hash1(x); }
uint32_t orc_hash3(uint32_t x, uint32_t y, uint32_t z) { return hash3(x, y, z); }
float orc_random3(float x, float y, float z) { return random3(x, y, z); }

double orc_ndf_ggx(double NdotH, double a) { return NDF_GGX<double>(NdotH, a); }
float orc_ndf_ggx_f(float NdotH, float a) { return NDF_GGX<float>(NdotH, a); }
double orc_vis_smith(double NdotV, double NdotL, double a) { return Visibility_G_SmithGGX<double>(NdotV, NdotL, a); }
double orc_fresnel(double u, double f0) { return Fr_Fresnel_Schlick<double>(u, v3s(f0)).x; }
double orc_fd_burley(double NdotV, double NdotL, double LdotH, double a) { return Fd_Burley<double>(NdotV, NdotL, LdotH, a); }
double orc_vpl_atten(double dist, double radius) { return vplDiffuseAttenuation<double>(v3(dist, 0.0, 0.0), radius); }

// BRDF_Burley * lightColor at one point (direct light kernel body) in FP64.
void orc_brdf_burley(const double lightDir[3], const double viewDir[3], const double normal[3],
                     const double baseColor[3], double refl, double roughness, double metallic,
                     double minRoughness, int perceptual, double out[3]) {
    Settings<double> st{minRoughness, perceptual != 0, 0.0, v3s(0.0)};
    const V3<double> r = BRDF_Burley<double>(v3(lightDir[0], lightDir[1], lightDir[2]), v3(viewDir[0], viewDir[1], viewDir[2]),
                                            v3(normal[0], normal[1], normal[2]), v3(baseColor[0], baseColor[1], baseColor[2]),
                                            v3s(refl), roughness, metallic, v3s(1.0), v3s(1.0), st);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// calculateVirtualPointLighting2 in FP64 (VPL BRDF with demodulated albedo).
void orc_vpl_lighting(const double fragPos[3], const double baseColor[3], const double normal[3], const double viewDir[3],
                      const double specPos[3], const double diffPos[3], const double lightColor[3], double radius,
                      double roughness, double metallic, double refl, double minRoughness, int perceptual,
                      double out[3]) {
    Settings<double> st{minRoughness, perceptual != 0, 0.0, v3s(0.0)};
    const V3<double> r = calculateVirtualPointLighting2<double>(
        v3(fragPos[0], fragPos[1], fragPos[2]), v3(baseColor[0], baseColor[1], baseColor[2]),
        v3(normal[0], normal[1], normal[2]), v3(viewDir[0], viewDir[1], viewDir[2]), v3(specPos[0], specPos[1], specPos[2]),
        v3(diffPos[0], diffPos[1], diffPos[2]), v3(lightColor[0], lightColor[1], lightColor[2]), 0.0, radius, roughness,
        metallic, 0.0, 0.0, v3s(refl), st);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

double orc_cube_fetch(const sgi_cube_atlases* cubes, int atlas, int layer, double dx, double dy, double dz) {
    return cubeFetch<double>(*cubes, atlas, (double)layer, v3(dx, dy, dz));
}
float orc_shadow1(const sgi_cube_atlases* cubes, int atlas, int layer, float farPlane, const float fragPos[3],
                  const float lightPos[3], float lndp, float minBias) {
    return calculateShadowValue1Sample<float>(*cubes, atlas, layer, farPlane, ld3<float>(fragPos), ld3<float>(lightPos),
                                              lndp, minBias);
}
double orc_shadow_compare(const float* csm, int res, double s, double t, int layer, double ref) {
    return shadowCompare<double>(csm, res, s, t, (double)layer, ref);
}
float orc_infinite_shadow(const sgi_csm* csm, const float sunDir[3], const float pos[3], const float normal[3], int useBias) {
    const V3<float> p = ld3<float>(pos);
    return calculateInfiniteShadowValue<float>(*csm, ld3<float>(sunDir), p, cascadeBlendsAt<float>(*csm, p),
                                               ld3<float>(normal), useBias != 0);
}
void orc_world_pos(float u, float v, float depth, const float invPV[16], float out[3]) {
    const V3<float> p = worldPositionFromDepth<float>(u, v, depth, invPV);
    out[0] = p.x; out[1] = p.y; out[2] = p.z;
}

// -- pass-level functions --
// The reads the GLSL passes make of the reference's G-buffer attachments
// (StratusRendererBackend.cpp:384-424): an RGB8 / RGBA8 / RG8 UNORM code k is
// k / 255 (GL 4.6 section 2.3.5.1, c / (2^8 - 1)); RG16F velocity widens to
// FP32.  Texel layouts as sgi_gbuffer_packed states them.
void orc_gbuffer_unpack(const sgi_gbuffer_packed* g, float* normal, float* albedo, float* refl, float* rma,
                        float* vel) {
    const long n = (long)g->width * g->height;
    const auto code = [](uint32_t v, int b) { return (float)((v >> (8 * b)) & 0xFFu) / 255.0f; };
    for (long p = 0; p < n; ++p) {
        for (int k = 0; k < 3; ++k) {
            normal[3 * p + k] = code(g->normals[p], k);
            rma[3 * p + k] = code(g->roughnessMetallicAmbient[p], k);
        }
        for (int k = 0; k < 4; ++k) albedo[4 * p + k] = code(g->albedo[p], k);
        for (int k = 0; k < 2; ++k) refl[2 * p + k] = code(g->baseReflectivity[p], k);
        if (vel && g->velocity)
            for (int k = 0; k < 2; ++k) vel[2 * p + k] = halfToFloat(g->velocity[2 * p + k]);
    }
}

// The geometry pass's attachment writes (pbr_geometry_pass.fs:127-137 outputs
// into the formats of StratusRendererBackend.cpp:384-424): GL's conversion of
// a fragment output into a UNORM8 channel, clamp to [0, 1] then
// floor(v * 255 + 0.5) in FP32 (NaN clamps to 0), and into RG16F, the
// nearest binary16 with ties to even.
void orc_gbuffer_pack(const sgi_gbuffer* planes, uint32_t* n8, uint32_t* a8, uint16_t* f8, uint32_t* m8,
                      uint16_t* v16) {
    const long n = (long)planes->width * planes->height;
    const auto code = [](float v) -> uint32_t {
        const float c = v >= 0.0f ? (v <= 1.0f ? v : 1.0f) : 0.0f;
        const float s = c * 255.0f;
        return (uint32_t)std::floor(s + 0.5f);
    };
    for (long p = 0; p < n; ++p) {
        uint32_t nv = 0, mv = 0, av = 0;
        for (int k = 0; k < 3; ++k) {
            nv |= code(planes->normal[3 * p + k]) << (8 * k);
            mv |= code(planes->roughnessMetallicAmbient[3 * p + k]) << (8 * k);
        }
        for (int k = 0; k < 4; ++k) av |= code(planes->albedo[4 * p + k]) << (8 * k);
        n8[p] = nv;
        a8[p] = av;
        m8[p] = mv;
        f8[p] = (uint16_t)(code(planes->baseReflectivity[2 * p]) | (code(planes->baseReflectivity[2 * p + 1]) << 8));
        if (v16 && planes->velocity)
            for (int k = 0; k < 2; ++k) v16[2 * p + k] = floatToHalfRN(planes->velocity[2 * p + k]);
    }
}

// taa.fs:12-100 with aa_common.glsl's tonemap c / (c + 1) and inverse
// c / (1 - c): velocity texel (NEAREST), screen at texel centres (the 4
// textureOffset taps clamp to the edge), prevScreen LINEAR at uv - velocity,
// clamp(prev, tonemap(min), tonemap(max)), mix(cur, prev, 0.9) = cur (1 - 0.9)
// + prev 0.9.
void orc_taa(const float* screen, const float* prev, const float* vel, int W, int H, float* out) {
    const Plane S{screen, W, H, 3}, P{prev, W, H, 3};
    const auto tm = [](float c) { return c / (c + 1.0f); };
    const auto itm = [](float c) { return c / (1.0f - c); };
    const float a = 0.9f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t p = (size_t)y * W + x;
            const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
            const float vu = vel ? vel[2 * p] : 0.0f, vv = vel ? vel[2 * p + 1] : 0.0f;
            for (int c = 0; c < 3; ++c) {
                const float cur = S.at(x, y, c);
                const float c1 = S.at(x, y + 1, c), c2 = S.at(x, y - 1, c), c3 = S.at(x + 1, y, c), c4 = S.at(x - 1, y, c);
                const float prevColor = tm(bilinear<float>(P, u - vu, v - vv, c));
                const float mn = tm(std::fmin(cur, std::fmin(c1, std::fmin(c2, std::fmin(c3, c4)))));
                const float mx = tm(std::fmax(cur, std::fmax(c1, std::fmax(c2, std::fmax(c3, c4)))));
                const float pc = std::fmin(std::fmax(prevColor, mn), mx);
                out[p * 3 + c] = itm(tm(cur)) * (1.0f - a) + itm(pc) * a;
            }
        }
}

// gammaTonemap.fs:20-29, 112-114, 125-142: per channel the ACES film fit in
// the shader's order, saturate, pow(x, 1.0 / gamma), the RGBA8 UNORM store
// floor(x * 255 + 0.5).  film (may be NULL) receives the value before the store.
uint16_t orc_half_rn(float f) { return floatToHalfRN(f); }

// The sun cascades' depth render, RenderCSMDepth_ (StratusRendererBackend.cpp:
// 1081-1158) with csm.vs:29-41: per cascade, each triangle through
// cascadeProjViews[c] (mat4 * vec4 as GLSL sums it, then the divide by w);
// culling as the pass sets it, each mesh's mode reversed (RenderImmediate_(...,
// true), :1029-1042): the NDC signed area ex1 ey2 - ex2 ey1 > 0 is
// counter-clockwise; a texel centre (2 (i + 0.5) / res - 1, ...) is covered
// when the three edge functions share a sign; z is interpolated by them;
// glPolygonOffset(slope, 0) (:1097) adds slope * max(|dz_w/dx_w|, |dz_w/dy_w|)
// = slope * max(|A|, |B|) / res for the NDC plane z = z0 + A dx + B dy; the
// window depth z * 0.5 + 0.5 + offset is clamped to [0, 1] (GL_DEPTH_CLAMP,
// :1094, which also keeps every caster unclipped) and kept when below the
// stored one (LESS, clear 1).  GL's rasteriser's exact edge ownership
// (top-left rule) is implementation-defined; here an edge texel belongs to both
// triangles and the first wins ties.
void orc_render_csm(const float* pos, int count, int meshCull, float slope, const sgi_csm* csm, float* depth) {
    const int res = csm->resolution;
    const int keepSign = meshCull == SGI_CULLING_CCW ? -1 : (meshCull == SGI_CULLING_CW ? 1 : 0);
    auto edge = [](float ax, float ay, float bx, float by, float sx, float sy) {
        return (bx - ax) * (sy - ay) - (by - ay) * (sx - ax);
    };
    auto work = [&](int c) {
        std::vector<float> t;  // kept triangles: 9 NDC values + offset
        for (int k = 0; k < count; ++k) {
            float v[10];
            for (int i = 0; i < 3; ++i) {
                float cl[4];
                mulMat4<float>(csm->cascadeProjViews[c], pos[9 * k + 3 * i], pos[9 * k + 3 * i + 1],
                               pos[9 * k + 3 * i + 2], 1.0f, cl);
                v[3 * i] = cl[0] / cl[3];
                v[3 * i + 1] = cl[1] / cl[3];
                v[3 * i + 2] = cl[2] / cl[3];
            }
            const float ex1 = v[3] - v[0], ey1 = v[4] - v[1], ez1 = v[5] - v[2];
            const float ex2 = v[6] - v[0], ey2 = v[7] - v[1], ez2 = v[8] - v[2];
            const float det = ex1 * ey2 - ex2 * ey1;
            const bool facing = keepSign > 0 ? det > 0.0f : (keepSign < 0 ? det < 0.0f : det != 0.0f);
            if (!facing) continue;
            const float A = (ez1 * ey2 - ez2 * ey1) / det, B = (ex1 * ez2 - ex2 * ez1) / det;
            v[9] = slope * (std::fmax(std::fabs(A), std::fabs(B)) / (float)res);
            t.insert(t.end(), v, v + 10);
        }
        const size_t n = t.size() / 10;
        float* layer = depth + (size_t)c * res * res;
        for (int j = 0; j < res; ++j)
            for (int i = 0; i < res; ++i) {
                const float sx = 2.0f * ((float)i + 0.5f) / (float)res - 1.0f;
                const float sy = 2.0f * ((float)j + 0.5f) / (float)res - 1.0f;
                float best = 1.0f;
                for (size_t q = 0; q < n; ++q) {
                    const float* v = &t[10 * q];
                    const float w0 = edge(v[3], v[4], v[6], v[7], sx, sy);
                    const float w1 = edge(v[6], v[7], v[0], v[1], sx, sy);
                    const float w2 = edge(v[0], v[1], v[3], v[4], sx, sy);
                    const bool in = (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) || (w0 <= 0.0f && w1 <= 0.0f && w2 <= 0.0f);
                    const float sum = w0 + w1 + w2;
                    if (!in || sum == 0.0f) continue;
                    const float z = (w0 * v[2] + w1 * v[5] + w2 * v[8]) / sum;
                    float dw = z * 0.5f + 0.5f + v[9];
                    dw = dw < 0.0f ? 0.0f : (dw > 1.0f ? 1.0f : dw);
                    if (dw < best) best = dw;
                }
                layer[(size_t)j * res + i] = best;
            }
    };
    std::vector<std::thread> th;
    for (int c = 0; c < SGI_NUM_CASCADES; ++c) th.emplace_back(work, c);
    for (auto& x : th) x.join();
}

// The VPL shadow pre-pass: the cube update loop of StratusRendererBackend.cpp:
// 1426-1575, shadow.vs:22-33 and shadowVpl.fs:17-37, for the lights
// vpls[updates[s]] into cubes entries[updates[s]].  Face f of a light is the
// view lookAt(pos, pos + axis_f, up_f) (:873-886) under perspective(90 deg, 1,
// near, farPlane) (:1450): the lookAt bases are signed permutations, so the
// ray through the centre of window pixel (i, j) is the GL major-axis table's
// direction d of cube texel (i, j) with major component +-1, and a point at
// view depth t along it is pos + t d.  Each layer starts cleared (distance 1,
// colour 0, :1453-1459); each triangle, in order, gives a fragment at texel
// (i, j) when the ray meets it (the three signed volumes d . (b x c),
// d . (c x a), d . (a x b) of the light-relative vertices of one sign, not all
// zero) at view depth t = det(a, b, c) / (their sum) within [near, far] (GL
// clipping; GL_DEPTH_CLAMP is off here), with value saturate(t |d| / far)
// (gl_FragDepth = saturate(|fsPosition - lightPos| / farPlane)) stored in the
// distance format (F32 value, binary16 or the UNORM16 code rint(v 65535)) and
// kept when below the stored one (LESS); its colour goes to the RGBA8 layer
// with alpha 255.  The exact pixel coverage of GL's rasteriser on shared
// edges (top-left rule, sub-pixel snapping) is implementation-defined; here a
// ray on an edge meets both triangles and the first in order wins ties.
namespace {
uint32_t depthKey(int fmt, float v) {
    v = v < 1.0f ? v : 1.0f;
    if (fmt == SGI_TEXEL_F32) { uint32_t b; std::memcpy(&b, &v, 4); return b; }
    if (fmt == SGI_TEXEL_F16) return floatToHalfRN(v);
    return (uint32_t)std::nearbyint(v * 65535.0f);
}
V3<float> faceRay(int f, int i, int j, int res) {
    const float sc = 2.0f * ((float)i + 0.5f) / (float)res - 1.0f;
    const float tc = 2.0f * ((float)j + 0.5f) / (float)res - 1.0f;
    switch (f) {
        case 0: return v3(1.0f, -tc, -sc);
        case 1: return v3(-1.0f, -tc, sc);
        case 2: return v3(sc, 1.0f, tc);
        case 3: return v3(sc, -1.0f, -tc);
        case 4: return v3(sc, -tc, 1.0f);
        default: return v3(-sc, -tc, -1.0f);
    }
}
V3<float> cross3(V3<float> a, V3<float> b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
float dot3(V3<float> a, V3<float> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
}  // namespace

void orc_render_vpl_cubes(const float* pos, const uint32_t* colors, int count, int cull, const sgi_vpl* vpls,
                          const int32_t* updates, int numUpdates, const sgi_atlas_entry* entries, float nearPlane,
                          const sgi_cube_atlases* dist, const sgi_cube_atlases* diffuse) {
    const int res = dist->resolution;
    auto work = [&](int job) {
        const int s = job / 6, f = job % 6;
        const sgi_vpl& v = vpls[updates[s]];
        const sgi_atlas_entry e = entries[updates[s]];
        if (e.index < 0 || e.index >= dist->num_atlases || e.layer < 0 || e.layer >= dist->layers_per_atlas) return;
        const V3<float> L = v3(v.position[0], v.position[1], v.position[2]);
        const float farP = v.farPlane;
        const size_t layerOff = ((size_t)e.layer * 6 + f) * res * res;
        for (int j = 0; j < res; ++j)
            for (int i = 0; i < res; ++i) {
                const V3<float> d = faceRay(f, i, j, res);
                const float dlen = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                uint32_t best = depthKey(dist->format, 1.0f), col = 0;
                for (int k = 0; k < count; ++k) {
                    const float* q = pos + (size_t)k * 9;
                    const V3<float> a = v3(q[0] - L.x, q[1] - L.y, q[2] - L.z);
                    const V3<float> b = v3(q[3] - L.x, q[4] - L.y, q[5] - L.z);
                    const V3<float> c = v3(q[6] - L.x, q[7] - L.y, q[8] - L.z);
                    const V3<float> C0 = cross3(b, c), C1 = cross3(c, a), C2 = cross3(a, b);
                    // SetCullState (StratusRendererBackend.cpp:901-918): seen from the light the
                    // triangle winds counter-clockwise (its front under GL_CCW) when a . (b x c) < 0
                    const float V = dot3(a, C0);
                    if ((cull == SGI_CULLING_CCW && !(V < 0.0f)) || (cull == SGI_CULLING_CW && !(V > 0.0f))) continue;
                    const float e0 = dot3(d, C0), e1 = dot3(d, C1), e2 = dot3(d, C2);
                    const bool in = (e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) || (e0 <= 0.0f && e1 <= 0.0f && e2 <= 0.0f);
                    const float den = e0 + e1 + e2;
                    if (!in || den == 0.0f) continue;
                    const float t = V / den;
                    if (!(t >= nearPlane && t <= farP)) continue;
                    const uint32_t key = depthKey(dist->format, (t * dlen) / farP);
                    if (key < best) {
                        best = key;
                        col = colors ? colors[k] : 0u;
                    }
                }
                const size_t off = layerOff + (size_t)j * res + i;
                if (dist->format == SGI_TEXEL_F32) std::memcpy((float*)dist->atlas[e.index] + off, &best, 4);
                else ((uint16_t*)dist->atlas[e.index])[off] = (uint16_t)best;
                if (diffuse) ((uint32_t*)diffuse->atlas[e.index])[off] = col | 0xFF000000u;
            }
    };
    const int jobs = numUpdates * 6;
    const int nt = std::max(1, std::min<int>(jobs, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::atomic<int> next{0};
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int job = next++; job < jobs; job = next++) work(job);
        });
    for (auto& t : th) t.join();
}

void orc_tonemap(const float* lit, long n, float gamma, uint32_t* out, float* film) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    const float invGamma = 1.0f / gamma;
    const auto sat = [](float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); };
    for (long p = 0; p < n; ++p) {
        uint32_t t = 0xFF000000u;
        for (int k = 0; k < 3; ++k) {
            const float x = lit[3 * p + k];
            const float v = sat(std::pow(sat((x * (a * x + b)) / (x * (c * x + d) + e)), invGamma));
            if (film) film[3 * p + k] = v;
            t |= (uint32_t)(v * 255.0f + 0.5f) << (8 * k);
        }
        out[p] = t;
    }
}

// The history copy of PerformTaaPostFx_ (StratusRendererBackend.cpp:2342-2349):
// the TAA output lives in an RGBA16F target (:708-711), so it is first rounded
// to binary16; the NEAREST blit into the RGB8 UNORM previousFrameBuffer
// (:449-453) clamps to [0, 1] and stores floor(v * 255 + 0.5); the next frame's
// TAA reads that texel back as k / 255.  prev receives the values read back.
void orc_taa_history(const float* taa, long n, float* prev) {
    for (long q = 0; q < 3 * n; ++q) {
        const float h = halfToFloat(floatToHalfRN(taa[q]));
        const float s = std::fmin(std::fmax(h, 0.0f), 1.0f);
        const uint32_t k = (uint32_t)(s * 255.0f + 0.5f);
        prev[q] = (float)k / 255.0f;
    }
}

// aa_common.glsl:7-9, in the shader's order.
inline float fxaaLuminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

inline uint32_t unorm8Store(float v) {
    v = std::fmin(std::fmax(v, 0.0f), 1.0f);
    return (uint32_t)(v * 255.0f + 0.5f);
}

// fxaa_luminance.fs:16-20 (PerformFxaaPostFx_'s first pass,
// StratusRendererBackend.cpp:2298-2303): the tonemapped RGBA8 colour (read as
// k / 255) with the luminance of its saturated rgb in alpha, stored into the
// RGBA8 target fxaaFbo1 (:683-688).
void orc_fxaa_luminance(const uint32_t* in, long n, uint32_t* out) {
    for (long p = 0; p < n; ++p) {
        float c[3];
        for (int k = 0; k < 3; ++k) c[k] = (float)((in[p] >> (8 * k)) & 0xFFu) / 255.0f;
        const float sat[3] = {std::fmin(std::fmax(c[0], 0.0f), 1.0f), std::fmin(std::fmax(c[1], 0.0f), 1.0f),
                              std::fmin(std::fmax(c[2], 0.0f), 1.0f)};
        const float lum = fxaaLuminance(sat[0], sat[1], sat[2]);
        out[p] = unorm8Store(c[0]) | (unorm8Store(c[1]) << 8) | (unorm8Store(c[2]) << 16) | (unorm8Store(lum) << 24);
    }
}

// fxaa_smoothing.fs:72-277 (the second pass, :2308-2316) over fxaaFbo1 (RGBA8,
// LINEAR, CLAMP_TO_EDGE, read as k / 255), stored into the RGBA8 target
// fxaaFbo2 with alpha 1.  The centre and textureOffset taps sit on texel
// centres and read the texel; the half-texel edge walk and the final offset
// sample are GL LINEAR taps (bilinear, FP32 weights).  GLSL's order of
// operations is kept; smoothstep(0, 1, x) = t t (3 - 2 t) with t = clamp(x, 0, 1).
void orc_fxaa_smoothing(const uint32_t* lum, int W, int H, float contrastThreshold, float relativeThreshold,
                        float subpixelBlending, uint32_t* out) {
    std::vector<float> tex((size_t)W * H * 4);
    for (size_t p = 0; p < (size_t)W * H; ++p)
        for (int k = 0; k < 4; ++k) tex[4 * p + k] = (float)((lum[p] >> (8 * k)) & 0xFFu) / 255.0f;
    const Plane T{tex.data(), W, H, 4};
    const float tsx = 1.0f / (float)W, tsy = 1.0f / (float)H;  // computeTexelSize (common.glsl:111-115)
    const float edgeStepOffsets[10] = {1.0f, 1.5f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 4.0f};
    const float edgeGuessJumpSize = 8.0f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t p = (size_t)y * W + x;
            const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
            const float lumaCenter = T.at(x, y, 3);
            const float lumaRight = T.at(x + 1, y, 3), lumaLeft = T.at(x - 1, y, 3);
            const float lumaTop = T.at(x, y + 1, 3), lumaBot = T.at(x, y - 1, 3);
            const float lumaMax = std::fmax(std::fmax(std::fmax(std::fmax(lumaCenter, lumaRight), lumaLeft), lumaTop),
                                            lumaBot);
            const float lumaMin = std::fmin(std::fmin(std::fmin(std::fmin(lumaCenter, lumaRight), lumaLeft), lumaTop),
                                            lumaBot);
            const float contrast = lumaMax - lumaMin;
            const float threshold = std::fmax(contrastThreshold, relativeThreshold * lumaMax);
            if (contrast < threshold) {
                out[p] = unorm8Store(T.at(x, y, 0)) | (unorm8Store(T.at(x, y, 1)) << 8) |
                         (unorm8Store(T.at(x, y, 2)) << 16) | 0xFF000000u;
                continue;
            }
            const float lumaTopRight = T.at(x + 1, y + 1, 3), lumaBotRight = T.at(x + 1, y - 1, 3);
            const float lumaTopLeft = T.at(x - 1, y + 1, 3), lumaBotLeft = T.at(x - 1, y - 1, 3);
            float average = 2.0f * (lumaRight + lumaLeft + lumaTop + lumaBot) + lumaTopRight + lumaBotRight +
                            lumaTopLeft + lumaBotLeft;
            average = average / 12.0f;
            const float centerAvgContrast = std::fabs(average - lumaCenter);
            const float normalized = std::fmin(std::fmax(centerAvgContrast / contrast, 0.0f), 1.0f);
            const float t = std::fmin(std::fmax(normalized, 0.0f), 1.0f);
            float pixelBlendFactor = t * t * (3.0f - 2.0f * t);
            pixelBlendFactor = pixelBlendFactor * pixelBlendFactor * subpixelBlending;
            const float horizontal = std::fabs(lumaTop + lumaBot - 2.0f * lumaCenter) * 2.0f +
                                     std::fabs(lumaTopRight + lumaBotRight - 2.0f * lumaRight) +
                                     std::fabs(lumaTopLeft + lumaBotLeft - 2.0f * lumaLeft);
            const float vertical = std::fabs(lumaRight + lumaLeft - 2.0f * lumaCenter) * 2.0f +
                                   std::fabs(lumaTopRight + lumaTopLeft - 2.0f * lumaTop) +
                                   std::fabs(lumaBotRight + lumaBotLeft - 2.0f * lumaBot);
            const bool isHorizontal = horizontal >= vertical;
            float pixelStep = isHorizontal ? tsy : tsx;
            const float positiveLuma = isHorizontal ? lumaTop : lumaRight;
            const float negativeLuma = isHorizontal ? lumaBot : lumaLeft;
            const float positiveGradient = std::fabs(positiveLuma - lumaCenter);
            const float negativeGradient = std::fabs(negativeLuma - lumaCenter);
            float gradient, oppositeLuma;
            if (positiveGradient < negativeGradient) {
                pixelStep = -pixelStep;
                oppositeLuma = negativeLuma;
                gradient = negativeGradient;
            } else {
                oppositeLuma = positiveLuma;
                gradient = positiveGradient;
            }
            float ue = u, ve = v, esx, esy;
            if (isHorizontal) {
                ve += pixelStep * 0.5f;
                esx = tsx;
                esy = 0.0f;
            } else {
                ue += pixelStep * 0.5f;
                esx = 0.0f;
                esy = tsy;
            }
            const float edgeLumaAvg = (lumaCenter + oppositeLuma) * 0.5f;
            const float gradientThreshold = gradient * 0.25f;
            float dist[2], deltas[2];
            const float signs[2] = {1.0f, -1.0f};
            for (int i = 0; i < 2; ++i) {
                const float sx = esx * signs[i], sy = esy * signs[i];  // edgeSteps[i] = +-edgeStep
                float pu = ue + sx * edgeStepOffsets[0], pv = ve + sy * edgeStepOffsets[0];
                float delta = bilinear<float>(T, pu, pv, 3) - edgeLumaAvg;
                bool atEnd = std::fabs(delta) >= gradientThreshold;
                for (int j = 0; j < 9 && !atEnd; ++j) {
                    pu += sx * edgeStepOffsets[j + 1];
                    pv += sy * edgeStepOffsets[j + 1];
                    delta = bilinear<float>(T, pu, pv, 3) - edgeLumaAvg;
                    atEnd = std::fabs(delta) >= gradientThreshold;
                }
                if (!atEnd) {
                    pu += sx * edgeGuessJumpSize;
                    pv += sy * edgeGuessJumpSize;
                }
                deltas[i] = delta;
                dist[i] = isHorizontal ? (signs[i] * pu) + (-signs[i] * u) : (signs[i] * pv) + (-signs[i] * v);
            }
            float shortest;
            bool deltaSign;
            if (dist[0] <= dist[1]) {
                shortest = dist[0];
                deltaSign = deltas[0] >= 0.0f;
            } else {
                shortest = dist[1];
                deltaSign = deltas[1] >= 0.0f;
            }
            const float edgeBlendFactor =
                deltaSign == (lumaCenter - edgeLumaAvg >= 0.0f) ? 0.0f : 0.5f - shortest / (dist[0] + dist[1]);
            const float finalBlend = std::fmax(pixelBlendFactor, edgeBlendFactor);
            float fu = u, fv = v;
            if (isHorizontal) fv += pixelStep * finalBlend;
            else fu += pixelStep * finalBlend;
            out[p] = unorm8Store(bilinear<float>(T, fu, fv, 0)) | (unorm8Store(bilinear<float>(T, fu, fv, 1)) << 8) |
                     (unorm8Store(bilinear<float>(T, fu, fv, 2)) << 16) | 0xFF000000u;
        }
}

void orc_color_vpls(const sgi_frame* fr, int count, const sgi_atlas_entry* entries, const sgi_cube_atlases* diffuse,
                    const sgi_cube_atlases* dist, float colorMultiplier, sgi_vpl* vpls) {
    colorVpls(*fr, count, entries, *diffuse, *dist, colorMultiplier, vpls);
}

int orc_cull_global(const sgi_frame* fr, const sgi_csm* csm, const sgi_vpl* in, int n, sgi_vpl* out, int32_t* idx) {
    return cullGlobal<float>(*fr, *csm, in, n, out, idx);
}

// Tile culling for tiles [t0, t1) of a ceil-divided grid; tiles = int32[T][1+K].
void orc_cull_tiles(const sgi_frame* fr, const sgi_gbuffer* g, int ts, int K, const sgi_vpl* vpls,
                    const int32_t* visible, int remap, int t0, int t1, sgi_tile_stage1* stage1, int32_t* tiles) {
    const GB gb = wrap(*g);
    const int tx = (g->width + ts - 1) / ts;
    forRows(t0, t1, [&](int t) {
        float pos[3], nrm[3];
        tileStage1<float>(*fr, gb, ts, t % tx, t / tx, pos, nrm);
        if (stage1) {
            for (int k = 0; k < 3; ++k) {
                stage1[t].averageLocalPosition[k] = pos[k];
                stage1[t].averageLocalNormal[k] = nrm[k];
            }
            stage1[t].averageLocalPosition[3] = 1.0f;
            stage1[t].averageLocalNormal[3] = 1.0f;
        }
        tileStage2<float>(pos, nrm, vpls, visible, remap, K, tiles + (size_t)t * (1 + K));
    });
}

// Gather over pixel rows [y0, y1).  mode: 0 random (reference), 1 tile lists, 2 every visible VPL,
// 3 `samples` hashed draws over every visible VPL (rejectionLimit as gatherPixel).
// precision 0 = FP32 restatement (float outputs), 1 = FP64 twin (double outputs).
void orc_gather(const sgi_frame* fr, const sgi_gbuffer* g, const sgi_vpl* vpls, int numVisible,
                const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases* cubes, const int32_t* tiles, int ts,
                int K, int mode, int samples, int rejectionLimit, int precision, int y0, int y1, void* color,
                void* reservoir) {
    const GB gb = wrap(*g);
    const int W = g->width, tx = (W + ts - 1) / ts;
    forRows(y0, y1, [&](int y) {
        for (int x = 0; x < W; ++x) {
            const int32_t* list = nullptr;
            if (mode == TILES) list = tiles + (size_t)((y / ts) * tx + (x / ts)) * (1 + K);
            const size_t p = (size_t)(y - y0) * W + x;
            if (precision == 0) {
                float c[3], r[4];
                gatherPixel<float>(*fr, gb, x, y, vpls, numVisible, shadowIndices, *cubes, list, mode, samples,
                                   rejectionLimit, c, r);
                if (color) for (int k = 0; k < 3; ++k) ((float*)color)[p * 3 + k] = c[k];
                for (int k = 0; k < 4; ++k) ((float*)reservoir)[p * 4 + k] = r[k];
            } else {
                double c[3], r[4];
                gatherPixel<double>(*fr, gb, x, y, vpls, numVisible, shadowIndices, *cubes, list, mode, samples,
                                    rejectionLimit, c, r);
                if (color) for (int k = 0; k < 3; ++k) ((double*)color)[p * 3 + k] = c[k];
                for (int k = 0; k < 4; ++k) ((double*)reservoir)[p * 4 + k] = r[k];
            }
        }
    });
}

void orc_direct_light(const sgi_frame* fr, const sgi_csm* csm, const sgi_gbuffer* g, int y0, int y1, float* lit) {
    const GB gb = wrap(*g);
    const int W = g->width;
    forRows(y0, y1, [&](int y) {
        for (int x = 0; x < W; ++x) {
            float c[3];
            directPixel<float>(*fr, *csm, gb, x, y, c);
            for (int k = 0; k < 3; ++k) lit[((size_t)(y - y0) * W + x) * 3 + k] = c[k];
        }
    });
}

void orc_denoise(const sgi_frame* fr, const sgi_gbuffer* cur, const sgi_gbuffer* prev, int numVisible,
                 const float* screen, const float* color, const float* reservoir, const float* prevGi,
                 const float* prevHistory, float* combined, float* giOut, float* resOut, float* histOut) {
    denoise<float>(*fr, wrap(*cur), wrap(prev ? *prev : *cur), numVisible, screen, color, reservoir, prevGi,
                   prevHistory, combined, giOut, resOut, histOut);
}

}  // extern "C"
