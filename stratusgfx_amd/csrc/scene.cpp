// Seeded analytic scene generator: scene construction, camera and cascade
// matrices, VPL sampling on sun-lit surfaces, and the host ray casts.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "scene_internal.h"

namespace {

using sgi_rt::Box;

// splitmix64: a counter-based generator, so every draw depends only on the seed.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {  // Box-Muller
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

Box makeBox(float cx, float cy, float cz, float hx, float hy, float hz, float yawDeg,
            float r, float g, float b, float rough, float metal, float id) {
    Box bx{};
    bx.c[0] = cx; bx.c[1] = cy; bx.c[2] = cz;
    bx.h[0] = hx; bx.h[1] = hy; bx.h[2] = hz;
    const double a = yawDeg * 3.141592653589793 / 180.0;
    bx.cosy = (float)std::cos(a);
    bx.siny = (float)std::sin(a);
    if (yawDeg == 0.0f) { bx.cosy = 1.0f; bx.siny = 0.0f; }
    bx.albedo[0] = r; bx.albedo[1] = g; bx.albedo[2] = b;
    bx.roughness = rough;
    bx.metallic = metal;
    bx.id = id;
    return bx;
}

void normalize3(float v[3]) {
    const double l = std::sqrt((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] = (float)(v[k] / l);
}

void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

void normalize3d(double v[3]) {
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] /= l;
}

// Column-major products and inverse, in double.
void matmul(const double a[16], const double b[16], double o[16]) {
    double t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            t[c * 4 + r] = s;
        }
    std::memcpy(o, t, sizeof(t));
}

bool invert(const double m[16], double inv[16]) {
    // Gauss-Jordan with partial pivoting on a row-major copy.
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (std::fabs(a[p][c]) < 1e-300) return false;
        if (p != c)
            for (int k = 0; k < 8; ++k) std::swap(a[p][k], a[c][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[c * 4 + r] = a[r][c + 4];
    return true;
}

void toFloat16(const double* d, float* f) {
    for (int i = 0; i < 16; ++i) f[i] = (float)d[i];
}

struct ViewMats {
    double view[16], proj[16], viewProj[16], invViewProj[16];
    double eye[3], fwd[3], right[3], up[3];
};

void forwardFromYawPitch(double yawDeg, double pitchDeg, double f[3]) {
    const double y = yawDeg * 3.141592653589793 / 180.0, p = pitchDeg * 3.141592653589793 / 180.0;
    f[0] = std::sin(y) * std::cos(p);
    f[1] = std::sin(p);
    f[2] = -std::cos(y) * std::cos(p);
}

ViewMats buildView(const sgi_view& v) {
    ViewMats m{};
    for (int k = 0; k < 3; ++k) m.eye[k] = v.eye[k];
    forwardFromYawPitch(v.yaw_deg, v.pitch_deg, m.fwd);
    const double worldUp[3] = {0.0, 1.0, 0.0};
    cross3(m.fwd, worldUp, m.right);
    normalize3d(m.right);
    cross3(m.right, m.fwd, m.up);
    // lookAt (right-handed, camera looks down -Z)
    std::memset(m.view, 0, sizeof(m.view));
    for (int k = 0; k < 3; ++k) {
        m.view[k * 4 + 0] = m.right[k];
        m.view[k * 4 + 1] = m.up[k];
        m.view[k * 4 + 2] = -m.fwd[k];
    }
    m.view[12] = -(m.right[0] * m.eye[0] + m.right[1] * m.eye[1] + m.right[2] * m.eye[2]);
    m.view[13] = -(m.up[0] * m.eye[0] + m.up[1] * m.eye[1] + m.up[2] * m.eye[2]);
    m.view[14] = (m.fwd[0] * m.eye[0] + m.fwd[1] * m.eye[1] + m.fwd[2] * m.eye[2]);
    m.view[15] = 1.0;
    // perspective (GL clip z in [-1, 1])
    const double f = 1.0 / std::tan(v.fovy_deg * 3.141592653589793 / 360.0);
    const double aspect = (double)v.width / (double)v.height;
    const double n = v.znear, fa = v.zfar;
    std::memset(m.proj, 0, sizeof(m.proj));
    m.proj[0] = f / aspect;
    m.proj[5] = f;
    m.proj[10] = -(fa + n) / (fa - n);
    m.proj[11] = -1.0;
    m.proj[14] = -2.0 * fa * n / (fa - n);
    matmul(m.proj, m.view, m.viewProj);
    invert(m.viewProj, m.invViewProj);
    return m;
}

void buildGrid(sgi_scene& s) {
    const int n = (int)s.boxes.size() - s.numAlways;
    if (n <= 0) {
        s.grid.nx = 0;
        return;
    }
    const float e = s.desc.city_half_extent;
    const float cell = 4.0f;
    const int nx = (int)std::ceil(2.0f * e / cell);
    s.grid.x0 = -e;
    s.grid.z0 = -e;
    s.grid.cell = cell;
    s.grid.nx = nx;
    s.grid.nz = nx;
    float ymax = 0.0f;
    std::vector<std::vector<int>> cells((size_t)nx * nx);
    for (int b = s.numAlways; b < (int)s.boxes.size(); ++b) {
        const Box& bx = s.boxes[b];
        const float ex = std::fabs(bx.cosy) * bx.h[0] + std::fabs(bx.siny) * bx.h[2];
        const float ez = std::fabs(bx.siny) * bx.h[0] + std::fabs(bx.cosy) * bx.h[2];
        ymax = std::max(ymax, bx.c[1] + bx.h[1]);
        int i0 = (int)std::floor((bx.c[0] - ex - s.grid.x0) / cell);
        int i1 = (int)std::floor((bx.c[0] + ex - s.grid.x0) / cell);
        int k0 = (int)std::floor((bx.c[2] - ez - s.grid.z0) / cell);
        int k1 = (int)std::floor((bx.c[2] + ez - s.grid.z0) / cell);
        i0 = std::max(i0, 0); k0 = std::max(k0, 0);
        i1 = std::min(i1, nx - 1); k1 = std::min(k1, nx - 1);
        for (int k = k0; k <= k1; ++k)
            for (int i = i0; i <= i1; ++i) cells[(size_t)k * nx + i].push_back(b);
    }
    s.grid.ymax = ymax + 1e-3f;
    s.cellStart.assign((size_t)nx * nx + 1, 0);
    s.cellItems.clear();
    for (size_t c = 0; c < cells.size(); ++c) {
        s.cellStart[c] = (int)s.cellItems.size();
        s.cellItems.insert(s.cellItems.end(), cells[c].begin(), cells[c].end());
    }
    s.cellStart[cells.size()] = (int)s.cellItems.size();
}

void buildCornell(sgi_scene& s) {
    const float W = 0.73f;
    auto& B = s.boxes;
    // 5 walls (open toward +Z, where the camera and the sun are)
    B.push_back(makeBox(0.0f, -0.1f, -5.0f, 5.2f, 0.1f, 5.2f, 0.0f, W, W, W, 0.8f, 0.0f, 1.0f));     // floor
    B.push_back(makeBox(0.0f, 10.1f, -5.0f, 5.2f, 0.1f, 5.2f, 0.0f, W, W, W, 0.8f, 0.0f, 2.0f));     // ceiling
    B.push_back(makeBox(0.0f, 5.0f, -10.1f, 5.2f, 5.2f, 0.1f, 0.0f, W, W, W, 0.8f, 0.0f, 3.0f));     // back
    B.push_back(makeBox(-5.1f, 5.0f, -5.0f, 0.1f, 5.2f, 5.2f, 0.0f, 0.63f, 0.065f, 0.05f, 0.8f, 0.0f, 4.0f));  // red
    B.push_back(makeBox(5.1f, 5.0f, -5.0f, 0.1f, 5.2f, 5.2f, 0.0f, 0.14f, 0.45f, 0.091f, 0.8f, 0.0f, 5.0f));  // green
    // 2 rotated boxes; the seed jitters their yaw a little
    Rng rng(s.desc.seed);
    const float yawTall = 17.0f + (float)rng.uniform(-3.0, 3.0);
    const float yawShort = -18.0f + (float)rng.uniform(-3.0, 3.0);
    B.push_back(makeBox(-1.8f, 3.0f, -6.5f, 1.5f, 3.0f, 1.5f, yawTall, W, W, W, 0.5f, 0.0f, 6.0f));
    B.push_back(makeBox(1.8f, 1.5f, -3.5f, 1.5f, 1.5f, 1.5f, yawShort, W, W, W, 0.35f, 0.0f, 7.0f));
    s.numAlways = (int)B.size();
    float sun[3] = {0.25f, 0.55f, 1.0f};
    normalize3(sun);
    std::memcpy(s.sunDir, sun, sizeof(sun));
    s.sunColor[0] = 3.0f; s.sunColor[1] = 2.85f; s.sunColor[2] = 2.55f;
    const float bmin[3] = {-5.3f, -0.3f, -10.3f}, bmax[3] = {5.3f, 10.3f, 0.3f};
    std::memcpy(s.boundsMin, bmin, sizeof(bmin));
    std::memcpy(s.boundsMax, bmax, sizeof(bmax));
    s.grid.nx = 0;
}

void buildCity(sgi_scene& s) {
    const float e = s.desc.city_half_extent;
    auto& B = s.boxes;
    B.push_back(makeBox(0.0f, -0.5f, 0.0f, e + 10.0f, 0.5f, e + 10.0f, 0.0f, 0.45f, 0.45f, 0.42f, 0.9f, 0.0f, 1.0f));
    s.numAlways = 1;
    Rng rng(s.desc.seed);
    float maxH = 0.0f;
    for (int i = 0; i < s.desc.num_city_boxes; ++i) {
        const float w = (float)rng.uniform(1.0, 8.0), d = (float)rng.uniform(1.0, 8.0);
        const float h = (float)std::exp(2.0 + 0.6 * rng.normal());
        const float x = (float)rng.uniform(-e + 4.0, e - 4.0), z = (float)rng.uniform(-e + 4.0, e - 4.0);
        const float r = (float)rng.uniform(0.1, 0.9), g = (float)rng.uniform(0.1, 0.9),
                    b = (float)rng.uniform(0.1, 0.9);
        const float rough = (float)rng.uniform(0.05, 1.0);
        const float metal = rng.uniform() < 0.2 ? 1.0f : 0.0f;
        B.push_back(makeBox(x, 0.5f * h, z, 0.5f * w, 0.5f * h, 0.5f * d, 0.0f, r, g, b, rough, metal,
                            (float)(i + 2)));
        maxH = std::max(maxH, h);
    }
    float sun[3] = {0.35f, 0.8f, 0.45f};
    normalize3(sun);
    std::memcpy(s.sunDir, sun, sizeof(sun));
    s.sunColor[0] = 3.0f; s.sunColor[1] = 2.85f; s.sunColor[2] = 2.55f;
    const float bmin[3] = {-e - 10.0f, -1.0f, -e - 10.0f}, bmax[3] = {e + 10.0f, maxH + 1.0f, e + 10.0f};
    std::memcpy(s.boundsMin, bmin, sizeof(bmin));
    std::memcpy(s.boundsMax, bmax, sizeof(bmax));
    buildGrid(s);
}

// Area-weighted sampling of box faces, kept when the face sees the sun unoccluded.
void sampleVpls(sgi_scene& s) {
    const sgi_rt::SceneView sv = s.hostView();
    std::vector<double> faceArea;  // 6 faces per box
    faceArea.reserve(s.boxes.size() * 6);
    double total = 0.0;
    for (const Box& b : s.boxes) {
        const double ax[3] = {4.0 * b.h[1] * b.h[2], 4.0 * b.h[0] * b.h[2], 4.0 * b.h[0] * b.h[1]};
        for (int f = 0; f < 6; ++f) {
            total += ax[f / 2];
            faceArea.push_back(total);
        }
    }
    Rng rng(s.desc.seed ^ 0x5851F42D4C957F2Dull);
    const float eps = 0.05f, radius = 500.0f;
    const long maxTries = 2000000L + 400L * s.desc.num_vpls;
    for (long tries = 0; tries < maxTries && (int)s.vpls.size() < s.desc.num_vpls; ++tries) {
        const double u = rng.uniform() * total;
        const size_t fi = (size_t)(std::lower_bound(faceArea.begin(), faceArea.end(), u) - faceArea.begin());
        const Box& b = s.boxes[std::min(fi / 6, s.boxes.size() - 1)];
        const int f = (int)(fi % 6), axis = f / 2;
        const float sgn = (f % 2) ? 1.0f : -1.0f;
        float lp[3], ln[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 3; ++k) lp[k] = (float)rng.uniform(-1.0, 1.0) * b.h[k];
        lp[axis] = sgn * b.h[axis];
        ln[axis] = sgn;
        const float p[3] = {b.c[0] + b.cosy * lp[0] + b.siny * lp[2], b.c[1] + lp[1],
                            b.c[2] - b.siny * lp[0] + b.cosy * lp[2]};
        const float n[3] = {b.cosy * ln[0] + b.siny * ln[2], ln[1], -b.siny * ln[0] + b.cosy * ln[2]};
        const float ndl = n[0] * s.sunDir[0] + n[1] * s.sunDir[1] + n[2] * s.sunDir[2];
        if (ndl <= 0.05f) continue;
        const float o[3] = {p[0] + eps * n[0], p[1] + eps * n[1], p[2] + eps * n[2]};
        // inside another box (overlapping city boxes) or out of the scene
        if (o[1] < 0.0f && s.desc.kind == SGI_SCENE_CITY) continue;
        if (s.desc.kind == SGI_SCENE_CORNELL &&  // only the room's interior faces
            (std::fabs(o[0]) > 5.0f || o[1] < 0.0f || o[1] > 10.0f || o[2] < -10.0f || o[2] > 0.2f))
            continue;
        const sgi_rt::Hit toSun = sgi_rt::traceScene(sv, o, s.sunDir, 0.0f, 1e4f);
        if (toSun.box >= 0) continue;
        bool inside = false;
        for (const Box& q : s.boxes) {
            const float dx = o[0] - q.c[0], dy = o[1] - q.c[1], dz = o[2] - q.c[2];
            const float lx = q.cosy * dx - q.siny * dz, lz = q.siny * dx + q.cosy * dz;
            if (std::fabs(lx) < q.h[0] && std::fabs(dy) < q.h[1] && std::fabs(lz) < q.h[2]) { inside = true; break; }
        }
        if (inside) continue;
        sgi_vpl v{};
        v.position[0] = o[0]; v.position[1] = o[1]; v.position[2] = o[2]; v.position[3] = 1.0f;
        for (int k = 0; k < 3; ++k) v.color[k] = b.albedo[k] * s.sunColor[k] * ndl * s.desc.vpl_color_multiplier;
        v.color[3] = 1.0f;
        // vpl_light_color.cs:55,71: specular position pushed along -sun by 1.1 x the
        // distance the VPL's own cube map records in direction -sun.
        const float down[3] = {-s.sunDir[0], -s.sunDir[1], -s.sunDir[2]};
        const sgi_rt::Hit below = sgi_rt::traceScene(sv, o, down, 0.0f, radius);
        const float mag = below.box >= 0 ? below.t : radius;
        for (int k = 0; k < 3; ++k) v.specularPosition[k] = o[k] - 1.1f * mag * s.sunDir[k];
        v.specularPosition[3] = 1.0f;
        v.radius = radius;
        v.farPlane = radius;
        v.intensity = 1.0f;
        s.vpls.push_back(v);
    }
}

void buildCsm(sgi_scene& s, const sgi_view& v, const ViewMats& m, sgi_csm* out) {
    sgi_rt::CsmRays& r = s.csm;
    // light basis: fwd away from the sun
    double fwd[3] = {-s.sunDir[0], -s.sunDir[1], -s.sunDir[2]};
    normalize3d(fwd);
    double upRef[3] = {0.0, 1.0, 0.0};
    if (std::fabs(fwd[1]) > 0.99) { upRef[0] = 1.0; upRef[1] = 0.0; }
    // The reference's light camera looks at the sun (infiniteLightDirection = its forward,
    // StratusRendererBackend.cpp:2455-2457, Camera::GetDirection) and is right-handed, so its light
    // space has x cross y = the direction away from the sun: right = upRef x fwd, up = fwd x right.
    // That fixes the window winding its (reversed) face culling acts on.
    double right[3], up[3];
    cross3(upRef, fwd, right);
    normalize3d(right);
    cross3(fwd, right, up);
    for (int k = 0; k < 3; ++k) { r.right[k] = (float)right[k]; r.up[k] = (float)up[k]; r.fwd[k] = (float)fwd[k]; }
    // traced range: the scene bounds along fwd (every caster; depths map through the slice range below)
    double dmin = 1e300, dmax = -1e300;
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? s.boundsMax[0] : s.boundsMin[0], (c & 2) ? s.boundsMax[1] : s.boundsMin[1],
                             (c & 4) ? s.boundsMax[2] : s.boundsMin[2]};
        const double d = p[0] * fwd[0] + p[1] * fwd[1] + p[2] * fwd[2];
        dmin = std::min(dmin, d);
        dmax = std::max(dmax, d);
    }
    dmin -= 1.0;
    dmax += 1.0;
    // StratusRendererFrontend.cpp:622-640: lambda-blended log/uniform splits, 4 m overlap
    const double znear = 1.0, zfar = v.zfar, lambda = 0.5;
    double ends[4], begins[4];
    for (int i = 0; i < 4; ++i) {
        const double p = (i + 1) / 4.0;
        const double lg = znear * std::pow(zfar / znear, p), un = znear + (zfar - znear) * p;
        ends[i] = std::floor(lambda * lg + (1.0 - lambda) * un);
    }
    begins[0] = 0.0;
    for (int i = 1; i < 4; ++i) begins[i] = ends[i - 1] - 4.0;
    const double ar = (double)v.width / (double)v.height;
    const double tanHalf = std::tan(v.fovy_deg * 3.141592653589793 / 360.0);
    const int res = v.csm_resolution;
    for (int i = 0; i < 4; ++i) {
        const double a = begins[i], b = ends[i];
        double corners[8][3];
        int ci = 0;
        for (int far = 0; far < 2; ++far) {
            const double z = far ? b : a;
            const double xh = z * ar * tanHalf, yh = z * tanHalf;
            for (int sy = -1; sy <= 1; sy += 2)
                for (int sx = -1; sx <= 1; sx += 2) {
                    for (int k = 0; k < 3; ++k)
                        corners[ci][k] = m.eye[k] + m.right[k] * sx * xh + m.up[k] * sy * yh + m.fwd[k] * z;
                    ++ci;
                }
        }
        double maxLen = 0.0;
        for (int p = 0; p < 8; ++p)
            for (int q = p + 1; q < 8; ++q) {
                const double dx = corners[p][0] - corners[q][0], dy = corners[p][1] - corners[q][1],
                             dz = corners[p][2] - corners[q][2];
                maxLen = std::max(maxLen, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        const double dk = std::ceil(maxLen);
        double mnx = 1e300, mxx = -1e300, mny = 1e300, mxy = -1e300, mne = 1e300, mxe = -1e300;
        for (int p = 0; p < 8; ++p) {
            const double x = corners[p][0] * right[0] + corners[p][1] * right[1] + corners[p][2] * right[2];
            const double y = corners[p][0] * up[0] + corners[p][1] * up[1] + corners[p][2] * up[2];
            const double e = corners[p][0] * fwd[0] + corners[p][1] * fwd[1] + corners[p][2] * fwd[2];
            mnx = std::min(mnx, x); mxx = std::max(mxx, x);
            mny = std::min(mny, y); mxy = std::max(mxy, y);
            mne = std::min(mne, e); mxe = std::max(mxe, e);
        }
        const double T = dk / res;  // texel-snapped centre keeps the cascade stable
        const double cx = std::floor((mxx + mnx) / (2.0 * T)) * T, cy = std::floor((mxy + mny) / (2.0 * T)) * T;
        r.cx[i] = (float)cx; r.cy[i] = (float)cy; r.dk[i] = (float)dk;
        r.dmin[i] = (float)dmin; r.dmax[i] = (float)dmax;
        r.emin[i] = (float)mne; r.erange[i] = (float)(mxe - mne);
        // StratusRendererFrontend.cpp:767-784: the cascade camera sk = (cx, cy, minZ) in light space and
        // the orthographic projection diag(2 / dk, 2 / dk, 1 / (maxZ - minZ)): clip = M p with
        // x = (right.p - cx) 2 / dk, y = (up.p - cy) 2 / dk, z = (fwd.p - minZ) / (maxZ - minZ) in [0, 1]
        // over the slice (window depth 0.5 .. 1; GL_DEPTH_CLAMP keeps casters outside it at 0 .. 1)
        double M[16] = {0};
        const double erange = (double)r.erange[i];
        for (int k = 0; k < 3; ++k) {
            M[k * 4 + 0] = right[k] * 2.0 / dk;
            M[k * 4 + 1] = up[k] * 2.0 / dk;
            M[k * 4 + 2] = fwd[k] / erange;
        }
        M[12] = -cx * 2.0 / dk;
        M[13] = -cy * 2.0 / dk;
        M[14] = -(double)r.emin[i] / erange;
        M[15] = 1.0;
        if (out) toFloat16(M, out->cascadeProjViews[i]);
        if (i > 0 && out) {
            // StratusRendererFrontend.cpp:794-803: plane along the view direction,
            // 0 at the start of cascade i and 1 at the end of cascade i-1
            const double scale = 1.0 / (ends[i - 1] - a);
            const double dnc = -(m.fwd[0] * m.eye[0] + m.fwd[1] * m.eye[1] + m.fwd[2] * m.eye[2]);
            out->cascadePlanes[i - 1][0] = (float)(m.fwd[0] * scale);
            out->cascadePlanes[i - 1][1] = (float)(m.fwd[1] * scale);
            out->cascadePlanes[i - 1][2] = (float)(m.fwd[2] * scale);
            out->cascadePlanes[i - 1][3] = (float)((dnc - a) * scale);
        }
    }
    s.csmResolution = res;
    if (out) {
        out->resolution = res;
        out->depth = nullptr;
        out->znear = (float)znear;
        out->zfar = (float)zfar;
        out->depthBias = 0.0f;
    }
}

template <class F>
void parallelFor(long n, F f) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || hw == 1) {
        for (long i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < hw; ++t)
        ts.emplace_back([=, &f]() {
            for (long i = t; i < n; i += hw) f(i);
        });
    for (auto& t : ts) t.join();
}

uint16_t floatToHalf(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int32_t exp = (int32_t)((x >> 23) & 0xFF) - 127 + 15;
    uint32_t mant = x & 0x7FFFFFu;
    if (((x >> 23) & 0xFF) == 0xFF) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u : 0u));
    if (exp >= 31) return (uint16_t)(sign | 0x7C00u);
    if (exp <= 0) {
        if (exp < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        const int shift = 14 - exp;
        uint32_t h = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)exp << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

bool checkCubes(const sgi_scene* s, const sgi_cube_atlases* c, int first, int count) {
    if (!s || !c) { sgi_set_error("sgi_scene_cubes: NULL argument"); return false; }
    if (c->resolution <= 0 ||
        (c->format != SGI_TEXEL_F32 && c->format != SGI_TEXEL_F16 && c->format != SGI_TEXEL_RGBA8 &&
         c->format != SGI_TEXEL_UNORM16) ||
        c->layers_per_atlas <= 0 || c->num_atlases <= 0 || c->num_atlases > SGI_MAX_TOTAL_SHADOW_ATLASES) {
        sgi_set_error("sgi_scene_cubes: bad atlas description");
        return false;
    }
    if (first < 0 || count < 0 || first + count > (int)s->vpls.size() ||
        (long)(first + count) > (long)c->layers_per_atlas * c->num_atlases) {
        sgi_set_error("sgi_scene_cubes: VPL range exceeds the scene's VPLs or the atlases");
        return false;
    }
    for (int a = 0; a < c->num_atlases; ++a)
        if (!c->atlas[a]) { sgi_set_error("sgi_scene_cubes: NULL atlas"); return false; }
    return true;
}

}  // namespace

extern "C" {

int sgi_scene_create(const sgi_scene_desc* desc, sgi_scene** out) {
    if (!desc || !out) { sgi_set_error("sgi_scene_create: NULL argument"); return SGI_E_INVALID; }
    if (desc->kind != SGI_SCENE_CORNELL && desc->kind != SGI_SCENE_CITY) {
        sgi_set_error("sgi_scene_create: unknown scene kind");
        return SGI_E_INVALID;
    }
    if (desc->num_vpls < 0 || desc->num_vpls > SGI_MAX_TOTAL_VPLS_BEFORE_CULLING) {
        sgi_set_error("sgi_scene_create: num_vpls out of range [0, 10000]");
        return SGI_E_INVALID;
    }
    if (desc->kind == SGI_SCENE_CITY && (desc->num_city_boxes < 0 || desc->city_half_extent <= 8.0f)) {
        sgi_set_error("sgi_scene_create: bad city parameters");
        return SGI_E_INVALID;
    }
    sgi_scene* s = new sgi_scene();
    s->desc = *desc;
    if (desc->kind == SGI_SCENE_CORNELL) buildCornell(*s);
    else buildCity(*s);
    sampleVpls(*s);
    if ((int)s->vpls.size() < desc->num_vpls) {
        sgi_set_error("sgi_scene_create: could not place all VPLs on sun-lit surfaces");
        delete s;
        return SGI_E_INVALID;
    }
    *out = s;
    return SGI_OK;
}

void sgi_scene_destroy(sgi_scene* s) {
    if (!s) return;
    if (s->freeDevice) s->freeDevice(s);
    delete s;
}

int sgi_scene_num_boxes(const sgi_scene* s) { return s ? (int)s->boxes.size() : 0; }

int sgi_scene_boxes(const sgi_scene* s, float* out) {
    if (!s || !out) { sgi_set_error("sgi_scene_boxes: NULL argument"); return SGI_E_INVALID; }
    static_assert(sizeof(sgi_rt::Box) == 16 * sizeof(float), "box record is 16 floats");
    std::memcpy(out, s->boxes.data(), s->boxes.size() * sizeof(sgi_rt::Box));
    return SGI_OK;
}

int sgi_scene_default_view(const sgi_scene* s, int32_t width, int32_t height, sgi_view* v) {
    if (!s || !v || width <= 0 || height <= 0) { sgi_set_error("sgi_scene_default_view: bad argument"); return SGI_E_INVALID; }
    std::memset(v, 0, sizeof(*v));
    v->width = width;
    v->height = height;
    v->znear = 1.0f;
    v->zfar = 1000.0f;
    if (s->desc.kind == SGI_SCENE_CORNELL) {
        v->eye[0] = 0.0f; v->eye[1] = 5.0f; v->eye[2] = 9.5f;
        v->yaw_deg = 0.0f;
        v->pitch_deg = -2.0f;
        v->fovy_deg = 60.0f;
        v->csm_resolution = 1024;
    } else {
        v->eye[0] = -55.0f; v->eye[1] = 22.0f; v->eye[2] = 75.0f;
        v->yaw_deg = 35.0f;
        v->pitch_deg = -12.0f;
        v->fovy_deg = 75.0f;
        v->csm_resolution = 1024;
    }
    return SGI_OK;
}

int sgi_scene_set_view(sgi_scene* s, const sgi_view* v, const sgi_view* prev, sgi_frame* frame, sgi_csm* csm) {
    if (!s || !v || v->width <= 0 || v->height <= 0 || v->csm_resolution <= 0 || v->znear <= 0.0f ||
        v->zfar <= v->znear || v->fovy_deg <= 0.0f || v->fovy_deg >= 180.0f) {
        sgi_set_error("sgi_scene_set_view: bad view");
        return SGI_E_INVALID;
    }
    if (prev && (prev->width != v->width || prev->height != v->height)) {
        sgi_set_error("sgi_scene_set_view: previous view has a different size");
        return SGI_E_INVALID;
    }
    const ViewMats m = buildView(*v);
    const ViewMats pm = buildView(prev ? *prev : *v);
    s->view = *v;
    s->hasView = true;
    sgi_rt::CameraRays& c = s->cam;
    toFloat16(m.invViewProj, c.invViewProj);
    toFloat16(m.viewProj, c.viewProj);
    toFloat16(pm.viewProj, c.prevViewProj);
    for (int k = 0; k < 3; ++k) c.eye[k] = (float)m.eye[k];
    c.width = v->width;
    c.height = v->height;
    c.farDist = v->zfar;
    c.reflectance = 0.05f;
    if (frame) {
        std::memset(frame, 0, sizeof(*frame));
        toFloat16(m.invViewProj, frame->invProjectionView);
        toFloat16(pm.invViewProj, frame->prevInvProjectionView);
        for (int k = 0; k < 3; ++k) {
            frame->viewPosition[k] = (float)m.eye[k];
            frame->infiniteLightDirection[k] = s->sunDir[k];
            frame->infiniteLightColor[k] = s->sunColor[k];
            frame->fogColor[k] = 0.5f;
        }
        frame->time = 0.0f;
        frame->framesPerSecond = 60.0f;
        frame->frameCount = 0;
        frame->minRoughness = 0.08f;
        frame->usePerceptualRoughness = 1;
        frame->minGiOcclusionFactor = 0.95f;
        frame->fogDensity = 0.0f;
        frame->emissionStrength = 0.0f;
    }
    if (csm) std::memset(csm, 0, sizeof(*csm));
    buildCsm(*s, *v, m, csm);
    return SGI_OK;
}

int sgi_scene_vpls(const sgi_scene* s, int32_t layersPerAtlas, sgi_vpl* vpls, sgi_atlas_entry* entries) {
    if (!s || layersPerAtlas <= 0 || layersPerAtlas > SGI_MAX_TOTAL_SHADOWS_PER_ATLAS) {
        sgi_set_error("sgi_scene_vpls: bad argument (layers_per_atlas in [1, 320])");
        return SGI_E_INVALID;
    }
    const int n = (int)s->vpls.size();
    if (vpls) std::memcpy(vpls, s->vpls.data(), sizeof(sgi_vpl) * n);
    if (entries)
        for (int j = 0; j < n; ++j) {
            entries[j].index = j / layersPerAtlas;
            entries[j].layer = j % layersPerAtlas;
        }
    return SGI_OK;
}

int sgi_scene_gbuffer_host(const sgi_scene* s, float* depth, float* normal, float* albedo, float* refl,
                           float* rma, float* velocity, float* id) {
    if (!s || !s->hasView) { sgi_set_error("sgi_scene_gbuffer_host: no view set"); return SGI_E_STATE; }
    const sgi_rt::SceneView sv = s->hostView();
    const sgi_rt::CameraRays cam = s->cam;
    const long W = cam.width, n = (long)cam.width * cam.height;
    parallelFor(n, [&](long p) {
        sgi_rt::GbufTexel g;
        sgi_rt::shadeGbufferTexel(sv, cam, (int)(p % W), (int)(p / W), g);
        if (depth) depth[p] = g.depth;
        if (normal) for (int k = 0; k < 3; ++k) normal[p * 3 + k] = g.normal[k];
        if (albedo) for (int k = 0; k < 4; ++k) albedo[p * 4 + k] = g.albedo[k];
        if (refl) for (int k = 0; k < 2; ++k) refl[p * 2 + k] = g.refl[k];
        if (rma) for (int k = 0; k < 3; ++k) rma[p * 3 + k] = g.rma[k];
        if (velocity) for (int k = 0; k < 2; ++k) velocity[p * 2 + k] = g.velocity[k];
        if (id) id[p] = g.id;
    });
    return SGI_OK;
}

int sgi_scene_cubes_host(const sgi_scene* s, const sgi_cube_atlases* c, int32_t first, int32_t count) {
    if (!checkCubes(s, c, first, count)) return SGI_E_INVALID;
    const sgi_rt::SceneView sv = s->hostView();
    const int res = c->resolution;
    const long perCube = 6L * res * res;
    parallelFor((long)count * perCube, [&](long q) {
        const int j = first + (int)(q / perCube);
        const long r = q % perCube;
        const int face = (int)(r / ((long)res * res));
        const int t = (int)((r / res) % res), i = (int)(r % res);
        const sgi_vpl& v = s->vpls[j];
        const int atlas = j / c->layers_per_atlas, layer = j % c->layers_per_atlas;
        const size_t off = (((size_t)layer * 6 + face) * res + t) * res + i;
        if (c->format == SGI_TEXEL_RGBA8) {
            ((uint32_t*)c->atlas[atlas])[off] = sgi_rt::cubeTexelColor(sv, v.position, v.farPlane, face, i, t, res);
            return;
        }
        const float val = sgi_rt::cubeTexelValue(sv, v.position, v.farPlane, face, i, t, res);
        if (c->format == SGI_TEXEL_F32) ((float*)c->atlas[atlas])[off] = val;
        else if (c->format == SGI_TEXEL_UNORM16) ((uint16_t*)c->atlas[atlas])[off] = sgi_rt::encodeUnorm16(val);
        else ((uint16_t*)c->atlas[atlas])[off] = floatToHalf(val);
    });
    return SGI_OK;
}

int sgi_scene_csm_host(const sgi_scene* s, float* depth) {
    if (!s || !s->hasView || !depth) { sgi_set_error("sgi_scene_csm_host: no view set or NULL"); return SGI_E_STATE; }
    const sgi_rt::SceneView sv = s->hostView();
    const int res = s->csmResolution;
    const long per = (long)res * res;
    parallelFor(4 * per, [&](long q) {
        const int c = (int)(q / per);
        const int j = (int)((q % per) / res), i = (int)(q % res);
        depth[q] = sgi_rt::csmTexelValue(sv, s->csm, c, i, j, res);
    });
    return SGI_OK;
}

}  // extern "C"
