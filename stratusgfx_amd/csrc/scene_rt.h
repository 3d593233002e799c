// Analytic scene ray caster shared by the host generator (g++) and the HIP
// generator kernels.  It produces the pass's inputs (G-buffer, VPL distance
// cube maps, sun cascades), which the reference rasterises with
// pbr_geometry_pass.fs, shadowVpl.fs and csm.fs; those producers are not part
// of the GI pass, so this is a plain ray caster over boxes rotated about Y,
// with a 2-D uniform grid over XZ for large box counts.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SGI_HD __host__ __device__
#else
#define SGI_HD
#endif

namespace sgi_rt {

struct Box {                // 64 B, one cache-friendly record per box
    float c[3];             // centre
    float cosy;             // rotation about +Y (local -> world), cos
    float h[3];             // half extents (local)
    float siny;             // rotation about +Y, sin
    float albedo[3];
    float roughness;
    float metallic;
    float id;               // mesh id written to the G-buffer (>= 1)
    float pad[2];
};

struct Grid {               // 2-D uniform grid over XZ, cells hold box indices
    float x0, z0, cell;     // origin and cell edge
    int nx, nz;
    float ymax;             // max box top of gridded boxes
    const int* cellStart;   // [nx*nz + 1]
    const int* cellItems;   // [cellStart[nx*nz]]
};

struct SceneView {
    const Box* boxes;
    int numAlways;          // boxes [0, numAlways) are tested by every ray
    int numBoxes;           // boxes [numAlways, numBoxes) are in the grid
    Grid grid;              // nx == 0 -> no grid
};

struct Hit {
    float t;
    int box;                // -1 = miss
    float n[3];             // world normal of the hit face
};

// Slab test of one box; updates hit when a nearer entry t in (tmin, hit.t) exists.
SGI_HD inline void intersectBox(const Box& b, int idx, const float o[3], const float d[3],
                                float tmin, Hit& hit) {
    const float dx = o[0] - b.c[0], dy = o[1] - b.c[1], dz = o[2] - b.c[2];
    // world -> local is R^T with R = [[c,0,s],[0,1,0],[-s,0,c]]
    const float lo[3] = {b.cosy * dx - b.siny * dz, dy, b.siny * dx + b.cosy * dz};
    const float ld[3] = {b.cosy * d[0] - b.siny * d[2], d[1], b.siny * d[0] + b.cosy * d[2]};
    float tn = -INFINITY, tf = INFINITY;
    int axis = -1;
    float sgn = 0.0f;
    for (int a = 0; a < 3; ++a) {
        if (ld[a] == 0.0f) {
            if (lo[a] < -b.h[a] || lo[a] > b.h[a]) return;
            continue;
        }
        const float inv = 1.0f / ld[a];
        float t0 = (-b.h[a] - lo[a]) * inv;
        float t1 = (b.h[a] - lo[a]) * inv;
        float s = -1.0f;  // entering through the -h face when t0 < t1
        if (t0 > t1) { const float tmp = t0; t0 = t1; t1 = tmp; s = 1.0f; }
        if (t0 > tn) { tn = t0; axis = a; sgn = s; }
        if (t1 < tf) tf = t1;
    }
    if (tn > tf || axis < 0) return;
    if (tn <= tmin || tn >= hit.t) return;
    float ln[3] = {0.0f, 0.0f, 0.0f};
    ln[axis] = sgn;
    hit.t = tn;
    hit.box = idx;
    hit.n[0] = b.cosy * ln[0] + b.siny * ln[2];
    hit.n[1] = ln[1];
    hit.n[2] = -b.siny * ln[0] + b.cosy * ln[2];
}

// Nearest hit along o + t d (d need not be normalised) with t in (tmin, tmax).
SGI_HD inline Hit traceScene(const SceneView& s, const float o[3], const float d[3],
                             float tmin, float tmax) {
    Hit hit;
    hit.t = tmax;
    hit.box = -1;
    hit.n[0] = hit.n[1] = hit.n[2] = 0.0f;
    for (int i = 0; i < s.numAlways; ++i) intersectBox(s.boxes[i], i, o, d, tmin, hit);
    const Grid& g = s.grid;
    if (g.nx <= 0) return hit;

    // Clip the ray to the grid's XZ extent and below ymax.
    const float gx1 = g.x0 + g.cell * g.nx, gz1 = g.z0 + g.cell * g.nz;
    float t0 = tmin, t1 = hit.t;
    const float lo[3] = {g.x0, -1e30f, g.z0};
    const float hi[3] = {gx1, g.ymax, gz1};
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0f) {
            if (o[a] < lo[a] || o[a] > hi[a]) return hit;
            continue;
        }
        const float inv = 1.0f / d[a];
        float ta = (lo[a] - o[a]) * inv, tb = (hi[a] - o[a]) * inv;
        if (ta > tb) { const float tmp = ta; ta = tb; tb = tmp; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
    }
    if (t0 > t1) return hit;

    // 2-D DDA over the cells crossed between t0 and t1.
    const float px = o[0] + d[0] * t0, pz = o[2] + d[2] * t0;
    int ix = (int)floorf((px - g.x0) / g.cell);
    int iz = (int)floorf((pz - g.z0) / g.cell);
    ix = ix < 0 ? 0 : (ix >= g.nx ? g.nx - 1 : ix);
    iz = iz < 0 ? 0 : (iz >= g.nz ? g.nz - 1 : iz);
    const int stepX = d[0] > 0.0f ? 1 : -1, stepZ = d[2] > 0.0f ? 1 : -1;
    float tMaxX = INFINITY, tMaxZ = INFINITY, tDX = INFINITY, tDZ = INFINITY;
    if (d[0] != 0.0f) {
        const float nextX = g.x0 + g.cell * (float)(ix + (stepX > 0 ? 1 : 0));
        tMaxX = (nextX - o[0]) / d[0];
        tDX = g.cell / fabsf(d[0]);
    }
    if (d[2] != 0.0f) {
        const float nextZ = g.z0 + g.cell * (float)(iz + (stepZ > 0 ? 1 : 0));
        tMaxZ = (nextZ - o[2]) / d[2];
        tDZ = g.cell / fabsf(d[2]);
    }
    for (int guard = 0; guard < 4 * (g.nx + g.nz) + 8; ++guard) {
        const int cidx = iz * g.nx + ix;
        for (int k = g.cellStart[cidx]; k < g.cellStart[cidx + 1]; ++k) {
            const int b = g.cellItems[k];
            intersectBox(s.boxes[b], b, o, d, tmin, hit);
        }
        const float tExit = tMaxX < tMaxZ ? tMaxX : tMaxZ;
        if (hit.t <= tExit || tExit > t1) break;
        if (tMaxX < tMaxZ) {
            ix += stepX;
            if (ix < 0 || ix >= g.nx) break;
            tMaxX += tDX;
        } else {
            iz += stepZ;
            if (iz < 0 || iz >= g.nz) break;
            tMaxZ += tDZ;
        }
    }
    return hit;
}

// Column-major 4x4 times (x, y, z, 1).
SGI_HD inline void xform4(const float m[16], const float p[3], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = m[r] * p[0] + m[4 + r] * p[1] + m[8 + r] * p[2] + m[12 + r];
}

// Direction of texel (i, j) of cube face f (GL major-axis table inverted).
SGI_HD inline void cubeTexelDir(int f, int i, int j, int res, float d[3]) {
    const float sc = 2.0f * ((float)i + 0.5f) / (float)res - 1.0f;
    const float tc = 2.0f * ((float)j + 0.5f) / (float)res - 1.0f;
    switch (f) {
        case 0: d[0] = 1.0f;  d[1] = -tc;   d[2] = -sc;   break;  // +X: sc=-rz, tc=-ry
        case 1: d[0] = -1.0f; d[1] = -tc;   d[2] = sc;    break;  // -X: sc=+rz, tc=-ry
        case 2: d[0] = sc;    d[1] = 1.0f;  d[2] = tc;    break;  // +Y: sc=+rx, tc=+rz
        case 3: d[0] = sc;    d[1] = -1.0f; d[2] = -tc;   break;  // -Y: sc=+rx, tc=-rz
        case 4: d[0] = sc;    d[1] = -tc;   d[2] = 1.0f;  break;  // +Z: sc=+rx, tc=-ry
        default: d[0] = -sc;  d[1] = -tc;   d[2] = -1.0f; break;  // -Z: sc=-rx, tc=-ry
    }
    const float inv = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] *= inv; d[1] *= inv; d[2] *= inv;
}

// Everything one G-buffer texel needs, from the camera's inverse view-projection.
struct GbufTexel {
    float depth;
    float normal[3];
    float albedo[4];
    float refl[2];
    float rma[3];
    float velocity[2];
    float id;
};

struct CameraRays {
    float invViewProj[16];   // NDC -> world
    float viewProj[16];      // world -> clip (depth)
    float prevViewProj[16];  // previous frame world -> clip (velocity)
    float eye[3];
    int width, height;
    float farDist;           // rays stop here (background)
    float reflectance;       // material reflectance (StratusMaterial.h:101 default 0.05)
};

SGI_HD inline void shadeGbufferTexel(const SceneView& s, const CameraRays& cam, int x, int y,
                                     GbufTexel& g) {
    const float u = ((float)x + 0.5f) / (float)cam.width;
    const float v = ((float)y + 0.5f) / (float)cam.height;
    const float ndc[3] = {2.0f * u - 1.0f, 2.0f * v - 1.0f, 1.0f};
    float w[4];
    xform4(cam.invViewProj, ndc, w);
    float d[3] = {w[0] / w[3] - cam.eye[0], w[1] / w[3] - cam.eye[1], w[2] / w[3] - cam.eye[2]};
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= len; d[1] /= len; d[2] /= len;
    const Hit h = traceScene(s, cam.eye, d, 0.0f, cam.farDist);
    for (int k = 0; k < 4; ++k) g.albedo[k] = 0.0f;
    g.normal[0] = g.normal[1] = g.normal[2] = 0.0f;
    g.refl[0] = g.refl[1] = 0.0f;
    g.rma[0] = g.rma[1] = g.rma[2] = 0.0f;
    g.velocity[0] = g.velocity[1] = 0.0f;
    g.id = 0.0f;
    if (h.box < 0) {  // cleared G-buffer: depth 1, everything else 0
        g.depth = 1.0f;
        return;
    }
    const Box& b = s.boxes[h.box];
    const float p[3] = {cam.eye[0] + d[0] * h.t, cam.eye[1] + d[1] * h.t, cam.eye[2] + d[2] * h.t};
    float clip[4];
    xform4(cam.viewProj, p, clip);
    g.depth = (clip[2] / clip[3]) * 0.5f + 0.5f;
    for (int k = 0; k < 3; ++k) g.normal[k] = h.n[k] * 0.5f + 0.5f;
    for (int k = 0; k < 3; ++k) g.albedo[k] = b.albedo[k];
    g.albedo[3] = 0.0f;  // emissive.r
    // pbr_geometry_pass.fs:129-132: reflectance mixed toward 0.8 by smoothness and metalness
    const float r = b.roughness, m = b.metallic;
    const float smooth = (1.0f - r) * 0.5f;
    const float base = cam.reflectance * (1.0f - smooth) + 0.8f * smooth;
    g.refl[0] = base * (1.0f - m) + 0.8f * m;
    g.refl[1] = 0.0f;    // emissive.g
    g.rma[0] = r;
    g.rma[1] = m;
    g.rma[2] = 0.0f;     // emissive.b
    float pc[4];
    xform4(cam.prevViewProj, p, pc);
    const float pu = (pc[0] / pc[3]) * 0.5f + 0.5f, pv = (pc[1] / pc[3]) * 0.5f + 0.5f;
    g.velocity[0] = u - pu;
    g.velocity[1] = v - pv;
    g.id = b.id;
}

// Distance (t) of the nearest hit from a VPL along a cube texel direction, as
// the stored value dist / farPlane clamped to 1 (a cleared depth attachment).
// A depth value as the GL_DEPTH_COMPONENT16 code the reference's VPL cubes hold:
// clamped to [0, 1], times 65535, rounded to nearest (ties to even; GL leaves
// ties to the implementation).
SGI_HD inline uint16_t encodeUnorm16(float v) {
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    return (uint16_t)rintf(v * 65535.0f);
}

SGI_HD inline float cubeTexelValue(const SceneView& s, const float pos[3], float farPlane,
                                   int face, int i, int j, int res) {
    float d[3];
    cubeTexelDir(face, i, j, res, d);
    const Hit h = traceScene(s, pos, d, 0.0f, farPlane);
    if (h.box < 0) return 1.0f;
    const float v = h.t / farPlane;
    return v < 1.0f ? v : 1.0f;
}

// Colour texel of a VPL's diffuse cube (shadowVpl.fs:37 writes the base colour
// of the nearest surface into the RGB8 attachment, StratusRendererBackend.cpp:
// 2419): UNORM8 r, g, b in bytes 0..2 (round to nearest), byte 3 = 255.  Where
// nothing lies within farPlane the layer keeps its clear value 0 (:1453-1454);
// the reference draws its skybox there, and these scenes have none.
SGI_HD inline uint32_t cubeTexelColor(const SceneView& s, const float pos[3], float farPlane,
                                      int face, int i, int j, int res) {
    float d[3];
    cubeTexelDir(face, i, j, res, d);
    const Hit h = traceScene(s, pos, d, 0.0f, farPlane);
    if (h.box < 0) return 0u;
    uint32_t rgba = 0xFF000000u;
    for (int k = 0; k < 3; ++k) {
        float a = s.boxes[h.box].albedo[k];
        a = a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
        rgba |= (uint32_t)(a * 255.0f + 0.5f) << (8 * k);
    }
    return rgba;
}

struct CsmRays {
    float right[3], up[3], fwd[3];  // light basis, fwd points away from the sun
    float cx[4], cy[4], dk[4];      // cascade centre (light x, y) and width
    float dmin[4], dmax[4];         // traced range along fwd (the scene's bounds)
    float emin[4], erange[4];       // the cascade's depth mapping: its slice's corners along fwd
};

// Window depth of the first surface seen by cascade c's texel (i, j), as the
// reference's cascade render stores it: NDC z = (e - emin) / erange with e the
// surface's position along fwd (StratusRendererFrontend.cpp:769-773), window
// z = 0.5 z + 0.5, clamped to [0, 1] by GL_DEPTH_CLAMP
// (StratusRendererBackend.cpp:1094); 1 (the clear) where nothing is hit.
SGI_HD inline float csmTexelValue(const SceneView& s, const CsmRays& r, int c, int i, int j, int res) {
    const float x = r.cx[c] + (((float)i + 0.5f) / (float)res * 2.0f - 1.0f) * 0.5f * r.dk[c];
    const float y = r.cy[c] + (((float)j + 0.5f) / (float)res * 2.0f - 1.0f) * 0.5f * r.dk[c];
    float o[3];
    for (int k = 0; k < 3; ++k) o[k] = r.right[k] * x + r.up[k] * y + r.fwd[k] * r.dmin[c];
    const float range = r.dmax[c] - r.dmin[c];
    const Hit h = traceScene(s, o, r.fwd, 0.0f, range);
    if (h.box < 0) return 1.0f;
    const float w = 0.5f * ((r.dmin[c] + h.t - r.emin[c]) / r.erange[c]) + 0.5f;
    return w < 0.0f ? 0.0f : (w > 1.0f ? 1.0f : w);
}

}  // namespace sgi_rt
