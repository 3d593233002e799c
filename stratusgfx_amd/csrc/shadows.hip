// The shadow-map producers of the pass's inputs: the VPL cube pre-pass and the
// sun cascades' depth render, from world-space triangles.
//
// VPL distance / diffuse cube producer — the reference's shadow pre-pass for
// virtual point lights (StratusRendererBackend.cpp:1426-1575): per updated
// light, the six faces rendered with perspective(90 deg, 1, near, far) x
// lookAt(pos, pos + axis, up) (:873-886, :1450), each face's depth layer
// cleared to 1 and its colour layer to 0 (:1453-1459), then every static
// triangle drawn with gl_FragDepth = saturate(|P - L| / farPlane) and colour =
// its base colour (shadowVpl.fs:17-37 through shadow.vs:22-33), depth test
// LESS, into layer * 6 + face of the light's atlas entry (:1511-1518).
//
// The six lookAt bases are signed permutations of the axes, so the face ray of
// texel (i, j) — the ray through its pixel centre — is exactly the GL
// major-axis table's direction with major component +-1 (window pixel (i, j)
// of face f is cube texel (i, j) of face f).  A fragment at that texel exists
// where the ray meets a triangle: with a, b, c the vertices relative to the
// light, e_k = d . (b x c, c x a, a x b) all of one sign, the view depth
// t = det(a, b, c) / (e0 + e1 + e2) within [near, far] (GL clipping), and the
// fragment's distance t |d|.  The depth test keeps the smallest stored value,
// the first triangle in draw order on ties (LESS), exactly as a per-texel
// minimum over the triangles in their order.
//
// MI355X structure: (1) one 1024-thread block per (light, face) compacts, in
// triangle order, the triangles that can overlap the face's frustum (ballot +
// block prefix; the order keeps the LESS tie rule); (2) one 256-thread block
// per (light, face, 32x32-texel region) walks that list in chunks of 256,
// keeps the triangles that can overlap its region's sub-frustum (again in
// order, via ballot into LDS, with their edge vectors precomputed once per
// chunk), and each lane depth-tests its 2x2 texels against every staged
// triangle with broadcast LDS reads.  Culling is conservative (a relative
// margin far above rounding); the per-texel test alone decides every value,
// in the restatement's operation order (-ffp-contract=off).
#include "launch.h"

namespace {

using namespace sgi_dev;

constexpr int kListBlock = 1024;
constexpr int kRegion = 32;   // texels per region edge
constexpr int kStage = 256;   // triangles staged per chunk

// Unnormalised face ray of texel (i, j): the GL major-axis table inverted.
__device__ __forceinline__ f3 faceRay(int f, int i, int j, int res) {
    const float sc = 2.0f * ((float)i + 0.5f) / (float)res - 1.0f;
    const float tc = 2.0f * ((float)j + 0.5f) / (float)res - 1.0f;
    switch (f) {
        case 0: return f3{1.0f, -tc, -sc};
        case 1: return f3{-1.0f, -tc, sc};
        case 2: return f3{sc, 1.0f, tc};
        case 3: return f3{sc, -1.0f, -tc};
        case 4: return f3{sc, -tc, 1.0f};
        default: return f3{-sc, -tc, -1.0f};
    }
}

// A point relative to the light in face f's coordinates: m along the face's
// axis, (s, t) with s / m, t / m the face's NDC-like coordinates (the table
// faceRay inverts).
__device__ __forceinline__ void faceCoords(int f, f3 p, float& m, float& s, float& t) {
    switch (f) {
        case 0: m = p.x; s = -p.z; t = -p.y; break;
        case 1: m = -p.x; s = p.z; t = -p.y; break;
        case 2: m = p.y; s = p.x; t = p.z; break;
        case 3: m = -p.y; s = p.x; t = -p.z; break;
        case 4: m = p.z; s = p.x; t = -p.y; break;
        default: m = -p.z; s = -p.x; t = -p.y; break;
    }
}

__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// True when no point of the triangle can lie inside the frustum
// s_lo m <= s <= s_hi m, t_lo m <= t <= t_hi m, near <= m <= far (all three
// vertices strictly beyond one plane, by a relative margin of 2^-12).
__device__ __forceinline__ bool outsideFrustum(const float m[3], const float s[3], const float t[3], float sLo,
                                               float sHi, float tLo, float tHi, float nearP, float farP) {
    bool o0 = true, o1 = true, o2 = true, o3 = true, o4 = true, o5 = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float mar = (fabsf(m[k]) + fabsf(s[k]) + fabsf(t[k])) * 0x1p-12f;
        o0 = o0 && (s[k] - sLo * m[k] < -mar);
        o1 = o1 && (sHi * m[k] - s[k] < -mar);
        o2 = o2 && (t[k] - tLo * m[k] < -mar);
        o3 = o3 && (tHi * m[k] - t[k] < -mar);
        o4 = o4 && (m[k] < nearP - mar);
        o5 = o5 && (m[k] > farP + mar);
    }
    return o0 || o1 || o2 || o3 || o4 || o5;
}

struct CubeJob {
    const float* pos;        // [count][3][3]
    const uint32_t* colors;  // [count] or NULL
    int count;
    const sgi_vpl* vpls;
    const int32_t* updates;  // [numUpdates] VPL indices
    int first;               // first update of this batch
    int batch;               // updates in this batch
    const sgi_atlas_entry* entries;
    float nearPlane;
    int cull;                // SGI_CULLING_* of the meshes
    int32_t* lists;          // [batch][6][count]
    int32_t* listCounts;     // [batch][6]
};

__device__ __forceinline__ f3 vertexRel(const float* pos, int k, int v, f3 L) {
    const float* q = pos + (size_t)k * 9 + v * 3;
    return f3{q[0] - L.x, q[1] - L.y, q[2] - L.z};
}

// The face culling of SetCullState (StratusRendererBackend.cpp:901-918) for a
// triangle seen from the light: with a, b, c relative to the light,
// V = a . (b x c) = (P - L) . n for n the winding's normal, so the triangle is
// counter-clockwise in window space (its front under glFrontFace(GL_CCW)) when
// V < 0.  The cube faces' views are rotations and their projections keep
// orientation, so window winding is the world winding seen from the light.
__device__ __forceinline__ bool culledFromLight(int cull, float V) {
    if (cull == SGI_CULLING_CCW) return !(V < 0.0f);
    if (cull == SGI_CULLING_CW) return !(V > 0.0f);
    return false;
}

// (1) per (light, face): the triangles that may overlap the face, in order.
__global__ __launch_bounds__(kListBlock) void cubeFaceListKernel(CubeJob job) {
    const int u = blockIdx.x / 6, f = blockIdx.x % 6;
    const sgi_vpl& v = job.vpls[job.updates[job.first + u]];
    const f3 L{v.position[0], v.position[1], v.position[2]};
    const float farP = v.farPlane;
    int32_t* list = job.lists + ((size_t)u * 6 + f) * job.count;
    __shared__ int waveCount[kListBlock / 64];
    __shared__ int base;
    if (threadIdx.x == 0) base = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c0 = 0; c0 < job.count; c0 += kListBlock) {
        const int k = c0 + (int)threadIdx.x;
        bool keep = false;
        if (k < job.count) {
            const f3 a = vertexRel(job.pos, k, 0, L), b = vertexRel(job.pos, k, 1, L), c = vertexRel(job.pos, k, 2, L);
            float m[3], s[3], t[3];
            faceCoords(f, a, m[0], s[0], t[0]);
            faceCoords(f, b, m[1], s[1], t[1]);
            faceCoords(f, c, m[2], s[2], t[2]);
            keep = !outsideFrustum(m, s, t, -1.0f, 1.0f, -1.0f, 1.0f, job.nearPlane, farP) &&
                   !culledFromLight(job.cull, dot(a, cross3(b, c)));
        }
        const uint64_t bal = __ballot(keep);
        if (lane == 0) waveCount[wave] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += waveCount[w];
        if (keep) list[off + __popcll(bal & ((1ull << lane) - 1ull))] = k;
        __syncthreads();
        if (threadIdx.x == 0) {
            int total = 0;
            for (int w = 0; w < kListBlock / 64; ++w) total += waveCount[w];
            base += total;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) job.listCounts[u * 6 + f] = base;
}

// Stored value of a fragment at v = dist / far (>= 0) as an ordered key: the
// depth attachment's own representation, so that key order is LESS's order.
template <int FMT>
__device__ __forceinline__ uint32_t depthKey(float v) {
    v = v < 1.0f ? v : 1.0f;  // saturate (v >= 0)
    if constexpr (FMT == SGI_TEXEL_F32) return __float_as_uint(v);
    else if constexpr (FMT == SGI_TEXEL_F16) return (uint32_t)__half_as_ushort(__float2half_rn(v));
    else return (uint32_t)rintf(v * 65535.0f);
}

struct StagedTri {
    float4 c0;  // b x c, V = det(a, b, c)
    float4 c1;  // c x a, colour bits
    float4 c2;  // a x b, unused
};

// (2) per (light, face, 32x32 region): depth-test every texel of the region.
template <int FMT>
__global__ __launch_bounds__(256) void cubeRasterKernel(CubeJob job, int res, int regionsPerEdge, CubeArgs dist,
                                                        CubeArgs diffuse, int withColor) {
    const int perFace = regionsPerEdge * regionsPerEdge;
    const int r = blockIdx.x % perFace, f = (blockIdx.x / perFace) % 6, u = blockIdx.x / (perFace * 6);
    const int rx = r % regionsPerEdge, ry = r / regionsPerEdge;
    const sgi_vpl& v = job.vpls[job.updates[job.first + u]];
    const f3 L{v.position[0], v.position[1], v.position[2]};
    const float farP = v.farPlane;
    const int32_t* list = job.lists + ((size_t)u * 6 + f) * job.count;
    const int n = job.listCounts[u * 6 + f];
    // the region's texel-edge bounds in the face's NDC-like coordinates
    const int x0 = rx * kRegion, y0 = ry * kRegion;
    const float sLo = 2.0f * (float)x0 / (float)res - 1.0f, sHi = 2.0f * (float)min(x0 + kRegion, res) / (float)res - 1.0f;
    const float tLo = 2.0f * (float)y0 / (float)res - 1.0f, tHi = 2.0f * (float)min(y0 + kRegion, res) / (float)res - 1.0f;
    // this lane's 2x2 texels
    const int lx = x0 + ((int)threadIdx.x % 16) * 2, ly = y0 + ((int)threadIdx.x / 16) * 2;
    f3 d[4];
    float dlen[4];
    uint32_t best[4], col[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        d[q] = faceRay(f, lx + (q & 1), ly + (q >> 1), res);
        dlen[q] = __builtin_sqrtf(d[q].x * d[q].x + d[q].y * d[q].y + d[q].z * d[q].z);
        best[q] = depthKey<FMT>(1.0f);  // the depth clear (:1455-1456)
        col[q] = 0u;                    // the colour clear (:1453-1454)
    }
    __shared__ StagedTri stage[kStage];
    __shared__ int waveCount[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c0 = 0; c0 < n; c0 += kStage) {
        const int idx = c0 + (int)threadIdx.x;
        bool keep = false;
        f3 a, b, c;
        if (idx < n) {
            const int k = list[idx];
            a = vertexRel(job.pos, k, 0, L);
            b = vertexRel(job.pos, k, 1, L);
            c = vertexRel(job.pos, k, 2, L);
            float m[3], s[3], t[3];
            faceCoords(f, a, m[0], s[0], t[0]);
            faceCoords(f, b, m[1], s[1], t[1]);
            faceCoords(f, c, m[2], s[2], t[2]);
            keep = !outsideFrustum(m, s, t, sLo, sHi, tLo, tHi, job.nearPlane, farP);
        }
        const uint64_t bal = __ballot(keep);
        if (lane == 0) waveCount[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            off += w < wave ? waveCount[w] : 0;
            total += waveCount[w];
        }
        if (keep) {
            const f3 C0 = cross3(b, c), C1 = cross3(c, a), C2 = cross3(a, b);
            const float V = dot(a, C0);
            const uint32_t rgb = withColor ? job.colors[list[idx]] : 0u;
            StagedTri st;
            st.c0 = make_float4(C0.x, C0.y, C0.z, V);
            st.c1 = make_float4(C1.x, C1.y, C1.z, __uint_as_float(rgb));
            st.c2 = make_float4(C2.x, C2.y, C2.z, 0.0f);
            stage[off + __popcll(bal & ((1ull << lane) - 1ull))] = st;
        }
        __syncthreads();
        for (int q = 0; q < total; ++q) {
            const StagedTri st = stage[q];
            const f3 C0{st.c0.x, st.c0.y, st.c0.z}, C1{st.c1.x, st.c1.y, st.c1.z}, C2{st.c2.x, st.c2.y, st.c2.z};
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const float e0 = dot(d[p], C0), e1 = dot(d[p], C1), e2 = dot(d[p], C2);
                const bool in = (e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) || (e0 <= 0.0f && e1 <= 0.0f && e2 <= 0.0f);
                const float den = e0 + e1 + e2;
                if (!in || den == 0.0f) continue;
                const float tv = st.c0.w / den;
                if (!(tv >= job.nearPlane && tv <= farP)) continue;  // GL clipping to [near, far]
                const uint32_t key = depthKey<FMT>((tv * dlen[p]) / farP);
                if (key < best[p]) {
                    best[p] = key;
                    col[p] = __float_as_uint(st.c1.w);
                }
            }
        }
        __syncthreads();
    }
    const sgi_atlas_entry e = job.entries[job.updates[job.first + u]];
    // an entry outside the atlases names no cube: nothing is written for it
    if (e.index < 0 || e.index >= dist.numAtlases || e.layer < 0 || e.layer >= dist.layers) return;
    void* zb = const_cast<void*>(cubeLayerBase(dist, e.index, e.layer, texelBytes(FMT)));
    uint32_t* cb = withColor ? (uint32_t*)const_cast<void*>(cubeLayerBase(diffuse, e.index, e.layer, 4)) : nullptr;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int i = lx + (p & 1), j = ly + (p >> 1);
        if (i >= res || j >= res) continue;
        const size_t off = ((size_t)f * res + j) * res + i;
        if constexpr (FMT == SGI_TEXEL_F32) ((float*)zb)[off] = __uint_as_float(best[p]);
        else ((uint16_t*)zb)[off] = (uint16_t)best[p];
        if (cb) cb[off] = col[p] | 0xFF000000u;
    }
}

// ---- sun cascades: RenderCSMDepth_ (StratusRendererBackend.cpp:1081-1158) with csm.vs:29-41
//
// Per cascade c, every triangle is transformed by cascadeProjViews[c]
// (projectionViewRender, an orthographic light projection) and rasterised
// into layer c: the window depth z * 0.5 + 0.5 interpolated at the texel
// centre, plus glPolygonOffset(factor, 0)'s slope term factor * m with
// m = max(|dz_w/dx_w|, |dz_w/dy_w|) (the form GL allows in place of the
// gradient's length), clamped to [0, 1] by GL_DEPTH_CLAMP (which also
// disables near / far clipping), depth test LESS against the clear 1.  The
// pass draws with each mesh's culling reversed (RenderImmediate_(..., true),
// :1029-1042): CCW meshes cull their light-facing (counter-clockwise) sides.
// Coverage: the 2-D edge functions of the NDC vertices at the texel centre all
// of one sign (a texel centre on an edge belongs to both triangles, the first
// in order winning LESS ties).
struct CsmJob {
    const float* pos;
    int count;
    int keepSign;        // +1: keep counter-clockwise (NDC area > 0), -1: clockwise, 0: both
    float slope;         // glPolygonOffset factor
    int res;
    float m[SGI_NUM_CASCADES][16];
    float* tris;         // [4][count][10]: NDC x, y, z of the 3 vertices and the polygon offset
    int32_t* counts;     // [4]
};

__device__ __forceinline__ float edgeFn(float ax, float ay, float bx, float by, float sx, float sy) {
    return (bx - ax) * (sy - ay) - (by - ay) * (sx - ax);
}

// (1) per cascade: the triangles that face the kept way and overlap the
// cascade's window, in order, with their NDC vertices and polygon offset.
__global__ __launch_bounds__(kListBlock) void csmListKernel(CsmJob job) {
    const int c = blockIdx.x;
    float* out = job.tris + (size_t)c * job.count * 10;
    __shared__ int waveCount[kListBlock / 64];
    __shared__ int base;
    if (threadIdx.x == 0) base = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c0 = 0; c0 < job.count; c0 += kListBlock) {
        const int k = c0 + (int)threadIdx.x;
        bool keep = false;
        float v[10];
        if (k < job.count) {
            const float* q = job.pos + (size_t)k * 9;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                float cl[4];
                mulMat4(job.m[c], q[3 * i], q[3 * i + 1], q[3 * i + 2], 1.0f, cl);
                v[3 * i] = cl[0] / cl[3];
                v[3 * i + 1] = cl[1] / cl[3];
                v[3 * i + 2] = cl[2] / cl[3];
            }
            const float ex1 = v[3] - v[0], ey1 = v[4] - v[1], ez1 = v[5] - v[2];
            const float ex2 = v[6] - v[0], ey2 = v[7] - v[1], ez2 = v[8] - v[2];
            const float det = ex1 * ey2 - ex2 * ey1;  // twice the signed NDC area
            const bool facing = job.keepSign > 0 ? det > 0.0f : (job.keepSign < 0 ? det < 0.0f : det != 0.0f);
            const float mnx = fminf(v[0], fminf(v[3], v[6])), mxx = fmaxf(v[0], fmaxf(v[3], v[6]));
            const float mny = fminf(v[1], fminf(v[4], v[7])), mxy = fmaxf(v[1], fmaxf(v[4], v[7]));
            const bool inWindow = mxx >= -1.001f && mnx <= 1.001f && mxy >= -1.001f && mny <= 1.001f;
            keep = facing && inWindow;
            if (keep) {
                const float A = (ez1 * ey2 - ez2 * ey1) / det;  // dz/dx of the NDC plane
                const float B = (ex1 * ez2 - ex2 * ez1) / det;  // dz/dy
                // window units: z_w = z / 2 + 1/2 and x_w = (x + 1) res / 2, so dz_w/dx_w = A / res
                v[9] = job.slope * (fmaxf(fabsf(A), fabsf(B)) / (float)job.res);
            }
        }
        const uint64_t bal = __ballot(keep);
        if (lane == 0) waveCount[wave] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += waveCount[w];
        if (keep) {
            float* o = out + (size_t)(off + __popcll(bal & ((1ull << lane) - 1ull))) * 10;
#pragma unroll
            for (int i = 0; i < 10; ++i) o[i] = v[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int total = 0;
            for (int w = 0; w < kListBlock / 64; ++w) total += waveCount[w];
            base += total;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) job.counts[c] = base;
}

// (2) per (cascade, 32x32 region): rasterise the list's triangles that can
// overlap the region (staged through LDS in order) at every texel centre.
__global__ __launch_bounds__(256) void csmRasterKernel(CsmJob job, int regionsPerEdge, float* __restrict__ depth) {
    const int perLayer = regionsPerEdge * regionsPerEdge;
    const int c = blockIdx.x / perLayer, r = blockIdx.x % perLayer;
    const int x0 = (r % regionsPerEdge) * kRegion, y0 = (r / regionsPerEdge) * kRegion;
    const int res = job.res;
    const float lox = 2.0f * (float)x0 / (float)res - 1.0f, hix = 2.0f * (float)min(x0 + kRegion, res) / (float)res - 1.0f;
    const float loy = 2.0f * (float)y0 / (float)res - 1.0f, hiy = 2.0f * (float)min(y0 + kRegion, res) / (float)res - 1.0f;
    const float* list = job.tris + (size_t)c * job.count * 10;
    const int n = job.counts[c];
    const int lx = x0 + ((int)threadIdx.x % 16) * 2, ly = y0 + ((int)threadIdx.x / 16) * 2;
    float sx[4], sy[4], best[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        sx[p] = 2.0f * ((float)(lx + (p & 1)) + 0.5f) / (float)res - 1.0f;
        sy[p] = 2.0f * ((float)(ly + (p >> 1)) + 0.5f) / (float)res - 1.0f;
        best[p] = 1.0f;  // the depth clear
    }
    __shared__ float stage[kStage][10];
    __shared__ int waveCount[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float mar = 1e-3f;
    for (int c0 = 0; c0 < n; c0 += kStage) {
        const int idx = c0 + (int)threadIdx.x;
        bool keep = false;
        float v[10];
        if (idx < n) {
#pragma unroll
            for (int i = 0; i < 10; ++i) v[i] = list[(size_t)idx * 10 + i];
            const float mnx = fminf(v[0], fminf(v[3], v[6])), mxx = fmaxf(v[0], fmaxf(v[3], v[6]));
            const float mny = fminf(v[1], fminf(v[4], v[7])), mxy = fmaxf(v[1], fmaxf(v[4], v[7]));
            keep = mxx >= lox - mar && mnx <= hix + mar && mxy >= loy - mar && mny <= hiy + mar;
        }
        const uint64_t bal = __ballot(keep);
        if (lane == 0) waveCount[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            off += w < wave ? waveCount[w] : 0;
            total += waveCount[w];
        }
        if (keep) {
            float* o = stage[off + __popcll(bal & ((1ull << lane) - 1ull))];
#pragma unroll
            for (int i = 0; i < 10; ++i) o[i] = v[i];
        }
        __syncthreads();
        for (int q = 0; q < total; ++q) {
            const float* t = stage[q];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const float w0 = edgeFn(t[3], t[4], t[6], t[7], sx[p], sy[p]);
                const float w1 = edgeFn(t[6], t[7], t[0], t[1], sx[p], sy[p]);
                const float w2 = edgeFn(t[0], t[1], t[3], t[4], sx[p], sy[p]);
                const bool in = (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) || (w0 <= 0.0f && w1 <= 0.0f && w2 <= 0.0f);
                const float sum = w0 + w1 + w2;
                if (!in || sum == 0.0f) continue;
                const float z = (w0 * t[2] + w1 * t[5] + w2 * t[8]) / sum;
                float dw = z * 0.5f + 0.5f + t[9];
                dw = dw < 0.0f ? 0.0f : (dw > 1.0f ? 1.0f : dw);  // GL_DEPTH_CLAMP
                if (dw < best[p]) best[p] = dw;
            }
        }
        __syncthreads();
    }
    float* layer = depth + (size_t)c * res * res;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int i = lx + (p & 1), j = ly + (p >> 1);
        if (i < res && j < res) layer[(size_t)j * res + i] = best[p];
    }
}

}  // namespace

namespace sgi_launch {

size_t csmScratchFloats(int count) { return (size_t)SGI_NUM_CASCADES * count * 10 + SGI_NUM_CASCADES; }

hipError_t renderCsm(const float* pos, int count, int keepSign, float slope, int res,
                     const float (*projViews)[16], float* scratch, float* depth, hipStream_t s) {
    CsmJob job{};
    job.pos = pos;
    job.count = count;
    job.keepSign = keepSign;
    job.slope = slope;
    job.res = res;
    for (int c = 0; c < SGI_NUM_CASCADES; ++c)
        for (int k = 0; k < 16; ++k) job.m[c][k] = projViews[c][k];
    job.tris = scratch;
    job.counts = (int32_t*)(scratch + (size_t)SGI_NUM_CASCADES * count * 10);
    hipLaunchKernelGGL(csmListKernel, dim3(SGI_NUM_CASCADES), dim3(kListBlock), 0, s, job);
    const int rpe = (res + kRegion - 1) / kRegion;
    hipLaunchKernelGGL(csmRasterKernel, dim3((unsigned)(SGI_NUM_CASCADES * rpe * rpe)), dim3(256), 0, s, job, rpe,
                       depth);
    return hipGetLastError();
}

size_t cubeListInts(int batch, int count) { return (size_t)batch * 6 * (size_t)count + (size_t)batch * 6; }

hipError_t renderVplCubes(const float* pos, const uint32_t* colors, int count, int cull, const sgi_vpl* vpls,
                          const int32_t* updates, int numUpdates, const sgi_atlas_entry* entries, float nearPlane,
                          const sgi_dev::CubeArgs& dist, int format, const sgi_dev::CubeArgs& diffuse,
                          bool withDiffuse, int batchMax, int32_t* scratch, hipStream_t s) {
    const int res = dist.res;
    const int rpe = (res + kRegion - 1) / kRegion;
    for (int first = 0; first < numUpdates; first += batchMax) {
        const int batch = min(batchMax, numUpdates - first);
        CubeJob job{pos, colors, count, vpls, updates, first, batch, entries, nearPlane, cull, scratch,
                    scratch + (size_t)batch * 6 * (size_t)count};
        hipLaunchKernelGGL(cubeFaceListKernel, dim3((unsigned)(batch * 6)), dim3(kListBlock), 0, s, job);
        const dim3 grid((unsigned)(batch * 6 * rpe * rpe));
        const int wc = withDiffuse ? 1 : 0;
        if (format == SGI_TEXEL_F32)
            hipLaunchKernelGGL(cubeRasterKernel<SGI_TEXEL_F32>, grid, dim3(256), 0, s, job, res, rpe, dist, diffuse, wc);
        else if (format == SGI_TEXEL_F16)
            hipLaunchKernelGGL(cubeRasterKernel<SGI_TEXEL_F16>, grid, dim3(256), 0, s, job, res, rpe, dist, diffuse, wc);
        else
            hipLaunchKernelGGL(cubeRasterKernel<SGI_TEXEL_UNORM16>, grid, dim3(256), 0, s, job, res, rpe, dist, diffuse,
                               wc);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace sgi_launch
