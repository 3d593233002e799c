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
    int32_t* lists;          // [batch][6][count]
    int32_t* listCounts;     // [batch][6]
};

__device__ __forceinline__ f3 vertexRel(const float* pos, int k, int v, f3 L) {
    const float* q = pos + (size_t)k * 9 + v * 3;
    return f3{q[0] - L.x, q[1] - L.y, q[2] - L.z};
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
            float m[3], s[3], t[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) faceCoords(f, vertexRel(job.pos, k, q, L), m[q], s[q], t[q]);
            keep = !outsideFrustum(m, s, t, -1.0f, 1.0f, -1.0f, 1.0f, job.nearPlane, farP);
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

}  // namespace

namespace sgi_launch {

size_t cubeListInts(int batch, int count) { return (size_t)batch * 6 * (size_t)count + (size_t)batch * 6; }

hipError_t renderVplCubes(const float* pos, const uint32_t* colors, int count, const sgi_vpl* vpls,
                          const int32_t* updates, int numUpdates, const sgi_atlas_entry* entries, float nearPlane,
                          const sgi_dev::CubeArgs& dist, int format, const sgi_dev::CubeArgs& diffuse,
                          bool withDiffuse, int batchMax, int32_t* scratch, hipStream_t s) {
    const int res = dist.res;
    const int rpe = (res + kRegion - 1) / kRegion;
    for (int first = 0; first < numUpdates; first += batchMax) {
        const int batch = min(batchMax, numUpdates - first);
        CubeJob job{pos, colors, count, vpls, updates, first, batch, entries, nearPlane, scratch,
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
