// GPU LBVH build for gfx950 (replaces the driver BLAS/TLAS build of
// DeviceScene::createAccelerationStructure, src/Scene.cpp:448-547).
//
//   1. world-space triangles in primitive order (light triangles first, then
//      each object instance's triangles transformed by its matrix)
//   2. centroid bounds: per-block min/max then one final block
//   3. 63-bit Morton codes of quantised centroids
//   4. stable LSD radix sort of (code, primitive) pairs: 8 passes of 8-bit
//      digits; per block a 256-bin histogram, one exclusive scan over
//      (digit, block), then a scatter whose in-block rank comes from 8 wave
//      ballots (equal-digit lane masks) plus per-wave digit counts in LDS
//   5. Karras topology, one thread per internal node
//   6. bottom-up refit: one thread per leaf climbs; the second arrival at a
//      node (device-scope atomic counter) merges both child boxes
// Every step is integer or exact min/max work, so the result is bit-identical
// to the CPU restatement of the same algorithm.
#include "rt_internal.h"
#include <algorithm>
#include <cstring>

using namespace rtd;

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

__device__ __forceinline__ uint64_t expandBits21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// 1. world triangles. instPrimOffset[k] = first primitive of instance k.
__global__ void k_world_tris(DevScene S, const uint32_t* __restrict__ instPrimOffset, uint32_t numInstances,
                             float4* __restrict__ primTris) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.numPrims) return;
    f3 v[3];
    uint32_t inst, prim;
    if (p < S.numLights) {
        const rt_triangle_light& L = S.lights[p];
        v[0] = ld3(L.v0); v[1] = ld3(L.v1); v[2] = ld3(L.v2);
        inst = 0; prim = p;
    } else {
        // binary search: last k with instPrimOffset[k] <= p
        uint32_t lo = 0, hi = numInstances;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (instPrimOffset[mid] <= p) lo = mid; else hi = mid;
        }
        const DevInstance& I = S.instances[lo];
        prim = p - instPrimOffset[lo];
        inst = lo + 1;
        for (int j = 0; j < 3; j++) {
            const rt_mesh_vertex& mv = S.vertices[S.indices[I.indexOffset + prim * 3 + j]];
            v[j] = colMul4(I.m, ld3(mv.pos), 1.0f);
        }
    }
    float4* o = primTris + 3 * (size_t)p;
    o[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(inst));
    o[1] = make_float4(v[1].x, v[1].y, v[1].z, __uint_as_float(prim));
    o[2] = make_float4(v[2].x, v[2].y, v[2].z, 0.0f);
}

__device__ __forceinline__ f3 centroid(const float4* t) {
    f3 a = mk3(t[0].x, t[0].y, t[0].z), b = mk3(t[1].x, t[1].y, t[1].z), c = mk3(t[2].x, t[2].y, t[2].z);
    return (a + b + c) * (1.0f / 3.0f);
}

// 2. bounds: partial[block] = (lo.xyz, hi.xyz)
__global__ void k_bounds_partial(const float4* __restrict__ primTris, uint32_t n, float* __restrict__ partial) {
    __shared__ float sm[6][kBlock];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < n; p += gridDim.x * kBlock) {
        f3 c = centroid(primTris + 3 * (size_t)p);
        lo[0] = fminf(lo[0], c.x); lo[1] = fminf(lo[1], c.y); lo[2] = fminf(lo[2], c.z);
        hi[0] = fmaxf(hi[0], c.x); hi[1] = fmaxf(hi[1], c.y); hi[2] = fmaxf(hi[2], c.z);
    }
    for (int a = 0; a < 3; a++) { sm[a][threadIdx.x] = lo[a]; sm[3 + a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int a = 0; a < 3; a++) {
                sm[a][threadIdx.x] = fminf(sm[a][threadIdx.x], sm[a][threadIdx.x + s]);
                sm[3 + a][threadIdx.x] = fmaxf(sm[3 + a][threadIdx.x], sm[3 + a][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ void k_bounds_final(const float* __restrict__ partial, uint32_t nb, float* __restrict__ box) {
    __shared__ float sm[6][kBlock];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < nb; b += kBlock)
        for (int a = 0; a < 3; a++) {
            v[a] = fminf(v[a], partial[b * 6 + a]);
            v[3 + a] = fmaxf(v[3 + a], partial[b * 6 + 3 + a]);
        }
    for (int a = 0; a < 6; a++) sm[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                sm[a][threadIdx.x] = fminf(sm[a][threadIdx.x], sm[a][threadIdx.x + s]);
                sm[3 + a][threadIdx.x] = fmaxf(sm[3 + a][threadIdx.x], sm[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = sm[threadIdx.x][0];
}

// 3. Morton codes
__global__ void k_morton(const float4* __restrict__ primTris, uint32_t n, const float* __restrict__ box,
                         unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float lo[3] = {box[0], box[1], box[2]};
    float sc[3];
    for (int a = 0; a < 3; a++) {
        float ext = box[3 + a] - lo[a];
        sc[a] = ext > 0.0f ? 2097151.0f / ext : 0.0f;
    }
    f3 c = centroid(primTris + 3 * (size_t)p);
    float cc[3] = {c.x, c.y, c.z};
    uint64_t q[3];
    for (int a = 0; a < 3; a++) {
        float f = (cc[a] - lo[a]) * sc[a];
        uint32_t u = (uint32_t)f;
        q[a] = u > 2097151u ? 2097151u : u;
    }
    keys[p] = (expandBits21(q[0]) << 2) | (expandBits21(q[1]) << 1) | expandBits21(q[2]);
    vals[p] = p;
}

// 4. radix sort
__global__ void k_radix_hist(const unsigned long long* __restrict__ keys, uint32_t n, int shift, uint32_t nBlocks,
                             uint32_t* __restrict__ blockHist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kTile;
    for (int k = 0; k < kItems; k++) {
        uint32_t i = base + k * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    blockHist[threadIdx.x * nBlocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of a uint32 array in one block of 1024 threads
__global__ void k_scan_single(uint32_t* __restrict__ data, uint32_t n) {
    __shared__ uint32_t part[1024];
    uint32_t per = (n + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += data[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = (threadIdx.x >= (uint32_t)off) ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;  // exclusive prefix of this chunk
    for (uint32_t i = b; i < e; i++) { uint32_t v = data[i]; data[i] = run; run += v; }
}

__global__ void k_radix_scatter(const unsigned long long* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                uint32_t n, int shift, uint32_t nBlocks, const uint32_t* __restrict__ blockOffset,
                                unsigned long long* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t waveCnt[kBlock / 64][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    running[tid] = blockOffset[tid * nBlocks + blockIdx.x];
    for (int w = 0; w < kBlock / 64; w++) waveCnt[w][tid] = 0;
    __syncthreads();
    const unsigned long long ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint32_t base = blockIdx.x * kTile;
    for (int k = 0; k < kItems; k++) {
        uint32_t i = base + k * kBlock + tid;
        bool valid = i < n;
        unsigned long long key = valid ? keysIn[i] : 0ull;
        uint32_t val = valid ? valsIn[i] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 0xffu;
        unsigned long long mask = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            unsigned long long bal = __ballot(bit);
            mask &= bit ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(mask & ltMask);
        if (valid && rank == 0) waveCnt[wave][d] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (valid) {
            uint32_t off = running[d] + rank;
            for (uint32_t w = 0; w < wave; w++) off += waveCnt[w][d];
            keysOut[off] = key;
            valsOut[off] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kBlock / 64; w++) { tot += waveCnt[w][tid]; waveCnt[w][tid] = 0; }
        running[tid] += tot;
        __syncthreads();
    }
}

// 5. Karras topology
__device__ __forceinline__ int kdelta(const unsigned long long* c, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    unsigned long long a = c[i], b = c[j];
    if (a == b) {
        uint32_t x = (uint32_t)i ^ (uint32_t)j;
        return 64 + (x == 0 ? 32 : __clz((int)x));
    }
    unsigned long long x = a ^ b;
    return __clzll((long long)x);
}

__global__ void k_karras(const unsigned long long* __restrict__ codes, int n, float4* __restrict__ nodes,
                         uint32_t* __restrict__ parentI, uint32_t* __restrict__ parentL) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(codes, n, i, i - d);
    int lmax = 2;
    while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(codes, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    uint32_t left = (min(i, j) == gamma) ? (RT_LEAF | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t right = (max(i, j) == gamma + 1) ? (RT_LEAF | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    nodes[4 * (size_t)i + 3] = make_float4(__uint_as_float(left), __uint_as_float(right), 0.0f, 0.0f);
    if (left & RT_LEAF) parentL[left & ~RT_LEAF] = (uint32_t)i; else parentI[left] = (uint32_t)i;
    if (right & RT_LEAF) parentL[right & ~RT_LEAF] = (uint32_t)i; else parentI[right] = (uint32_t)i;
    if (i == 0) parentI[0] = 0xffffffffu;
}

// leaves in Morton order
__global__ void k_gather_leaves(const float4* __restrict__ primTris, const uint32_t* __restrict__ order, uint32_t n,
                                float4* __restrict__ leafTris) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* s = primTris + 3 * (size_t)order[i];
    float4* d = leafTris + 3 * (size_t)i;
    d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
}

// node slot layout: n0 = (lmin.xyz, lmax.x) n1 = (lmax.yz, rmin.xy) n2 = (rmin.z, rmax.xyz)
__device__ __forceinline__ void writeSlot(float* nf, bool right, const float* lo, const float* hi) {
    int o = right ? 6 : 0;
    nf[o + 0] = lo[0]; nf[o + 1] = lo[1]; nf[o + 2] = lo[2];
    nf[o + 3] = hi[0]; nf[o + 4] = hi[1]; nf[o + 5] = hi[2];
}

// 6. refit
__global__ void k_refit(const float4* __restrict__ leafTris, uint32_t n, float4* nodes,
                        const uint32_t* __restrict__ parentI, const uint32_t* __restrict__ parentL, uint32_t* counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* t = leafTris + 3 * (size_t)i;
    float lo[3] = {fminf(fminf(t[0].x, t[1].x), t[2].x), fminf(fminf(t[0].y, t[1].y), t[2].y),
                   fminf(fminf(t[0].z, t[1].z), t[2].z)};
    float hi[3] = {fmaxf(fmaxf(t[0].x, t[1].x), t[2].x), fmaxf(fmaxf(t[0].y, t[1].y), t[2].y),
                   fmaxf(fmaxf(t[0].z, t[1].z), t[2].z)};
    uint32_t ref = RT_LEAF | i;
    uint32_t p = parentL[i];
    while (p != 0xffffffffu) {
        float* nf = reinterpret_cast<float*>(nodes + 4 * (size_t)p);
        uint32_t leftRef = __float_as_uint(nf[12]);
        writeSlot(nf, leftRef != ref, lo, hi);
        __threadfence();
        uint32_t old = atomicAdd(&counters[p], 1u);
        if (old == 0) return;
        __threadfence();
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(__hip_atomic_load(&nf[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                          __hip_atomic_load(&nf[6 + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            hi[a] = fmaxf(__hip_atomic_load(&nf[3 + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                          __hip_atomic_load(&nf[9 + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
        ref = p;
        p = parentI[p];
    }
}

// depth: each leaf climbs to the root counting internal nodes
__global__ void k_depth(const uint32_t* __restrict__ parentI, const uint32_t* __restrict__ parentL, uint32_t n,
                        uint32_t* maxDepth) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t d = 0, p = parentL[i];
    while (p != 0xffffffffu && d < 4096u) { d++; p = parentI[p]; }
    atomicMax(maxDepth, d);
}

// 7. collapse to a 4-wide tree. A binary node's candidate slots: its two
// children, then twice (while fewer than four) the internal slot with the
// largest surface area is replaced in place by its two children.
struct Slot4 { uint32_t ref; float lo[3], hi[3]; };

__device__ __forceinline__ float boxArea(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
__device__ __forceinline__ void readSlots(const float4* nodes, uint32_t n, Slot4& a, Slot4& b) {
    const float* f = reinterpret_cast<const float*>(nodes + 4 * (size_t)n);
    a.ref = __float_as_uint(f[12]); b.ref = __float_as_uint(f[13]);
    for (int k = 0; k < 3; k++) { a.lo[k] = f[k]; a.hi[k] = f[3 + k]; b.lo[k] = f[6 + k]; b.hi[k] = f[9 + k]; }
}
__device__ int collapseSlots(const float4* nodes, uint32_t n, Slot4 s[4]) {
    readSlots(nodes, n, s[0], s[1]);
    int cnt = 2;
#pragma unroll
    for (int step = 0; step < 2; step++) {
        int best = -1; float bestA = -1.0f;
        for (int k = 0; k < cnt; k++) {
            if (s[k].ref & RT_LEAF) continue;
            float A = boxArea(s[k].lo, s[k].hi);
            if (best < 0 || A > bestA) { best = k; bestA = A; }
        }
        if (best < 0) break;
        Slot4 l, r;
        readSlots(nodes, s[best].ref, l, r);
        for (int k = cnt; k > best + 1; k--) s[k] = s[k - 1];
        s[best] = l; s[best + 1] = r;
        cnt++;
    }
    return cnt;
}

// frontier expansion: marks the internal slots of every frontier node
__global__ void k_collapse_mark(const float4* __restrict__ nodes, const uint32_t* __restrict__ frontier, uint32_t nf,
                                uint32_t* __restrict__ keep, uint32_t* __restrict__ next, uint32_t* __restrict__ nNext) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    Slot4 s[4];
    int cnt = collapseSlots(nodes, frontier[i], s);
    for (int k = 0; k < cnt; k++) {
        if (s[k].ref & RT_LEAF) continue;
        keep[s[k].ref] = 1u;
        next[atomicAdd(nNext, 1u)] = s[k].ref;
    }
}

__global__ void k_collapse_emit(const float4* __restrict__ nodes, uint32_t nb, const uint32_t* __restrict__ keep,
                                const uint32_t* __restrict__ newId, float4* __restrict__ nodes4) {
    uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nb || !keep[n]) return;
    Slot4 s[4];
    int cnt = collapseSlots(nodes, n, s);
    float lo[3][4], hi[3][4];
    uint32_t ch[4];
    for (int k = 0; k < 4; k++) {
        if (k < cnt) {
            ch[k] = (s[k].ref & RT_LEAF) ? s[k].ref : newId[s[k].ref];
            for (int a = 0; a < 3; a++) { lo[a][k] = s[k].lo[a]; hi[a][k] = s[k].hi[a]; }
        } else {
            ch[k] = 0xffffffffu;
            for (int a = 0; a < 3; a++) { lo[a][k] = INFINITY; hi[a][k] = -INFINITY; }
        }
    }
    float4* o = nodes4 + 8 * (size_t)newId[n];
    for (int a = 0; a < 3; a++) {
        o[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        o[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    o[6] = make_float4(__uint_as_float(ch[0]), __uint_as_float(ch[1]), __uint_as_float(ch[2]), __uint_as_float(ch[3]));
    o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

namespace rti {

// Collapses the binary LBVH (s->nodes) into the 4-wide tree s->nodes4.
static bool collapseBVH4(rt_device_scene* s, uint32_t nb, hipStream_t stream) {
    uint32_t *keep = nullptr, *newId = nullptr, *fa = nullptr, *fb = nullptr, *cnt = nullptr;
    bool ok = hipCheck(hipMalloc(&keep, sizeof(uint32_t) * nb), "hipMalloc") &&
              hipCheck(hipMalloc(&newId, sizeof(uint32_t) * nb), "hipMalloc") &&
              hipCheck(hipMalloc(&fa, sizeof(uint32_t) * nb), "hipMalloc") &&
              hipCheck(hipMalloc(&fb, sizeof(uint32_t) * nb), "hipMalloc") &&
              hipCheck(hipMalloc(&cnt, sizeof(uint32_t)), "hipMalloc");
    uint32_t depth4 = 0, total = 0;
    if (ok) {
        uint32_t one = 1u, zero = 0u;
        ok = hipCheck(hipMemsetAsync(keep, 0, sizeof(uint32_t) * nb, stream), "memset") &&
             hipCheck(hipMemcpyAsync(keep, &one, 4, hipMemcpyHostToDevice, stream), "copy") &&
             hipCheck(hipMemcpyAsync(fa, &zero, 4, hipMemcpyHostToDevice, stream), "copy");
        uint32_t nf = 1;
        while (ok && nf > 0) {
            depth4++;
            ok = hipCheck(hipMemsetAsync(cnt, 0, 4, stream), "memset");
            if (!ok) break;
            k_collapse_mark<<<cdiv(nf, kBlock), kBlock, 0, stream>>>(s->nodes, fa, nf, keep, fb, cnt);
            ok = hipCheck(hipGetLastError(), "collapse mark") &&
                 hipCheck(hipMemcpyAsync(&nf, cnt, 4, hipMemcpyDeviceToHost, stream), "copy") &&
                 hipCheck(hipStreamSynchronize(stream), "sync");
            std::swap(fa, fb);
        }
    }
    if (ok) {
        ok = hipCheck(hipMemcpyAsync(newId, keep, sizeof(uint32_t) * nb, hipMemcpyDeviceToDevice, stream), "copy");
        k_scan_single<<<1, 1024, 0, stream>>>(newId, nb);
        uint32_t lastKeep = 0, lastId = 0;
        ok = ok && hipCheck(hipMemcpyAsync(&lastKeep, keep + nb - 1, 4, hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipMemcpyAsync(&lastId, newId + nb - 1, 4, hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipStreamSynchronize(stream), "sync");
        total = lastId + lastKeep;
    }
    if (ok) {
        ok = hipCheck(hipMalloc(&s->nodes4, sizeof(float4) * 8 * (size_t)total), "hipMalloc");
        if (ok) s->deviceBytes += sizeof(float4) * 8 * (size_t)total;
    }
    if (ok) {
        k_collapse_emit<<<cdiv(nb, kBlock), kBlock, 0, stream>>>(s->nodes, nb, keep, newId, s->nodes4);
        ok = hipCheck(hipGetLastError(), "collapse emit") && hipCheck(hipStreamSynchronize(stream), "sync");
    }
    s->numNodes4 = total;
    s->bvh4Depth = depth4;
    (void)hipFree(keep); (void)hipFree(newId); (void)hipFree(fa); (void)hipFree(fb); (void)hipFree(cnt);
    return ok;
}

}  // namespace rti

namespace rti {

bool buildLBVH(rt_device_scene* s, const std::vector<uint32_t>& instPrimOffset, hipStream_t stream) {
    const uint32_t n = s->numPrims;
    if (n == 0) return true;
    auto alloc = [&](void** p, size_t bytes) {
        if (!hipCheck(hipMalloc(p, bytes), "hipMalloc")) return false;
        s->deviceBytes += bytes;
        return true;
    };
    if (!alloc((void**)&s->primTris, sizeof(float4) * 3 * (size_t)n)) return false;
    if (!alloc((void**)&s->leafTris, sizeof(float4) * 3 * (size_t)n)) return false;
    if (!alloc((void**)&s->codes, sizeof(unsigned long long) * n)) return false;
    if (!alloc((void**)&s->order, sizeof(uint32_t) * n)) return false;
    const uint32_t nNodes = n > 1 ? n - 1 : 1;
    if (!alloc((void**)&s->nodes, sizeof(float4) * 4 * (size_t)nNodes)) return false;

    // scratch
    uint32_t* dOffsets = nullptr; unsigned long long* keysB = nullptr; uint32_t* valsB = nullptr;
    uint32_t* blockHist = nullptr; float* partial = nullptr; float* box = nullptr;
    uint32_t *parentI = nullptr, *parentL = nullptr, *counters = nullptr, *dDepth = nullptr;
    const uint32_t nTiles = cdiv(n, kTile);
    const uint32_t nbBounds = std::min<uint32_t>(1024u, cdiv(n, kBlock));
    bool ok = true;
    ok = ok && hipCheck(hipMalloc(&dOffsets, sizeof(uint32_t) * std::max<size_t>(1, instPrimOffset.size())), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&keysB, sizeof(unsigned long long) * n), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&valsB, sizeof(uint32_t) * n), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&blockHist, sizeof(uint32_t) * 256 * (size_t)nTiles), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&partial, sizeof(float) * 6 * nbBounds), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&box, sizeof(float) * 6), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&parentI, sizeof(uint32_t) * nNodes), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&parentL, sizeof(uint32_t) * n), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&counters, sizeof(uint32_t) * nNodes), "hipMalloc");
    ok = ok && hipCheck(hipMalloc(&dDepth, sizeof(uint32_t)), "hipMalloc");
    if (ok && !instPrimOffset.empty())
        ok = hipCheck(hipMemcpyAsync(dOffsets, instPrimOffset.data(), sizeof(uint32_t) * instPrimOffset.size(),
                                     hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    if (ok) {
        k_world_tris<<<cdiv(n, kBlock), kBlock, 0, stream>>>(s->dev, dOffsets, s->numInstances, s->primTris);
        k_bounds_partial<<<nbBounds, kBlock, 0, stream>>>(s->primTris, n, partial);
        k_bounds_final<<<1, kBlock, 0, stream>>>(partial, nbBounds, box);
        k_morton<<<cdiv(n, kBlock), kBlock, 0, stream>>>(s->primTris, n, box, s->codes, s->order);
        unsigned long long *kin = s->codes, *kout = keysB;
        uint32_t *vin = s->order, *vout = valsB;
        for (int pass = 0; pass < 8; pass++) {
            int shift = pass * 8;
            k_radix_hist<<<nTiles, kBlock, 0, stream>>>(kin, n, shift, nTiles, blockHist);
            k_scan_single<<<1, 1024, 0, stream>>>(blockHist, 256 * nTiles);
            k_radix_scatter<<<nTiles, kBlock, 0, stream>>>(kin, vin, n, shift, nTiles, blockHist, kout, vout);
            std::swap(kin, kout); std::swap(vin, vout);
        }
        // 8 passes: sorted data is back in s->codes / s->order
        ok = hipCheck(hipGetLastError(), "radix sort launch");
    }
    if (ok) {
        if (n == 1) {
            float4 h[8];
            ok = hipCheck(hipMemcpyAsync(h, s->primTris, sizeof(float4) * 3, hipMemcpyDeviceToHost, stream), "copy") &&
                 hipCheck(hipStreamSynchronize(stream), "sync");
            if (ok) {
                float lo[3] = {std::min(std::min(h[0].x, h[1].x), h[2].x), std::min(std::min(h[0].y, h[1].y), h[2].y),
                               std::min(std::min(h[0].z, h[1].z), h[2].z)};
                float hi[3] = {std::max(std::max(h[0].x, h[1].x), h[2].x), std::max(std::max(h[0].y, h[1].y), h[2].y),
                               std::max(std::max(h[0].z, h[1].z), h[2].z)};
                float nf[16] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], INFINITY, INFINITY, INFINITY,
                                -INFINITY, -INFINITY, -INFINITY, 0, 0, 0, 0};
                uint32_t leaf = RT_LEAF;
                std::memcpy(&nf[12], &leaf, 4); std::memcpy(&nf[13], &leaf, 4);
                ok = hipCheck(hipMemcpyAsync(s->nodes, nf, sizeof(nf), hipMemcpyHostToDevice, stream), "copy") &&
                     hipCheck(hipMemcpyAsync(s->leafTris, s->primTris, sizeof(float4) * 3, hipMemcpyDeviceToDevice, stream), "copy") &&
                     hipCheck(hipStreamSynchronize(stream), "sync");
                s->bvhDepth = 1;
            }
        } else {
            ok = hipCheck(hipMemsetAsync(counters, 0, sizeof(uint32_t) * nNodes, stream), "memset") &&
                 hipCheck(hipMemsetAsync(dDepth, 0, sizeof(uint32_t), stream), "memset");
            if (ok) {
                k_karras<<<cdiv(n - 1, kBlock), kBlock, 0, stream>>>(s->codes, (int)n, s->nodes, parentI, parentL);
                k_gather_leaves<<<cdiv(n, kBlock), kBlock, 0, stream>>>(s->primTris, s->order, n, s->leafTris);
                k_refit<<<cdiv(n, kBlock), kBlock, 0, stream>>>(s->leafTris, n, s->nodes, parentI, parentL, counters);
                k_depth<<<cdiv(n, kBlock), kBlock, 0, stream>>>(parentI, parentL, n, dDepth);
                ok = hipCheck(hipGetLastError(), "lbvh launch");
            }
            uint32_t depth = 0;
            if (ok) ok = hipCheck(hipMemcpyAsync(&depth, dDepth, 4, hipMemcpyDeviceToHost, stream), "copy") &&
                         hipCheck(hipStreamSynchronize(stream), "sync");
            s->bvhDepth = depth;
        }
    }
    if (ok) {
        float hb[6] = {0, 0, 0, 0, 0, 0};
        ok = hipCheck(hipMemcpyAsync(hb, box, sizeof(hb), hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipStreamSynchronize(stream), "sync");
        for (int a = 0; a < 3; a++) { s->sceneLo[a] = hb[a]; s->sceneHi[a] = hb[3 + a]; }
    }
    (void)hipFree(dOffsets); (void)hipFree(keysB); (void)hipFree(valsB); (void)hipFree(blockHist); (void)hipFree(partial); (void)hipFree(box);
    (void)hipFree(parentI); (void)hipFree(parentL); (void)hipFree(counters); (void)hipFree(dDepth);
    if (ok) ok = collapseBVH4(s, nNodes, stream);
    // a 4-wide node pushes at most three children: the stack never holds more
    // than 3 (depth - 1) entries
    if (ok && 3 * (uint64_t)s->bvh4Depth > RT_LDS_STACK + RT_SPILL_STACK) {
        setError("BVH depth " + std::to_string(s->bvh4Depth) + " exceeds the traversal stack");
        return false;
    }
    return ok;
}

}  // namespace rti
