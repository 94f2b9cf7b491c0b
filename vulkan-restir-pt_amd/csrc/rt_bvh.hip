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
//   6. bottom-up refit: per block of leaves, one thread per leaf climbs and
//      the second arrival at a node merges both child boxes; then the nodes
//      spanning several blocks, level by level from the deepest
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
constexpr uint32_t kNoParent = 0xffffffffu;

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

// Block-wide exclusive scan of one value per thread (blockDim.x a multiple of
// 64, at most 1024); *total receives the block's sum. Holds two barriers.
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t waveSum[16];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += t;
    }
    if (lane == 63u) waveSum[wave] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t w = 0; w < waves; w++) {
        uint32_t x = waveSum[w];
        if (w < wave) base += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
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

// One workgroup per digit: exclusive scan of the digit's per-block counts in
// place, digit total to digitTotal[digit].
__global__ void k_radix_digit_scan(uint32_t* __restrict__ blockHist, uint32_t nBlocks, uint32_t* __restrict__ digitTotal) {
    uint32_t* h = blockHist + (size_t)blockIdx.x * nBlocks;
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nBlocks; c += kBlock) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < nBlocks ? h[i] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan(v, &tot);
        if (i < nBlocks) h[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) digitTotal[blockIdx.x] = carry;
}

// Stable scatter: a key's destination is the exclusive count of smaller digits
// (scan of the digit totals), plus the count of its digit in earlier blocks,
// plus its rank among equal digits earlier in the block (8 wave ballots give
// the equal-digit lane mask; per-wave digit counts go through LDS).
__global__ void k_radix_scatter(const unsigned long long* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                uint32_t n, int shift, uint32_t nBlocks, const uint32_t* __restrict__ blockOffset,
                                const uint32_t* __restrict__ digitTotal, unsigned long long* __restrict__ keysOut,
                                uint32_t* __restrict__ valsOut) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t waveCnt[kBlock / 64][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t unused;
    const uint32_t digitBase = blockExclusiveScan(digitTotal[tid], &unused);
    running[tid] = digitBase + blockOffset[tid * nBlocks + blockIdx.x];
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

// Exclusive scan of n uint32 (in may equal out): per-chunk sums, one workgroup
// scanning the chunk sums, then each chunk rescanned with its base.
constexpr uint32_t kScanChunk = kBlock * kItems;

__global__ void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ sums) {
    const uint32_t b = blockIdx.x * kScanChunk + threadIdx.x * kItems;
    uint32_t s = 0;
    for (int k = 0; k < kItems; k++)
        if (b + k < n) s += in[b + k];
    uint32_t tot;
    (void)blockExclusiveScan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(uint32_t* __restrict__ sums, uint32_t m) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < m; c += blockDim.x) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < m ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan(v, &tot);
        if (i < m) sums[i] = carry + ex;
        carry += tot;
    }
}

// With compact non-null (and in holding 0/1 marks) also writes the marked
// indices, in order, to compact[out[i]].
__global__ void k_scan_apply(const uint32_t* in, uint32_t n, const uint32_t* __restrict__ sums, uint32_t* out,
                             uint32_t* __restrict__ compact) {
    const uint32_t b = blockIdx.x * kScanChunk + threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
    for (int k = 0; k < kItems; k++) {
        v[k] = b + k < n ? in[b + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + blockExclusiveScan(s, &tot);
    for (int k = 0; k < kItems; k++)
        if (b + k < n) {
            out[b + k] = run;
            if (compact && v[k]) compact[run] = b + k;
            run += v[k];
        }
}

// 5. Karras topology. Also records each node's leaf range.
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
                         uint32_t* __restrict__ parentI, uint32_t* __restrict__ parentL, uint2* __restrict__ range) {
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
    range[i] = make_uint2((uint32_t)min(i, j), (uint32_t)max(i, j));
    if (i == 0) parentI[0] = kNoParent;
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

// union of the node's two child boxes; returns the node's height (a leaf has
// height 0) from its children's
__device__ __forceinline__ uint32_t mergeNode(const float* nf, const uint32_t* height, float* lo, float* hi) {
    for (int a = 0; a < 3; a++) {
        lo[a] = fminf(nf[a], nf[6 + a]);
        hi[a] = fmaxf(nf[3 + a], nf[9 + a]);
    }
    const uint32_t l = __float_as_uint(nf[12]), r = __float_as_uint(nf[13]);
    const uint32_t hl = (l & RT_LEAF) ? 0u : height[l], hr = (r & RT_LEAF) ? 0u : height[r];
    return 1u + max(hl, hr);
}

// 6. refit, phase 1. Workgroup b owns leaves [b*S, b*S + S) and every internal
// node whose leaf range lies inside them (a Karras node is one end of its own
// range, so those nodes are numbered inside the same interval). One thread per
// leaf climbs: it writes its box into its slot of the parent; the second
// arrival at a node (counted in LDS) merges both slots and climbs on. Every
// arrival is ordered by workgroup-scope fences only, which on this part need no
// cache maintenance. A climb that reaches a node with leaves outside the
// workgroup leaves its slot written and stops there (phase 2 takes over).
constexpr uint32_t kRefitThreads = 1024;
constexpr uint32_t kRefitLeaves = 4096;

__global__ __launch_bounds__(kRefitThreads) void k_refit_block(const float4* __restrict__ tris,
                                                               const uint32_t* __restrict__ order, uint32_t n,
                                                               float4* nodes, const uint32_t* __restrict__ parentI,
                                                               const uint32_t* __restrict__ parentL,
                                                               const uint2* __restrict__ range, uint32_t* height) {
    __shared__ uint32_t arrivals[kRefitLeaves];
    const uint32_t L0 = blockIdx.x * kRefitLeaves, L1 = min(n, L0 + kRefitLeaves);
    for (uint32_t k = threadIdx.x; k < kRefitLeaves; k += kRefitThreads) arrivals[k] = 0u;
    __syncthreads();
    for (uint32_t i = L0 + threadIdx.x; i < L1; i += kRefitThreads) {
        const float4* t = tris + 3 * (size_t)(order ? order[i] : i);
        float lo[3] = {fminf(fminf(t[0].x, t[1].x), t[2].x), fminf(fminf(t[0].y, t[1].y), t[2].y),
                       fminf(fminf(t[0].z, t[1].z), t[2].z)};
        float hi[3] = {fmaxf(fmaxf(t[0].x, t[1].x), t[2].x), fmaxf(fmaxf(t[0].y, t[1].y), t[2].y),
                       fmaxf(fmaxf(t[0].z, t[1].z), t[2].z)};
        uint32_t ref = RT_LEAF | i;
        uint32_t p = parentL[i];
        while (true) {
            float* nf = reinterpret_cast<float*>(nodes + 4 * (size_t)p);
            writeSlot(nf, __float_as_uint(nf[12]) != ref, lo, hi);
            const uint2 r = range[p];
            if (r.x < L0 || r.y >= L1) break;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const uint32_t old = atomicAdd(&arrivals[p - L0], 1u);
            if (old == 0u) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            height[p] = mergeNode(nf, height, lo, hi);
            ref = p;
            p = parentI[p];
            if (p == kNoParent) break;
        }
    }
}

// Refit, phase 2: the nodes whose leaves span several phase-1 workgroups (the
// top part, a few per workgroup boundary), in one workgroup. They are listed
// level by level from the root (a child belongs to the top part when its leaf
// range spans blocks), then merged from the deepest level up: a node's slots
// then hold both children's boxes (phase 1 wrote the block subtrees' roots,
// the level below the top part's children), and its union goes into its
// parent's slot. Workgroup barriers order the levels. The binary tree is at
// most 98 levels high (one Karras split per common-prefix length of the 64-bit
// codes and 32-bit indices), so kMaxTopLevels bounds the level list.
constexpr uint32_t kMaxTopLevels = 128;

__global__ __launch_bounds__(kRefitThreads) void k_refit_top(float4* nodes, const uint32_t* __restrict__ parentI,
                                                             const uint2* __restrict__ range, uint32_t n,
                                                             uint32_t* height, uint32_t* list) {
    __shared__ uint32_t levelEnd[kMaxTopLevels];
    __shared__ uint32_t tail, levels;
    if (n <= kRefitLeaves) return;  // one block: phase 1 merged every node up to the root
    if (threadIdx.x == 0) {
        list[0] = 0u;  // the root
        tail = 1u;
        levels = 0u;
    }
    __syncthreads();
    uint32_t a = 0, b = 1;
    while (a < b) {
        if (threadIdx.x == 0) levelEnd[min(levels++, kMaxTopLevels - 1u)] = b;
        for (uint32_t i = a + threadIdx.x; i < b; i += blockDim.x) {
            const float* nf = reinterpret_cast<const float*>(nodes + 4 * (size_t)list[i]);
            for (int k = 0; k < 2; k++) {
                const uint32_t c = __float_as_uint(nf[12 + k]);
                if (c & RT_LEAF) continue;
                const uint2 r = range[c];
                if (r.x / kRefitLeaves != r.y / kRefitLeaves) list[atomicAdd(&tail, 1u)] = c;
            }
        }
        __syncthreads();
        a = b;
        b = tail;
        __syncthreads();
    }
    const int top = (int)min(levels, kMaxTopLevels);
    for (int L = top - 1; L >= 0; L--) {
        const uint32_t s0 = L ? levelEnd[L - 1] : 0u, s1 = levelEnd[L];
        for (uint32_t i = s0 + threadIdx.x; i < s1; i += blockDim.x) {
            const uint32_t p = list[i];
            float lo[3], hi[3];
            height[p] = mergeNode(reinterpret_cast<const float*>(nodes + 4 * (size_t)p), height, lo, hi);
            const uint32_t q = parentI[p];
            if (q == kNoParent) continue;
            float* qf = reinterpret_cast<float*>(nodes + 4 * (size_t)q);
            writeSlot(qf, __float_as_uint(qf[12]) != p, lo, hi);
        }
        __syncthreads();
    }
}

// a one-primitive scene: one node holding the leaf in its left slot
__global__ void k_single_leaf(const float4* __restrict__ tris, float4* __restrict__ nodes, uint32_t* __restrict__ height) {
    const float4* t = tris;
    float lo[3] = {fminf(fminf(t[0].x, t[1].x), t[2].x), fminf(fminf(t[0].y, t[1].y), t[2].y),
                   fminf(fminf(t[0].z, t[1].z), t[2].z)};
    float hi[3] = {fmaxf(fmaxf(t[0].x, t[1].x), t[2].x), fmaxf(fmaxf(t[0].y, t[1].y), t[2].y),
                   fmaxf(fmaxf(t[0].z, t[1].z), t[2].z)};
    nodes[0] = make_float4(lo[0], lo[1], lo[2], hi[0]);
    nodes[1] = make_float4(hi[1], hi[2], INFINITY, INFINITY);
    nodes[2] = make_float4(INFINITY, -INFINITY, -INFINITY, -INFINITY);
    nodes[3] = make_float4(__uint_as_float(RT_LEAF), __uint_as_float(RT_LEAF), 0.0f, 0.0f);
    height[0] = 1u;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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

// Keep marks, top-down in two launches. A kept node keeps the internal nodes
// among its collapsed slots; every other internal node is absorbed into the
// 4-wide node of its nearest kept ancestor. The nodes whose leaves span several
// kCollapseLeaves-leaf blocks (the top part, a few per block boundary) are
// expanded by one workgroup, level after level; a kept node it reaches inside a
// block is a seed of that block. Then one workgroup per block expands from its
// seeds: a block's nodes are numbered inside its leaf interval and its subtrees
// stay inside it, so the workgroup first copies each of its nodes' two slot
// references and slot box areas into LDS and then runs every level from LDS.
// Each node's 4-wide level (root 0) travels with it; depth4 receives the
// largest level + 1.
constexpr uint32_t kCollapseLeaves = kRefitLeaves;
constexpr uint32_t kCollapseThreads = 1024;

__device__ __forceinline__ uint32_t blockOf(uint32_t leaf) { return leaf / kCollapseLeaves; }

// list: nb entries (the top part's kept nodes, level after level); seeds: nb
// entries of (node, block << 8 | level); ctr[0] = seed count, ctr[1] = depth4
// (both zero). A 4-wide level is below the binary height (< 256).
__global__ __launch_bounds__(kCollapseThreads) void k_collapse_top(const float4* __restrict__ nodes,
                                                                   const uint2* __restrict__ range, uint32_t* keep,
                                                                   uint32_t* list, uint2* __restrict__ seeds,
                                                                   uint32_t* ctr) {
    __shared__ uint32_t tail, deepest;
    if (threadIdx.x == 0) {
        list[0] = 0u;  // the root; kept by the caller's fill
        tail = 1u;
        deepest = 0u;
    }
    __syncthreads();
    uint32_t a = 0, b = 1, level = 0;
    while (a < b) {
        for (uint32_t i = a + threadIdx.x; i < b; i += blockDim.x) {
            Slot4 sl[4];
            const int c = collapseSlots(nodes, list[i], sl);
            for (int k = 0; k < c; k++) {
                const uint32_t r = sl[k].ref;
                if (r & RT_LEAF) continue;
                keep[r] = 1u;
                const uint2 lr = range[r];
                if (blockOf(lr.x) != blockOf(lr.y)) list[atomicAdd(&tail, 1u)] = r;
                else seeds[atomicAdd(&ctr[0], 1u)] = make_uint2(r, blockOf(lr.x) << 8 | (level + 1u));
            }
        }
        if (threadIdx.x == 0) deepest = level;
        __syncthreads();
        a = b;
        b = tail;
        level++;
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&ctr[1], deepest + 1u);
}

// One workgroup per block of kCollapseLeaves leaves (n >= 2); nb = n - 1.
__global__ __launch_bounds__(kCollapseThreads) void k_collapse_block(const float4* __restrict__ nodes, uint32_t n,
                                                                     uint32_t nb, uint32_t* keep,
                                                                     const uint2* __restrict__ seeds,
                                                                     const uint32_t* ctr, uint32_t* depth4) {
    // the block's node ids [L0, min(L1, nb)): slot references and slot box areas
    __shared__ uint2 slotRef[kCollapseLeaves];
    __shared__ float2 slotArea[kCollapseLeaves];
    // a block holds at most kCollapseLeaves - 1 internal nodes, each in one frontier once
    __shared__ uint32_t front[2][kCollapseLeaves];
    __shared__ uint8_t lev[2][kCollapseLeaves];
    __shared__ uint32_t size[2], deepest;
    const uint32_t L0 = blockIdx.x * kCollapseLeaves, I1 = min(nb, min(n, L0 + kCollapseLeaves));
    if (threadIdx.x == 0) { size[0] = 0u; size[1] = 0u; deepest = 0u; }
    for (uint32_t i = L0 + threadIdx.x; i < I1; i += blockDim.x) {
        const float4* q = nodes + 4 * (size_t)i;
        const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        const float v[12] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
        slotRef[i - L0] = make_uint2(__float_as_uint(q3.x), __float_as_uint(q3.y));
        slotArea[i - L0] = make_float2(boxArea(v, v + 3), boxArea(v + 6, v + 9));
    }
    __syncthreads();
    const uint32_t ns = ctr[0];
    for (uint32_t e = threadIdx.x; e < ns; e += blockDim.x) {
        const uint2 sd = seeds[e];
        if ((sd.y >> 8) == blockIdx.x) {
            const uint32_t at = atomicAdd(&size[0], 1u);
            front[0][at] = sd.x;
            lev[0][at] = (uint8_t)(sd.y & 0xffu);
        }
    }
    __syncthreads();
    uint32_t cur = 0, myDeepest = 0;
    while (size[cur] != 0u) {
        const uint32_t nf = size[cur];
        for (uint32_t i = threadIdx.x; i < nf; i += blockDim.x) {
            const uint32_t level = lev[cur][i];
            myDeepest = max(myDeepest, level);
            // collapseSlots on the LDS copy: the same slots, areas and first-largest choice
            const uint32_t x = front[cur][i] - L0;
            uint32_t r[4] = {slotRef[x].x, slotRef[x].y, 0u, 0u};
            float A[4] = {slotArea[x].x, slotArea[x].y, 0.0f, 0.0f};
            int c = 2;
            for (int step = 0; step < 2; step++) {
                int best = -1; float bestA = -1.0f;
                for (int k = 0; k < c; k++) {
                    if (r[k] & RT_LEAF) continue;
                    if (best < 0 || A[k] > bestA) { best = k; bestA = A[k]; }
                }
                if (best < 0) break;
                const uint32_t y = r[best] - L0;
                for (int k = c; k > best + 1; k--) { r[k] = r[k - 1]; A[k] = A[k - 1]; }
                r[best] = slotRef[y].x; A[best] = slotArea[y].x;
                r[best + 1] = slotRef[y].y; A[best + 1] = slotArea[y].y;
                c++;
            }
            int m = 0;
            for (int k = 0; k < c; k++) m += (r[k] & RT_LEAF) ? 0 : 1;
            if (m) {
                uint32_t at = atomicAdd(&size[cur ^ 1u], (uint32_t)m);
                for (int k = 0; k < c; k++) {
                    if (r[k] & RT_LEAF) continue;
                    keep[r[k]] = 1u;
                    front[cur ^ 1u][at] = r[k];
                    lev[cur ^ 1u][at] = (uint8_t)(level + 1u);
                    at++;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) size[cur] = 0u;
        cur ^= 1u;
        __syncthreads();
    }
    if (myDeepest) atomicMax(&deepest, myDeepest);
    __syncthreads();
    if (threadIdx.x == 0 && deepest) atomicMax(depth4, deepest + 1u);
}

// One thread per kept node: kept[j] is the binary node of 4-wide node j; the
// count is the last node's exclusive-scan position plus its mark.
__global__ void k_collapse_emit(const float4* __restrict__ nodes, const uint32_t* __restrict__ keep,
                                const uint32_t* __restrict__ kept, uint32_t nb, const uint32_t* __restrict__ newId,
                                float4* __restrict__ nodes4) {
    const uint32_t total = newId[nb - 1] + keep[nb - 1];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) {
        const uint32_t n = kept[j];
        Slot4 s[4];
        int cnt4 = collapseSlots(nodes, n, s);
        float lo[3][4], hi[3][4];
        uint32_t ch[4];
        for (int k = 0; k < 4; k++) {
            if (k < cnt4) {
                ch[k] = (s[k].ref & RT_LEAF) ? s[k].ref : newId[s[k].ref];
                for (int a = 0; a < 3; a++) { lo[a][k] = s[k].lo[a]; hi[a][k] = s[k].hi[a]; }
            } else {
                ch[k] = 0xffffffffu;
                for (int a = 0; a < 3; a++) { lo[a][k] = INFINITY; hi[a][k] = -INFINITY; }
            }
        }
        float4* o = nodes4 + 8 * (size_t)j;
        for (int a = 0; a < 3; a++) {
            o[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            o[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        o[6] = make_float4(__uint_as_float(ch[0]), __uint_as_float(ch[1]), __uint_as_float(ch[2]), __uint_as_float(ch[3]));
        o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Exclusive scan of n uint32 (see k_scan_apply for compact); sums holds
// cdiv(n, kScanChunk) entries.
void scanExclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, uint32_t* compact,
                   hipStream_t stream) {
    const uint32_t chunks = cdiv(n, kScanChunk);
    k_scan_reduce<<<chunks, kBlock, 0, stream>>>(in, n, sums);
    k_scan_sums<<<1, 1024, 0, stream>>>(sums, chunks);
    k_scan_apply<<<chunks, kBlock, 0, stream>>>(in, n, sums, out, compact);
}

}  // namespace

namespace rti {

size_t carveLbvhScratch(uint32_t n, char* base, LbvhScratch* out) {
    const size_t nTiles = cdiv(n, kTile), nb = std::min<uint32_t>(1024u, cdiv(n, kBlock));
    const size_t nNodes = n > 1 ? n - 1 : 1;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    LbvhScratch s;
    s.keysB = reinterpret_cast<unsigned long long*>(take(8 * (size_t)n));
    s.valsB = reinterpret_cast<uint32_t*>(take(4 * (size_t)n));
    s.blockHist = reinterpret_cast<uint32_t*>(take(4 * 256 * nTiles));
    s.digitTotal = reinterpret_cast<uint32_t*>(take(4 * 256));
    s.partial = reinterpret_cast<float*>(take(4 * 6 * nb));
    s.box = reinterpret_cast<float*>(take(4 * 6));
    s.parentI = reinterpret_cast<uint32_t*>(take(4 * nNodes));
    s.parentL = reinterpret_cast<uint32_t*>(take(4 * (size_t)n));
    s.range = reinterpret_cast<uint2*>(take(8 * nNodes));
    s.height = reinterpret_cast<uint32_t*>(take(4 * nNodes));
    s.topList = reinterpret_cast<uint32_t*>(take(4 * nNodes));
    if (out) *out = s;
    return off;
}

// Binary LBVH of n world-space triangles (3 float4 per primitive): centroid
// bounds, Morton codes, radix sort, Karras topology and refit. Launches only:
// no allocation, no host synchronisation. Outputs codes, order (n entries),
// nodes (max(n - 1, 1) internal nodes); with leafTris non-null the triangles
// are also gathered into Morton order and the refit reads them there. The
// scene's box lands in sc.box, the tree's height in sc.height[0].
hipError_t launchBinaryBuild(const float4* primTris, uint32_t n, unsigned long long* codes, uint32_t* order,
                             float4* nodes, float4* leafTris, const LbvhScratch& sc, hipStream_t stream) {
    const uint32_t nTiles = cdiv(n, kTile);
    const uint32_t nbBounds = std::min<uint32_t>(1024u, cdiv(n, kBlock));
    k_bounds_partial<<<nbBounds, kBlock, 0, stream>>>(primTris, n, sc.partial);
    k_bounds_final<<<1, kBlock, 0, stream>>>(sc.partial, nbBounds, sc.box);
    k_morton<<<cdiv(n, kBlock), kBlock, 0, stream>>>(primTris, n, sc.box, codes, order);
    unsigned long long *kin = codes, *kout = sc.keysB;
    uint32_t *vin = order, *vout = sc.valsB;
    for (int pass = 0; pass < 8; pass++) {
        const int shift = pass * 8;
        k_radix_hist<<<nTiles, kBlock, 0, stream>>>(kin, n, shift, nTiles, sc.blockHist);
        k_radix_digit_scan<<<256, kBlock, 0, stream>>>(sc.blockHist, nTiles, sc.digitTotal);
        k_radix_scatter<<<nTiles, kBlock, 0, stream>>>(kin, vin, n, shift, nTiles, sc.blockHist, sc.digitTotal, kout,
                                                       vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // 8 passes: the sorted data is back in codes / order
    if (n == 1) {
        k_single_leaf<<<1, 1, 0, stream>>>(primTris, nodes, sc.height);
        if (leafTris) k_gather_leaves<<<1, kBlock, 0, stream>>>(primTris, order, n, leafTris);
        return hipGetLastError();
    }
    k_karras<<<cdiv(n - 1, kBlock), kBlock, 0, stream>>>(codes, (int)n, nodes, sc.parentI, sc.parentL, sc.range);
    const float4* tris = primTris;
    const uint32_t* refitOrder = order;
    if (leafTris) {
        k_gather_leaves<<<cdiv(n, kBlock), kBlock, 0, stream>>>(primTris, order, n, leafTris);
        tris = leafTris;
        refitOrder = nullptr;
    }
    k_refit_block<<<cdiv(n, kRefitLeaves), kRefitThreads, 0, stream>>>(tris, refitOrder, n, nodes, sc.parentI,
                                                                       sc.parentL, sc.range, sc.height);
    k_refit_top<<<1, kRefitThreads, 0, stream>>>(nodes, sc.parentI, sc.range, n, sc.height, sc.topList);
    return hipGetLastError();
}

// Scratch of the 4-wide collapse, allocated before the build starts.
struct CollapseScratch {
    uint32_t *keep = nullptr, *newId = nullptr, *kept = nullptr, *ctr = nullptr, *sums = nullptr;
    uint2* seeds = nullptr;
    bool alloc(uint32_t nb) {
        return hipCheck(hipMalloc(&keep, sizeof(uint32_t) * nb), "hipMalloc") &&
               hipCheck(hipMalloc(&newId, sizeof(uint32_t) * nb), "hipMalloc") &&
               hipCheck(hipMalloc(&kept, sizeof(uint32_t) * nb), "hipMalloc") &&
               hipCheck(hipMalloc(&seeds, sizeof(uint2) * nb), "hipMalloc") &&
               hipCheck(hipMalloc(&ctr, sizeof(uint32_t) * 2), "hipMalloc") &&
               hipCheck(hipMalloc(&sums, sizeof(uint32_t) * cdiv(nb, kScanChunk)), "hipMalloc");
    }
    ~CollapseScratch() {
        (void)hipFree(keep); (void)hipFree(newId); (void)hipFree(kept); (void)hipFree(seeds); (void)hipFree(ctr);
        (void)hipFree(sums);
    }
};

// Collapses the binary LBVH of n primitives (s->nodes, nb internal nodes) into
// the 4-wide tree s->nodes4 (allocated for nb nodes: the kept nodes are a
// subset of the binary internal nodes). Launches only: keep marks (top part,
// then per block), the exclusive scan of the marks that numbers the kept nodes
// in binary order and lists them, then one thread per 4-wide node.
static void launchCollapseBVH4(rt_device_scene* s, uint32_t n, uint32_t nb, const uint2* range, CollapseScratch& cs,
                               hipStream_t stream) {
    k_fill_u32<<<cdiv(nb, kBlock), kBlock, 0, stream>>>(cs.keep, nb, 0u);
    k_fill_u32<<<1, 1, 0, stream>>>(cs.keep, 1, 1u);
    k_fill_u32<<<1, 2, 0, stream>>>(cs.ctr, 2, 0u);
    k_collapse_top<<<1, kCollapseThreads, 0, stream>>>(s->nodes, range, cs.keep, cs.kept, cs.seeds, cs.ctr);
    if (n >= 2)
        k_collapse_block<<<cdiv(n, kCollapseLeaves), kCollapseThreads, 0, stream>>>(s->nodes, n, nb, cs.keep, cs.seeds,
                                                                                     cs.ctr, cs.ctr + 1);
    scanExclusive(cs.keep, cs.newId, nb, cs.sums, cs.kept, stream);
    k_collapse_emit<<<std::max(1u, std::min(2048u, cdiv(nb, kBlock))), kBlock, 0, stream>>>(s->nodes, cs.keep, cs.kept,
                                                                                           nb, cs.newId, s->nodes4);
}

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
    if (!alloc((void**)&s->nodes4, sizeof(float4) * 8 * (size_t)nNodes)) return false;

    uint32_t* dOffsets = nullptr;
    char* scratch = nullptr;
    LbvhScratch sc;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool ok = hipCheck(hipMalloc(&dOffsets, sizeof(uint32_t) * std::max<size_t>(1, instPrimOffset.size())), "hipMalloc") &&
              hipCheck(hipMalloc(&scratch, carveLbvhScratch(n, nullptr, nullptr)), "hipMalloc");
    CollapseScratch cs;
    ok = ok && cs.alloc(nNodes);
    for (int k = 0; ok && k < 3; k++) ok = hipCheck(hipEventCreate(&ev[k]), "hipEventCreate");
    if (ok) carveLbvhScratch(n, scratch, &sc);
    if (ok && !instPrimOffset.empty())
        ok = hipCheck(hipMemcpyAsync(dOffsets, instPrimOffset.data(), sizeof(uint32_t) * instPrimOffset.size(),
                                     hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    if (ok) {
        ok = hipCheck(hipEventRecord(ev[0], stream), "event");
        k_world_tris<<<cdiv(n, kBlock), kBlock, 0, stream>>>(s->dev, dOffsets, s->numInstances, s->primTris);
        ok = ok && hipCheck(launchBinaryBuild(s->primTris, n, s->codes, s->order, s->nodes, s->leafTris, sc, stream),
                            "lbvh launch") &&
             hipCheck(hipEventRecord(ev[1], stream), "event");
    }
    if (ok) {
        launchCollapseBVH4(s, n, nNodes, sc.range, cs, stream);
        ok = hipCheck(hipGetLastError(), "collapse launch") && hipCheck(hipEventRecord(ev[2], stream), "event");
    }
    if (ok) {
        float hb[6] = {0, 0, 0, 0, 0, 0};
        uint32_t depth = 0, depth4 = 0, lastKeep = 0, lastId = 0;
        ok = hipCheck(hipMemcpyAsync(hb, sc.box, sizeof(hb), hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipMemcpyAsync(&depth, sc.height, 4, hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipMemcpyAsync(&depth4, cs.ctr + 1, 4, hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipMemcpyAsync(&lastKeep, cs.keep + nNodes - 1, 4, hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipMemcpyAsync(&lastId, cs.newId + nNodes - 1, 4, hipMemcpyDeviceToHost, stream), "copy") &&
             hipCheck(hipStreamSynchronize(stream), "sync");
        for (int a = 0; a < 3; a++) { s->sceneLo[a] = hb[a]; s->sceneHi[a] = hb[3 + a]; }
        s->bvhDepth = depth;
        s->bvh4Depth = depth4;
        s->numNodes4 = lastId + lastKeep;
    }
    (void)hipFree(dOffsets);
    (void)hipFree(scratch);
    if (ok) {
        ok = hipCheck(hipEventElapsedTime(&s->bvhBinaryMs, ev[0], ev[1]), "event time") &&
             hipCheck(hipEventElapsedTime(&s->bvhBuildMs, ev[0], ev[2]), "event time");
    }
    for (int k = 0; k < 3; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    // a 4-wide node pushes at most three children: the stack never holds more
    // than 3 (depth - 1) entries
    if (ok && 3 * (uint64_t)s->bvh4Depth > RT_LDS_STACK + RT_SPILL_STACK) {
        setError("BVH depth " + std::to_string(s->bvh4Depth) + " exceeds the traversal stack");
        return false;
    }
    return ok;
}

}  // namespace rti
