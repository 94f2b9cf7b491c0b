// Software ray traversal (replaces the driver traversal behind
// rayQueryProceedEXT, src/shader/ray_query.glsl:6-70; CDNA4 has no ray
// tracing units).
//
// The traversed tree is the 4-wide collapse of the binary LBVH: one 128-B node
// (8 x float4: lo.x/lo.y/lo.z/hi.x/hi.y/hi.z of the four children as float4s,
// the four child references, padding) per visit. The hit children are ordered
// by entry distance (slot order on ties) with a five-comparator network;
// leaves are tested in that order, the nearest internal child is visited next
// and the other internal children are pushed farthest first.
//
// Per lane: a short stack in LDS (strided by the block size so a wave's pushes
// and pops hit distinct banks) that spills to a private array for deep trees.
// Closest-hit traversal keeps each pushed node's entry distance beside it (the
// top 16 bits of the float, i.e. rounded toward zero) and drops a popped node
// that lies beyond the closest hit found since it was pushed. That changes no
// result: a node's box contains its children's boxes, and the slab distances are
// monotone in the bounds under IEEE rounding, so every child of such a node
// would fail its own box test against the same closest hit.
// Leaves hold one triangle, tested with the watertight algorithm of Woop,
// Benthin and Wald (JCGT 2013) including its double-precision fallback. Box
// tests use (bound - o) * (1 / d) with IEEE minNum/maxNum so that 0 * inf slabs
// are dropped. Closest hit accepts tmin <= t < tBest.
#pragma once
#include "rt_dscene.h"

#ifndef RT_LDS_STACK
#define RT_LDS_STACK 12
#endif
#define RT_SPILL_STACK 240
#define RT_LEAF 0x80000000u
#define RT_EMPTY 0xffffffffu

namespace rtd {

// Work counters of the counting kernel builds (nodes, triangles, rays).
struct TraceCounters { unsigned long long nodes, tris, rays; };

// Pixel queue of a persistent pass. The work items (pixels in tile order) are
// split into kQueueParts contiguous bands of tiles, each with its own counter
// of claimed items in its own 128-B line, so that the waves' claims spread over
// kQueueParts atomics instead of one; `done` counts finished workgroups. All are
// zero between launches (the last workgroup of a launch clears them).
constexpr uint32_t kQueueParts = 8;
struct WorkQueue {
    struct Part { uint32_t next, pad[31]; } part[kQueueParts];
    uint32_t done, pad[31];
};

struct RayT {
    f3 o, d, invd;
    float tmin, tmax;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

RT_DEV float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

RT_DEV RayT makeRay(f3 o, f3 d, float tmin, float tmax) {
    RayT r;
    r.o = o; r.d = d; r.tmin = tmin; r.tmax = tmax;
    r.invd = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dz = comp(d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = comp(d, kx) / dz; r.Sy = comp(d, ky) / dz; r.Sz = 1.0f / dz;
    return r;
}

RT_DEV bool boxHit(const RayT& r, float lx, float ly, float lz, float hx, float hy, float hz, float tBest, float& tEnter) {
    float tx0 = (lx - r.o.x) * r.invd.x, tx1 = (hx - r.o.x) * r.invd.x;
    float ty0 = (ly - r.o.y) * r.invd.y, ty1 = (hy - r.o.y) * r.invd.y;
    float tz0 = (lz - r.o.z) * r.invd.z, tz1 = (hz - r.o.z) * r.invd.z;
    float tn = fmaxf(fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1)), r.tmin);
    float tf = fminf(fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)), tBest);
    tEnter = tn;
    return tn <= tf;
}

RT_DEV bool triHit(const RayT& r, f3 v0, f3 v1, f3 v2, float tBest, float& tOut, float& b1, float& b2) {
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = comp(A, r.kx) - r.Sx * Akz, Ay = comp(A, r.ky) - r.Sy * Akz;
    float Bx = comp(B, r.kx) - r.Sx * Bkz, By = comp(B, r.ky) - r.Sy * Bkz;
    float Cx = comp(C, r.kx) - r.Sx * Ckz, Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f) { T = -T; det = -det; V = -V; W = -W; }
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t < tBest)) return false;
    tOut = t; b1 = V * rcp; b2 = W * rcp;
    return true;
}

// LDS stack slice for this lane: node references and their entry distances
// (16-bit, see above), stride = threads per block.
struct TStack {
    uint32_t* lds;
    uint16_t* ldsT;
    uint32_t stride;
};

// The LDS of a block's traversal stacks.
template <int kThreads>
struct StackLDS {
    uint32_t ref[RT_LDS_STACK * kThreads];
    uint16_t t[RT_LDS_STACK * kThreads];
};

template <int kThreads>
RT_DEV TStack laneStack(StackLDS<kThreads>& s) {
    return TStack{s.ref + threadIdx.x, s.t + threadIdx.x, (uint32_t)kThreads};
}

// Entry distance as stored: the float's top 16 bits, which for t >= 0 is a
// value <= t; a negative t is stored as -inf so that it never drops a node.
RT_DEV uint16_t stackKey(float t) {
    const uint32_t b = __float_as_uint(t);
    return (b >> 31) ? (uint16_t)0xff80u : (uint16_t)(b >> 16);
}
RT_DEV float stackKeyValue(uint16_t k) { return __uint_as_float((uint32_t)k << 16); }

// (t, slot) key order; a missed child carries t = +inf and sorts last.
RT_DEV void cswap(float& ta, uint32_t& ra, float& tb, uint32_t& rb) {
    bool sw = tb < ta;
    float t = sw ? tb : ta; uint32_t r = sw ? rb : ra;
    tb = sw ? ta : tb; rb = sw ? ra : rb;
    ta = t; ra = r;
}

// kCull: closest-hit queries drop popped nodes beyond the closest hit (the
// result is the same either way; the reuse passes' replays, which rarely pop
// such nodes, leave it off to save registers).
template <bool kAny, bool kCount, bool kCull = true>
RT_DEV bool traverse(const DevScene& S, const RayT& r, bool skipLights, TStack st, float& tBest, Isec& hit,
                     TraceCounters* ctr) {
    uint32_t spill[RT_SPILL_STACK];
    constexpr bool kKeep = kCull && !kAny;  // entry distances kept on the stack
    uint16_t spillT[kKeep ? RT_SPILL_STACK : 1];
    int sp = 0;
    bool found = false;
    unsigned long long cn = 0, ct = 0;
    if (S.numPrims == 0) return false;
    // A non-finite ray hits nothing; reject it here, because the slab test
    // (fminf/fmaxf drop NaN operands) would accept every box and the ray would
    // walk the whole tree.
    if (!(isfinite(r.o.x) && isfinite(r.o.y) && isfinite(r.o.z) && isfinite(r.d.x) && isfinite(r.d.y) &&
          isfinite(r.d.z))) {
        if (kCount) atomicAdd(&ctr->rays, 1ull);
        return false;
    }
    uint32_t node = 0;
    for (;;) {
        const float4* np = S.nodes + 8 * (size_t)node;
        float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5], cr = np[6];
        if (kCount) cn++;
        uint32_t c0 = __float_as_uint(cr.x), c1 = __float_as_uint(cr.y), c2 = __float_as_uint(cr.z),
                 c3 = __float_as_uint(cr.w);
        // all four slab tests unconditionally (branch-free, so the node's seven loads issue
        // together); an empty slot is masked afterwards
        float t0, t1, t2, t3;
        bool h0 = boxHit(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, tBest, t0) & (c0 != RT_EMPTY);
        bool h1 = boxHit(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, tBest, t1) & (c1 != RT_EMPTY);
        bool h2 = boxHit(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, tBest, t2) & (c2 != RT_EMPTY);
        bool h3 = boxHit(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, tBest, t3) & (c3 != RT_EMPTY);
        // keys: (t, slot); strict < keeps slot order on equal t
        float k0 = h0 ? t0 : INFINITY, k1 = h1 ? t1 : INFINITY, k2 = h2 ? t2 : INFINITY, k3 = h3 ? t3 : INFINITY;
        uint32_t r0 = h0 ? c0 : RT_EMPTY, r1 = h1 ? c1 : RT_EMPTY, r2 = h2 ? c2 : RT_EMPTY, r3 = h3 ? c3 : RT_EMPTY;
        // stable 4-sort network of adjacent comparators (insertion order)
        cswap(k0, r0, k1, r1);
        cswap(k1, r1, k2, r2); cswap(k0, r0, k1, r1);
        cswap(k2, r2, k3, r3); cswap(k1, r1, k2, r2); cswap(k0, r0, k1, r1);
        uint32_t rs[4] = {r0, r1, r2, r3};
        const float ks[4] = {k0, k1, k2, k3};
        // leaves in sorted order, one triangle test per loop trip: a wave runs
        // as many trips as its lane with the most leaf hits (a leaf is tested
        // by ~1 lane in 5 per node visit), not one test per slot position
        uint32_t leafMask = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (rs[q] != RT_EMPTY && (rs[q] & RT_LEAF)) leafMask |= 1u << q;
        while (leafMask) {
            const uint32_t q = (uint32_t)__builtin_ctz(leafMask);
            leafMask &= leafMask - 1u;
            const uint32_t ref = q == 0 ? r0 : (q == 1 ? r1 : (q == 2 ? r2 : r3));
            uint32_t leaf = ref & ~RT_LEAF;
            const float4* tp = S.leafTris + 3 * (size_t)leaf;
            float4 a = tp[0], b = tp[1];
            uint32_t inst = __float_as_uint(a.w);
            if (skipLights && inst == 0) continue;
            float4 cc = tp[2];
            if (kCount) ct++;
            float t, b1, b2;
            if (triHit(r, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), tBest, t, b1, b2)) {
                found = true;
                if (kAny) {
                    if (kCount) { atomicAdd(&ctr->nodes, cn); atomicAdd(&ctr->tris, ct); atomicAdd(&ctr->rays, 1ull); }
                    return true;
                }
                tBest = t;
                hit.b1 = b1; hit.b2 = b2; hit.inst = inst; hit.prim = __float_as_uint(b.w);
            }
        }
        // nearest internal child next, the others pushed farthest first
        uint32_t im = 0;  // internal children, by sorted position
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (rs[q] != RT_EMPTY && !(rs[q] & RT_LEAF)) im |= 1u << q;
        uint32_t next = RT_EMPTY;
        if (im) {
            const uint32_t f = (uint32_t)__builtin_ctz(im);
            next = f == 0u ? r0 : (f == 1u ? r1 : (f == 2u ? r2 : r3));
        }
        const uint32_t pm = im & (im - 1u);  // pushed: every internal child but the nearest
        const int npush = __builtin_popcount(pm);
        if (sp + npush <= RT_LDS_STACK) {
            // all pushes fit in LDS: slot q lands at sp + the number of pushed slots farther than q
#pragma unroll
            for (int q = 3; q >= 0; q--) {
                if ((pm >> q) & 1u) {
                    const int pos = sp + __builtin_popcount(pm >> (q + 1));
                    st.lds[pos * st.stride] = rs[q];
                    if (kKeep) st.ldsT[pos * st.stride] = stackKey(ks[q]);
                }
            }
            sp += npush;
        } else {
#pragma unroll
            for (int q = 3; q >= 0; q--) {
                if (!((pm >> q) & 1u)) continue;
                if (sp < RT_LDS_STACK + RT_SPILL_STACK) {  // the build bounds the stack depth
                    if (sp < RT_LDS_STACK) {
                        st.lds[sp * st.stride] = rs[q];
                        if (kKeep) st.ldsT[sp * st.stride] = stackKey(ks[q]);
                    } else {
                        spill[sp - RT_LDS_STACK] = rs[q];
                        if (kKeep) spillT[sp - RT_LDS_STACK] = stackKey(ks[q]);
                    }
                    sp++;
                }
            }
        }
        // pop; a shadow ray's tBest is fixed, so only closest-hit rays drop nodes
        while (next == RT_EMPTY && sp > 0) {
            sp--;
            if (sp < RT_LDS_STACK) {
                if (!kKeep || !(stackKeyValue(st.ldsT[sp * st.stride]) > tBest)) next = st.lds[sp * st.stride];
            } else {
                if (!kKeep || !(stackKeyValue(spillT[sp - RT_LDS_STACK]) > tBest)) next = spill[sp - RT_LDS_STACK];
            }
        }
        if (next == RT_EMPTY) break;
        node = next;
    }
    if (kCount) { atomicAdd(&ctr->nodes, cn); atomicAdd(&ctr->tris, ct); atomicAdd(&ctr->rays, 1ull); }
    return found;
}

}  // namespace rtd
