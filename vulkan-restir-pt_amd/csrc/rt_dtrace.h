// Software ray traversal over the LBVH (replaces the driver traversal behind
// rayQueryProceedEXT, src/shader/ray_query.glsl:6-70; CDNA4 has no ray
// tracing units).
//
// Per lane: a short stack in LDS (kLdsStack entries, strided by the block size
// so a wave's pushes and pops hit distinct banks) that spills to a private
// array for deep trees. Both child boxes of a node arrive in one 64-B load;
// the nearer child is visited first (left on ties), the other is pushed.
// Leaves hold one triangle, tested with the watertight algorithm of Woop,
// Benthin and Wald (JCGT 2013) including its double-precision fallback.
// Box tests use (bound - o) * (1 / d) with IEEE minNum/maxNum so that 0 * inf
// slabs are dropped. Closest hit accepts tmin <= t < tBest.
#pragma once
#include "rt_dscene.h"

#ifndef RT_LDS_STACK
#define RT_LDS_STACK 16
#endif
#define RT_SPILL_STACK 96
#define RT_LEAF 0x80000000u

namespace rtd {

// Work counters of the counting kernel builds (nodes, triangles, rays).
struct TraceCounters { unsigned long long nodes, tris, rays; };

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

// LDS stack slice for this lane: stride = threads per block.
struct TStack {
    uint32_t* lds;
    uint32_t stride;
};

template <bool kAny, bool kCount>
RT_DEV bool traverse(const DevScene& S, const RayT& r, bool skipLights, TStack st, float& tBest, Isec& hit,
                     TraceCounters* ctr) {
    uint32_t spill[RT_SPILL_STACK];
    int sp = 0;
    bool found = false;
    unsigned long long cn = 0, ct = 0;
    if (S.numPrims == 0) return false;
    uint32_t node = 0;
    for (;;) {
        const float4* np = S.nodes + 4 * (size_t)node;
        float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
        if (kCount) cn++;
        float tl, tr;
        bool hl = boxHit(r, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, tBest, tl);
        bool hr = boxHit(r, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, tBest, tr);
        uint32_t cl = __float_as_uint(n3.x), cr = __float_as_uint(n3.y);
        uint32_t c[2]; int nc = 0;
        if (hl && hr) {
            bool lf = tl <= tr;
            c[0] = lf ? cl : cr; c[1] = lf ? cr : cl; nc = 2;
        } else if (hl) { c[0] = cl; nc = 1; }
        else if (hr) { c[0] = cr; nc = 1; }
        uint32_t next = 0xffffffffu;
        for (int k = 0; k < nc; k++) {
            uint32_t ch = c[k];
            if (ch & RT_LEAF) {
                uint32_t leaf = ch & ~RT_LEAF;
                const float4* tp = S.leafTris + 3 * (size_t)leaf;
                float4 a = tp[0], b = tp[1];
                uint32_t inst = __float_as_uint(a.w);
                if (!(skipLights && inst == 0)) {
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
            } else if (next == 0xffffffffu) {
                next = ch;
            } else if (sp < RT_LDS_STACK + RT_SPILL_STACK) {  // the build rejects deeper trees
                if (sp < RT_LDS_STACK) st.lds[sp * st.stride] = ch;
                else spill[sp - RT_LDS_STACK] = ch;
                sp++;
            }
        }
        if (next == 0xffffffffu) {
            if (sp == 0) break;
            sp--;
            next = (sp < RT_LDS_STACK) ? st.lds[sp * st.stride] : spill[sp - RT_LDS_STACK];
        }
        node = next;
    }
    if (kCount) { atomicAdd(&ctr->nodes, cn); atomicAdd(&ctr->tris, ct); atomicAdd(&ctr->rays, 1ull); }
    return found;
}

}  // namespace rtd
