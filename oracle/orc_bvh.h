// ORACLE — test infrastructure only. Nothing in the product links or calls this.
//
// CPU restatement of the acceleration structure that replaces the driver BVH
// built in src/Scene.cpp:448-547 and traversed by rayQueryProceedEXT in
// src/shader/ray_query.glsl:6-70.
//
// Primitive enumeration (mirrors the TLAS custom indices of
// src/Scene.cpp:491-541): first every light triangle (instance custom index
// 0, primitive index = light id), then for every object instance i (custom
// index i+1) its triangles 0..indexCount/3-1. Object triangles are transformed
// to world space with the instance transform (GLSL mat4*vec4 order).
//
// LBVH (Karras 2012, "Maximizing parallelism in the construction of BVHs"):
//   * centroid c = (v0 + v1 + v2) * (1/3)
//   * scene box = min/max over centroids; per-axis scale = 2097151 / extent
//     (0 when extent == 0); q = min(uint(trunc((c - lo) * scale)), 2097151)
//   * 63-bit Morton code, x in the top bit of each triple
//   * stable sort by code (ties keep primitive order)
//   * topology by Karras' delta with index tiebreak (64 + clz(i ^ j))
//   * node boxes = exact min/max unions (order independent)
// Traversal: stack based, both children box-tested, nearer child first (left
// on ties), closest hit accepts tmin <= t < tBest. Watertight ray/triangle test
// (Woop, Benthin, Wald 2013) with the paper's double-precision fallback for
// zero edge functions, two-sided.
#pragma once
#include "orc_scene.h"
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

struct WorldTri { vec3 v0, v1, v2; uint32_t inst, prim; };

struct BVHNode {
    float lmin[3], lmax[3], rmin[3], rmax[3];
    uint32_t left, right;  // high bit set => leaf index
};
static const uint32_t LeafFlag = 0x80000000u;

struct BVH4Node {
    float lo[3][4], hi[3][4];  // [axis][slot]
    uint32_t child[4];
};
static const uint32_t EmptyRef = 0xffffffffu;

struct BVH {
    std::vector<WorldTri> tris;           // primitive order
    std::vector<uint64_t> codes;          // sorted Morton codes
    std::vector<uint32_t> order;          // leaf i -> primitive id
    std::vector<BVHNode> nodes;           // N-1 internal nodes (root 0)
    std::vector<BVH4Node> nodes4;         // collapsed 4-wide tree (root 0)
    float sceneLo[3], sceneHi[3];
};

static inline void buildWorldTris(const SceneArrays& S, std::vector<WorldTri>& out) {
    out.clear();
    for (uint32_t i = 0; i < S.numLights; i++) {
        const TriangleLight& L = S.lights[i];
        out.push_back({L3(L.v0), L3(L.v1), L3(L.v2), 0u, i});
    }
    for (uint32_t k = 0; k < S.numInstances; k++) {
        const ObjectInstance& I = S.instances[k];
        uint32_t ntri = I.indexCount / 3;
        for (uint32_t t = 0; t < ntri; t++) {
            vec3 p[3];
            for (int j = 0; j < 3; j++) {
                const MeshVertex& v = S.vertices[S.indices[I.indexOffset + t * 3 + j]];
                p[j] = mat4MulPoint(I.transform, L3(v.pos), 1.0f);
            }
            out.push_back({p[0], p[1], p[2], k + 1, t});
        }
    }
}

static inline uint64_t expandBits21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
static inline vec3 centroidOf(const WorldTri& t) { return (t.v0 + t.v1 + t.v2) * (1.0f / 3.0f); }

static inline uint32_t quantAxis(float c, float lo, float scale) {
    float f = (c - lo) * scale;
    uint32_t q = (uint32_t)f;
    return q > 2097151u ? 2097151u : q;
}

static inline int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }
static inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

static inline int karrasDelta(const std::vector<uint64_t>& c, int i, int j) {
    int n = (int)c.size();
    if (j < 0 || j >= n) return -1;
    if (c[i] == c[j]) return 64 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz64(c[i] ^ c[j]);
}

static inline void triBounds(const WorldTri& t, float* lo, float* hi) {
    lo[0] = std::fmin(std::fmin(t.v0.x, t.v1.x), t.v2.x);
    lo[1] = std::fmin(std::fmin(t.v0.y, t.v1.y), t.v2.y);
    lo[2] = std::fmin(std::fmin(t.v0.z, t.v1.z), t.v2.z);
    hi[0] = std::fmax(std::fmax(t.v0.x, t.v1.x), t.v2.x);
    hi[1] = std::fmax(std::fmax(t.v0.y, t.v1.y), t.v2.y);
    hi[2] = std::fmax(std::fmax(t.v0.z, t.v1.z), t.v2.z);
}

static inline void buildBVH(BVH& B) {
    const int n = (int)B.tris.size();
    B.codes.clear(); B.order.clear(); B.nodes.clear();
    if (n == 0) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<vec3> cen(n);
    for (int i = 0; i < n; i++) {
        cen[i] = centroidOf(B.tris[i]);
        lo[0] = std::fmin(lo[0], cen[i].x); hi[0] = std::fmax(hi[0], cen[i].x);
        lo[1] = std::fmin(lo[1], cen[i].y); hi[1] = std::fmax(hi[1], cen[i].y);
        lo[2] = std::fmin(lo[2], cen[i].z); hi[2] = std::fmax(hi[2], cen[i].z);
    }
    for (int a = 0; a < 3; a++) { B.sceneLo[a] = lo[a]; B.sceneHi[a] = hi[a]; }
    float scale[3];
    for (int a = 0; a < 3; a++) {
        float ext = hi[a] - lo[a];
        scale[a] = ext > 0.0f ? 2097151.0f / ext : 0.0f;
    }
    std::vector<uint64_t> code(n);
    for (int i = 0; i < n; i++) {
        uint64_t x = quantAxis(cen[i].x, lo[0], scale[0]);
        uint64_t y = quantAxis(cen[i].y, lo[1], scale[1]);
        uint64_t z = quantAxis(cen[i].z, lo[2], scale[2]);
        code[i] = (expandBits21(x) << 2) | (expandBits21(y) << 1) | expandBits21(z);
    }
    B.order.resize(n);
    std::iota(B.order.begin(), B.order.end(), 0u);
    std::stable_sort(B.order.begin(), B.order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    B.codes.resize(n);
    for (int i = 0; i < n; i++) B.codes[i] = code[B.order[i]];

    if (n == 1) {
        BVHNode r{};
        float l0[3], l1[3];
        triBounds(B.tris[B.order[0]], l0, l1);
        for (int a = 0; a < 3; a++) { r.lmin[a] = l0[a]; r.lmax[a] = l1[a]; r.rmin[a] = INFINITY; r.rmax[a] = -INFINITY; }
        r.left = LeafFlag | 0u; r.right = LeafFlag | 0u;
        B.nodes.push_back(r);
        return;
    }
    B.nodes.assign(n - 1, BVHNode{});
    std::vector<int> parentI(n - 1, -1), parentL(n, -1);
    const auto& C = B.codes;
    for (int i = 0; i < n - 1; i++) {
        int d = (karrasDelta(C, i, i + 1) - karrasDelta(C, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = karrasDelta(C, i, i - d);
        int lmax = 2;
        while (karrasDelta(C, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (karrasDelta(C, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = karrasDelta(C, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (karrasDelta(C, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int gamma = i + s * d + std::min(d, 0);
        int lo_ = std::min(i, j), hi_ = std::max(i, j);
        uint32_t left = (lo_ == gamma) ? (LeafFlag | (uint32_t)gamma) : (uint32_t)gamma;
        uint32_t right = (hi_ == gamma + 1) ? (LeafFlag | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        B.nodes[i].left = left; B.nodes[i].right = right;
        if (left & LeafFlag) parentL[left & ~LeafFlag] = i; else parentI[left] = i;
        if (right & LeafFlag) parentL[right & ~LeafFlag] = i; else parentI[right] = i;
    }
    // bottom-up refit (recursive post-order; exact unions)
    std::vector<float> nlo((size_t)(n - 1) * 3), nhi((size_t)(n - 1) * 3);
    std::vector<int> stack;
    std::vector<uint8_t> done(n - 1, 0);
    stack.push_back(0);
    auto childBox = [&](uint32_t c, float* cl, float* ch) {
        if (c & LeafFlag) triBounds(B.tris[B.order[c & ~LeafFlag]], cl, ch);
        else for (int a = 0; a < 3; a++) { cl[a] = nlo[c * 3 + a]; ch[a] = nhi[c * 3 + a]; }
    };
    while (!stack.empty()) {
        int i = stack.back();
        BVHNode& nd = B.nodes[i];
        bool ready = true;
        if (!done[i]) {
            if (!(nd.left & LeafFlag) && !done[nd.left]) { stack.push_back(nd.left); ready = false; }
            if (!(nd.right & LeafFlag) && !done[nd.right]) { stack.push_back(nd.right); ready = false; }
        }
        if (!ready) continue;
        stack.pop_back();
        if (done[i]) continue;
        float a0[3], a1[3], b0[3], b1[3];
        childBox(nd.left, a0, a1);
        childBox(nd.right, b0, b1);
        for (int a = 0; a < 3; a++) {
            nd.lmin[a] = a0[a]; nd.lmax[a] = a1[a]; nd.rmin[a] = b0[a]; nd.rmax[a] = b1[a];
            nlo[i * 3 + a] = std::fmin(a0[a], b0[a]);
            nhi[i * 3 + a] = std::fmax(a1[a], b1[a]);
        }
        done[i] = 1;
    }
}

// ---- 4-wide collapse ------------------------------------------------------------------
// Every binary internal node n defines a candidate 4-wide node: start from its two
// children and, while fewer than four, replace the internal child with the largest
// box surface area 2(dx dy + dy dz + dz dx) (first such child on ties) by its two
// children, in place (left child at its position, right child right after). The
// 4-wide tree keeps the root and every internal child of a kept node; kept nodes are
// numbered in increasing binary index. Empty slots hold the ref 0xffffffff and an
// empty box (+inf, -inf).
static inline float boxArea(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct Slot { uint32_t ref; float lo[3], hi[3]; };

static inline int collapseSlots(const BVH& B, uint32_t n, Slot out[4]) {
    const BVHNode& nd = B.nodes[n];
    int cnt = 2;
    out[0].ref = nd.left; out[1].ref = nd.right;
    for (int a = 0; a < 3; a++) {
        out[0].lo[a] = nd.lmin[a]; out[0].hi[a] = nd.lmax[a];
        out[1].lo[a] = nd.rmin[a]; out[1].hi[a] = nd.rmax[a];
    }
    while (cnt < 4) {
        int best = -1; float bestA = -1.0f;
        for (int k = 0; k < cnt; k++) {
            if (out[k].ref & LeafFlag) continue;
            float A = boxArea(out[k].lo, out[k].hi);
            if (best < 0 || A > bestA) { best = k; bestA = A; }
        }
        if (best < 0) break;
        const BVHNode& c = B.nodes[out[best].ref];
        for (int k = cnt; k > best + 1; k--) out[k] = out[k - 1];
        out[best].ref = c.left; out[best + 1].ref = c.right;
        for (int a = 0; a < 3; a++) {
            out[best].lo[a] = c.lmin[a]; out[best].hi[a] = c.lmax[a];
            out[best + 1].lo[a] = c.rmin[a]; out[best + 1].hi[a] = c.rmax[a];
        }
        cnt++;
    }
    return cnt;
}

static inline void collapseBVH4(const BVH& B, std::vector<BVH4Node>& out) {
    out.clear();
    size_t nb = B.nodes.size();
    if (nb == 0) return;
    std::vector<uint8_t> keep(nb, 0);
    std::vector<uint32_t> stack = {0};
    keep[0] = 1;
    while (!stack.empty()) {
        uint32_t n = stack.back(); stack.pop_back();
        Slot s[4]; int cnt = collapseSlots(B, n, s);
        for (int k = 0; k < cnt; k++)
            if (!(s[k].ref & LeafFlag) && !keep[s[k].ref]) { keep[s[k].ref] = 1; stack.push_back(s[k].ref); }
    }
    std::vector<uint32_t> newId(nb, EmptyRef);
    uint32_t m = 0;
    for (size_t i = 0; i < nb; i++) if (keep[i]) newId[i] = m++;
    out.resize(m);
    for (size_t i = 0; i < nb; i++) {
        if (!keep[i]) continue;
        Slot s[4]; int cnt = collapseSlots(B, (uint32_t)i, s);
        BVH4Node& o = out[newId[i]];
        for (int k = 0; k < 4; k++) {
            if (k < cnt) {
                o.child[k] = (s[k].ref & LeafFlag) ? s[k].ref : newId[s[k].ref];
                for (int a = 0; a < 3; a++) { o.lo[a][k] = s[k].lo[a]; o.hi[a][k] = s[k].hi[a]; }
            } else {
                o.child[k] = EmptyRef;
                for (int a = 0; a < 3; a++) { o.lo[a][k] = INFINITY; o.hi[a][k] = -INFINITY; }
            }
        }
    }
}

// ---- ray / box, ray / triangle -------------------------------------------------
struct RayIn { vec3 o, d, invd; float tmin, tmax; };
static inline RayIn makeRayIn(vec3 o, vec3 d, float tmin, float tmax) {
    return {o, d, V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z), tmin, tmax};
}
// A ray with a non-finite origin or direction hits nothing (no triangle test
// can pass with NaN coordinates); it is rejected before traversal, because
// the slab test below would accept every box for it (minNum/maxNum drop the
// NaN operands) and walk the whole tree.
static inline bool rayFinite(const RayIn& r) {
    return std::isfinite(r.o.x) && std::isfinite(r.o.y) && std::isfinite(r.o.z) && std::isfinite(r.d.x) &&
           std::isfinite(r.d.y) && std::isfinite(r.d.z);
}
// Slab test; IEEE minNum/maxNum drop NaN operands (0 * inf).
static inline bool boxHit(const RayIn& r, const float* lo, const float* hi, float tBest, float& tEnter) {
    float tx0 = (lo[0] - r.o.x) * r.invd.x, tx1 = (hi[0] - r.o.x) * r.invd.x;
    float ty0 = (lo[1] - r.o.y) * r.invd.y, ty1 = (hi[1] - r.o.y) * r.invd.y;
    float tz0 = (lo[2] - r.o.z) * r.invd.z, tz1 = (hi[2] - r.o.z) * r.invd.z;
    float tn = std::fmax(std::fmax(std::fmax(std::fmin(tx0, tx1), std::fmin(ty0, ty1)), std::fmin(tz0, tz1)), r.tmin);
    float tf = std::fmin(std::fmin(std::fmin(std::fmax(tx0, tx1), std::fmax(ty0, ty1)), std::fmax(tz0, tz1)), tBest);
    tEnter = tn;
    return tn <= tf;
}

static inline int maxDimAbs(vec3 d) {
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    if (ax > ay) return ax > az ? 0 : 2;
    return ay > az ? 1 : 2;
}
static inline float comp(vec3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Returns true with t, b1, b2 when tmin <= t < tBest.
static inline bool triHit(const RayIn& r, const WorldTri& tri, float tBest, float& tOut, float& b1, float& b2) {
    int kz = maxDimAbs(r.d);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (comp(r.d, kz) < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    float dz = comp(r.d, kz);
    float Sx = comp(r.d, kx) / dz, Sy = comp(r.d, ky) / dz, Sz = 1.0f / dz;
    vec3 A = tri.v0 - r.o, Bv = tri.v1 - r.o, Cv = tri.v2 - r.o;
    float Ax = comp(A, kx) - Sx * comp(A, kz), Ay = comp(A, ky) - Sy * comp(A, kz);
    float Bx = comp(Bv, kx) - Sx * comp(Bv, kz), By = comp(Bv, ky) - Sy * comp(Bv, kz);
    float Cx = comp(Cv, kx) - Sx * comp(Cv, kz), Cy = comp(Cv, ky) - Sy * comp(Cv, kz);
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
    float Az = Sz * comp(A, kz), Bz = Sz * comp(Bv, kz), Cz = Sz * comp(Cv, kz);
    float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f) { T = -T; det = -det; U = -U; V = -V; W = -W; }
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t < tBest)) return false;
    tOut = t; b1 = V * rcp; b2 = W * rcp;
    return true;
}

struct TraceStats { uint64_t nodes = 0, tris = 0, rays = 0; };

struct Hit { Intersection isec; float t; bool hit; };

static inline Hit traceClosest(const BVH& B, const RayIn& r, bool skipLights, TraceStats* st) {
    Hit h; h.hit = false; h.t = r.tmax;
    h.isec.instanceIdx = InvalidHitIndex; h.isec.triangleIdx = 0; h.isec.bary[0] = h.isec.bary[1] = 0.0f;
    if (st) st->rays++;
    if (B.nodes.empty() || !rayFinite(r)) return h;
    uint32_t stack[128]; int sp = 0;
    uint32_t node = 0;
    float tBest = r.tmax;
    auto testLeaf = [&](uint32_t leaf) {
        const WorldTri& tri = B.tris[B.order[leaf]];
        if (skipLights && tri.inst == 0) return;
        if (st) st->tris++;
        float t, b1, b2;
        if (triHit(r, tri, tBest, t, b1, b2)) {
            tBest = t; h.hit = true; h.t = t;
            h.isec.bary[0] = b1; h.isec.bary[1] = b2; h.isec.instanceIdx = tri.inst; h.isec.triangleIdx = tri.prim;
        }
    };
    for (;;) {
        const BVHNode& nd = B.nodes[node];
        if (st) st->nodes++;
        float tl, tr;
        bool hl = boxHit(r, nd.lmin, nd.lmax, tBest, tl);
        bool hr = boxHit(r, nd.rmin, nd.rmax, tBest, tr);
        uint32_t next = 0xffffffffu;
        if (hl && hr) {
            uint32_t nearC = (tl <= tr) ? nd.left : nd.right;
            uint32_t farC = (tl <= tr) ? nd.right : nd.left;
            if (nearC & LeafFlag) testLeaf(nearC & ~LeafFlag); else next = nearC;
            if (farC & LeafFlag) testLeaf(farC & ~LeafFlag);
            else { if (next == 0xffffffffu) next = farC; else stack[sp++] = farC; }
        } else if (hl) {
            if (nd.left & LeafFlag) testLeaf(nd.left & ~LeafFlag); else next = nd.left;
        } else if (hr) {
            if (nd.right & LeafFlag) testLeaf(nd.right & ~LeafFlag); else next = nd.right;
        }
        if (next == 0xffffffffu) {
            if (sp == 0) break;
            next = stack[--sp];
        }
        node = next;
    }
    if (!h.hit) h.isec.instanceIdx = InvalidHitIndex;
    return h;
}

static inline bool traceAny(const BVH& B, const RayIn& r, TraceStats* st) {
    if (st) st->rays++;
    if (B.nodes.empty() || !rayFinite(r)) return false;
    uint32_t stack[128]; int sp = 0;
    uint32_t node = 0;
    const float tmax = r.tmax;
    auto leafHit = [&](uint32_t leaf) {
        if (st) st->tris++;
        float t, b1, b2;
        return triHit(r, B.tris[B.order[leaf]], tmax, t, b1, b2);
    };
    for (;;) {
        const BVHNode& nd = B.nodes[node];
        if (st) st->nodes++;
        float tl, tr;
        bool hl = boxHit(r, nd.lmin, nd.lmax, tmax, tl);
        bool hr = boxHit(r, nd.rmin, nd.rmax, tmax, tr);
        uint32_t next = 0xffffffffu;
        if (hl && hr) {
            uint32_t nearC = (tl <= tr) ? nd.left : nd.right;
            uint32_t farC = (tl <= tr) ? nd.right : nd.left;
            if (nearC & LeafFlag) { if (leafHit(nearC & ~LeafFlag)) return true; } else next = nearC;
            if (farC & LeafFlag) { if (leafHit(farC & ~LeafFlag)) return true; }
            else { if (next == 0xffffffffu) next = farC; else stack[sp++] = farC; }
        } else if (hl) {
            if (nd.left & LeafFlag) { if (leafHit(nd.left & ~LeafFlag)) return true; } else next = nd.left;
        } else if (hr) {
            if (nd.right & LeafFlag) { if (leafHit(nd.right & ~LeafFlag)) return true; } else next = nd.right;
        }
        if (next == 0xffffffffu) {
            if (sp == 0) break;
            next = stack[--sp];
        }
        node = next;
    }
    return false;
}

// 4-wide traversal. At each node the hit children are ordered by entry distance
// (slot order on ties); leaves are tested in that order as they come, the nearest
// internal child is visited next and the other internal children are pushed
// farthest first. A closest-hit query keeps each pushed node's entry distance,
// truncated to the float's top 16 bits (a value <= the distance; -inf for a
// negative one), and skips a popped node whose stored distance exceeds the
// closest hit so far: all of its children would miss their box tests, since a
// node's box holds its children's boxes and the slab distances are monotone in
// the bounds. Hits are those of the full walk; the work counters are lower.
static inline uint16_t stackKey(float t) {
    uint32_t b; std::memcpy(&b, &t, 4);
    return (b >> 31) ? (uint16_t)0xff80u : (uint16_t)(b >> 16);
}
static inline float stackKeyValue(uint16_t k) {
    const uint32_t b = (uint32_t)k << 16;
    float t; std::memcpy(&t, &b, 4);
    return t;
}

template <bool kAny>
static inline bool traverse4(const BVH& B, const RayIn& r, bool skipLights, float& tBest, Hit& h, TraceStats* st) {
    if (B.nodes4.empty() || !rayFinite(r)) return false;
    uint32_t stack[512]; uint16_t stackT[512]; int sp = 0;
    uint32_t node = 0;
    bool found = false;
    for (;;) {
        const BVH4Node& nd = B.nodes4[node];
        if (st) st->nodes++;
        int ord[4]; float tk[4]; int nh = 0;
        for (int k = 0; k < 4; k++) {
            if (nd.child[k] == EmptyRef) continue;
            float lo[3] = {nd.lo[0][k], nd.lo[1][k], nd.lo[2][k]}, hi[3] = {nd.hi[0][k], nd.hi[1][k], nd.hi[2][k]};
            float te;
            if (!boxHit(r, lo, hi, tBest, te)) continue;
            int j = nh++;
            while (j > 0 && tk[j - 1] > te) { tk[j] = tk[j - 1]; ord[j] = ord[j - 1]; j--; }
            tk[j] = te; ord[j] = k;
        }
        uint32_t next = EmptyRef;
        uint32_t pend[3]; uint16_t pendT[3]; int np = 0;
        for (int q = 0; q < nh; q++) {
            uint32_t ref = nd.child[ord[q]];
            if (ref & LeafFlag) {
                const WorldTri& tri = B.tris[B.order[ref & ~LeafFlag]];
                if (skipLights && tri.inst == 0) continue;
                if (st) st->tris++;
                float t, b1, b2;
                if (triHit(r, tri, tBest, t, b1, b2)) {
                    found = true;
                    if (kAny) return true;
                    tBest = t; h.t = t;
                    h.isec.bary[0] = b1; h.isec.bary[1] = b2; h.isec.instanceIdx = tri.inst; h.isec.triangleIdx = tri.prim;
                }
            } else if (next == EmptyRef) {
                next = ref;
            } else {
                pendT[np] = stackKey(tk[q]);
                pend[np++] = ref;
            }
        }
        while (np > 0) { np--; stack[sp] = pend[np]; stackT[sp] = pendT[np]; sp++; }
        while (next == EmptyRef && sp > 0) {
            sp--;
            if (kAny || !(stackKeyValue(stackT[sp]) > tBest)) next = stack[sp];
        }
        if (next == EmptyRef) break;
        node = next;
    }
    return found;
}

static inline Hit traceClosest4(const BVH& B, const RayIn& r, bool skipLights, TraceStats* st) {
    Hit h; h.hit = false; h.t = r.tmax;
    h.isec.instanceIdx = InvalidHitIndex; h.isec.triangleIdx = 0; h.isec.bary[0] = h.isec.bary[1] = 0.0f;
    float tBest = r.tmax;
    bool hit = traverse4<false>(B, r, skipLights, tBest, h, st);
    if (st) st->rays++;
    h.hit = hit;
    if (!hit) { h.isec.instanceIdx = InvalidHitIndex; h.isec.triangleIdx = 0; h.isec.bary[0] = h.isec.bary[1] = 0.0f; }
    return h;
}
static inline bool traceAny4(const BVH& B, const RayIn& r, TraceStats* st) {
    Hit h; float tBest = r.tmax;
    bool hit = traverse4<true>(B, r, false, tBest, h, st);
    if (st) st->rays++;
    return hit;
}

// Brute force over all primitives in primitive order (parity reference for
// traversal; ties are resolved by primitive order here).
static inline Hit traceBrute(const BVH& B, const RayIn& r, bool skipLights) {
    Hit h; h.hit = false; h.t = r.tmax; h.isec.instanceIdx = InvalidHitIndex; h.isec.triangleIdx = 0;
    h.isec.bary[0] = h.isec.bary[1] = 0.0f;
    float tBest = r.tmax;
    for (size_t i = 0; i < B.tris.size(); i++) {
        const WorldTri& tri = B.tris[i];
        if (skipLights && tri.inst == 0) continue;
        float t, b1, b2;
        if (triHit(r, tri, tBest, t, b1, b2)) {
            tBest = t; h.hit = true; h.t = t;
            h.isec.bary[0] = b1; h.isec.bary[1] = b2; h.isec.instanceIdx = tri.inst; h.isec.triangleIdx = tri.prim;
        }
    }
    return h;
}

}  // namespace orc
