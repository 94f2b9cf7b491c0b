// Procedural workloads (the reference loads XML scenes through assimp; these
// generators produce the same data model directly). All randomness comes from
// PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r) streams seeded by the
// scene seed, so every scene and camera path is reproducible.
#include "Scene.h"
#include <cmath>
#include <cstring>
#include <map>

namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t initstate, uint64_t initseq) {
        state = 0u; inc = (initseq << 1u) | 1u;
        next(); state += initstate; next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-(int32_t)rot) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
};

struct Mesh { std::vector<Vec3> pos; std::vector<uint32_t> tri; };

Mesh icosphere(int subdiv) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    Mesh m;
    float v[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) {
        float l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        m.pos.push_back({p[0] / l, p[1] / l, p[2] / l});
    }
    m.tri = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
             3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < subdiv; s++) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            Vec3 p = {(m.pos[a].x + m.pos[b].x) * 0.5f, (m.pos[a].y + m.pos[b].y) * 0.5f, (m.pos[a].z + m.pos[b].z) * 0.5f};
            float l = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
            m.pos.push_back({p.x / l, p.y / l, p.z / l});
            uint32_t id = (uint32_t)m.pos.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nt;
        for (size_t i = 0; i < m.tri.size(); i += 3) {
            uint32_t a = m.tri[i], b = m.tri[i + 1], c = m.tri[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nt.insert(nt.end(), add, add + 12);
        }
        m.tri.swap(nt);
    }
    return m;
}

rt_material lambert(float r, float g, float b) {
    rt_material m{};
    m.baseColor[0] = r; m.baseColor[1] = g; m.baseColor[2] = b;
    m.type = 1; m.textureIdx = 0xffffffffu; m.metallic = 0.0f; m.roughness = 1.0f; m.ior = 1.5f;
    return m;
}
rt_material ggx(float r, float g, float b, float roughness, float metallic) {
    rt_material m = lambert(r, g, b);
    m.type = 2; m.roughness = roughness; m.metallic = metallic;
    return m;
}

void identityM(float* m) { std::memset(m, 0, 64); m[0] = m[5] = m[10] = m[15] = 1.0f; }
void translateScale(float* m, Vec3 t, float s) {
    identityM(m);
    m[0] = m[5] = m[10] = s;
    m[12] = t.x; m[13] = t.y; m[14] = t.z;
}

// Adds a mesh (object-space vertices with normals) as a unique model; returns
// its index offset. Per-triangle material `mat`.
uint32_t addModel(Scene& S, const std::vector<rt_mesh_vertex>& verts, const std::vector<uint32_t>& tris,
                  const std::vector<int32_t>& triMat) {
    uint32_t vbase = (uint32_t)S.vertices.size();
    uint32_t ioff = (uint32_t)S.indices.size();
    S.vertices.insert(S.vertices.end(), verts.begin(), verts.end());
    for (uint32_t i : tris) S.indices.push_back(vbase + i);
    S.materialIndices.insert(S.materialIndices.end(), triMat.begin(), triMat.end());
    return ioff;
}
// Same vertices, new index range (a second unique model sharing vertex data).
uint32_t addModelSharedVerts(Scene& S, uint32_t vbase, const std::vector<uint32_t>& tris, int32_t mat) {
    uint32_t ioff = (uint32_t)S.indices.size();
    for (uint32_t i : tris) S.indices.push_back(vbase + i);
    for (size_t t = 0; t < tris.size() / 3; t++) S.materialIndices.push_back(mat);
    return ioff;
}
void addInstance(Scene& S, uint32_t indexOffset, uint32_t indexCount, const float* m) {
    rt_object_instance inst{};
    makeInstanceTransform(inst, m);
    inst.indexOffset = indexOffset;
    inst.indexCount = indexCount;
    S.objectInstances.push_back(inst);
}
rt_mesh_vertex vtx(Vec3 p, Vec3 n) {
    rt_mesh_vertex v{};
    v.pos[0] = p.x; v.pos[1] = p.y; v.pos[2] = p.z;
    v.norm[0] = n.x; v.norm[1] = n.y; v.norm[2] = n.z;
    return v;
}
// quad p0 p1 p2 p3 (counter-clockwise seen from the side `n` points to)
void addQuadModel(Scene& S, Vec3 p0, Vec3 p1, Vec3 p2, Vec3 p3, Vec3 n, int32_t mat) {
    std::vector<rt_mesh_vertex> v = {vtx(p0, n), vtx(p1, n), vtx(p2, n), vtx(p3, n)};
    uint32_t ioff = addModel(S, v, {0, 1, 2, 0, 2, 3}, {mat, mat});
    float m[16]; identityM(m);
    addInstance(S, ioff, 6, m);
}

// ---- value noise fBm for the heightfield ---------------------------------------------------
struct Noise {
    uint32_t seed;
    float lattice(int x, int y) const {
        uint32_t h = (uint32_t)x * 374761393u + (uint32_t)y * 668265263u + seed * 2246822519u;
        h = (h ^ (h >> 13)) * 1274126177u;
        h ^= h >> 16;
        return (float)(h & 0xffffffu) / 16777215.0f * 2.0f - 1.0f;
    }
    float value(float x, float y) const {
        int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
        float fx = x - x0, fy = y - y0;
        float sx = fx * fx * (3.0f - 2.0f * fx), sy = fy * fy * (3.0f - 2.0f * fy);
        float a = lattice(x0, y0), b = lattice(x0 + 1, y0), c = lattice(x0, y0 + 1), d = lattice(x0 + 1, y0 + 1);
        return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sy;
    }
    float fbm(float x, float y, int octaves) const {
        float s = 0.0f, amp = 0.5f, f = 1.0f;
        for (int o = 0; o < octaves; o++) { s += amp * value(x * f, y * f); f *= 2.0f; amp *= 0.5f; }
        return s;
    }
};

// Value noise on the integer lattice of 3-space (trilinear, smoothstep weights), in [-1, 1].
float valueNoise3(uint32_t seed, float x, float y, float z) {
    auto lat = [&](int i, int j, int k) {
        uint32_t h = (uint32_t)i * 374761393u + (uint32_t)j * 668265263u + (uint32_t)k * 3266489917u + seed * 2246822519u;
        h = (h ^ (h >> 13)) * 1274126177u;
        h ^= h >> 16;
        return (float)(h & 0xffffffu) / 16777215.0f * 2.0f - 1.0f;
    };
    int x0 = (int)std::floor(x), y0 = (int)std::floor(y), z0 = (int)std::floor(z);
    float fx = x - x0, fy = y - y0, fz = z - z0;
    float sx = fx * fx * (3.0f - 2.0f * fx), sy = fy * fy * (3.0f - 2.0f * fy), sz = fz * fz * (3.0f - 2.0f * fz);
    auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
    float c00 = lerp(lat(x0, y0, z0), lat(x0 + 1, y0, z0), sx), c10 = lerp(lat(x0, y0 + 1, z0), lat(x0 + 1, y0 + 1, z0), sx);
    float c01 = lerp(lat(x0, y0, z0 + 1), lat(x0 + 1, y0, z0 + 1), sx);
    float c11 = lerp(lat(x0, y0 + 1, z0 + 1), lat(x0 + 1, y0 + 1, z0 + 1), sx);
    return lerp(lerp(c00, c10, sy), lerp(c01, c11, sy), sz);
}

// The icosphere with each vertex pushed along its direction by up to +-30 % (3-D value noise,
// two octaves) and smooth normals rebuilt from the displaced faces (area-weighted sums). The
// surface stays star-shaped around the centre, so it does not intersect itself.
std::vector<rt_mesh_vertex> displacedIcosphere(const Mesh& ico, uint32_t seed) {
    std::vector<Vec3> p(ico.pos.size());
    for (size_t i = 0; i < p.size(); i++) {
        const Vec3& d = ico.pos[i];
        float n = 0.7f * valueNoise3(seed, 2.5f * d.x, 2.5f * d.y, 2.5f * d.z) +
                  0.3f * valueNoise3(seed + 1u, 5.0f * d.x, 5.0f * d.y, 5.0f * d.z);
        float r = 1.0f + 0.3f * n;
        p[i] = {d.x * r, d.y * r, d.z * r};
    }
    std::vector<Vec3> nrm(p.size(), Vec3{0.0f, 0.0f, 0.0f});
    for (size_t t = 0; t < ico.tri.size(); t += 3) {
        const Vec3 &a = p[ico.tri[t]], &b = p[ico.tri[t + 1]], &c = p[ico.tri[t + 2]];
        Vec3 e1 = {b.x - a.x, b.y - a.y, b.z - a.z}, e2 = {c.x - a.x, c.y - a.y, c.z - a.z};
        Vec3 f = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
        // outward: the face normal points away from the centre
        if (f.x * (a.x + b.x + c.x) + f.y * (a.y + b.y + c.y) + f.z * (a.z + b.z + c.z) < 0.0f) f = {-f.x, -f.y, -f.z};
        for (int k = 0; k < 3; k++) {
            Vec3& q = nrm[ico.tri[t + k]];
            q = {q.x + f.x, q.y + f.y, q.z + f.z};
        }
    }
    std::vector<rt_mesh_vertex> v;
    for (size_t i = 0; i < p.size(); i++) {
        float l = std::sqrt(nrm[i].x * nrm[i].x + nrm[i].y * nrm[i].y + nrm[i].z * nrm[i].z);
        v.push_back(vtx(p[i], {nrm[i].x / l, nrm[i].y / l, nrm[i].z / l}));
    }
    return v;
}

// ---- configs -----------------------------------------------------------------------
std::unique_ptr<Scene> boxRoom(uint32_t seed) {
    auto S = std::make_unique<Scene>();
    Pcg32 rng(seed, 1);
    // 6 faces of [-1,1] x [-1,1] x [0,2], each split into two quads, normals inward
    struct Face { Vec3 o, u, v, n; };
    Face faces[6] = {
        {{-1, -1, 0}, {2, 0, 0}, {0, 2, 0}, {0, 0, 1}},    // floor
        {{-1, 1, 2}, {2, 0, 0}, {0, -2, 0}, {0, 0, -1}},   // ceiling
        {{-1, 1, 0}, {2, 0, 0}, {0, 0, 2}, {0, -1, 0}},    // back wall (y = 1)
        {{1, -1, 0}, {-2, 0, 0}, {0, 0, 2}, {0, 1, 0}},    // front wall (y = -1)
        {{-1, -1, 0}, {0, 2, 0}, {0, 0, 2}, {1, 0, 0}},    // left wall (x = -1)
        {{1, 1, 0}, {0, -2, 0}, {0, 0, 2}, {-1, 0, 0}},    // right wall (x = 1)
    };
    for (int f = 0; f < 6; f++) {
        for (int h = 0; h < 2; h++) {
            const Face& F = faces[f];
            Vec3 u = {F.u.x * 0.5f, F.u.y * 0.5f, F.u.z * 0.5f};
            Vec3 o = {F.o.x + u.x * h, F.o.y + u.y * h, F.o.z + u.z * h};
            Vec3 p0 = o, p1 = {o.x + u.x, o.y + u.y, o.z + u.z};
            Vec3 p2 = {p1.x + F.v.x, p1.y + F.v.y, p1.z + F.v.z}, p3 = {o.x + F.v.x, o.y + F.v.y, o.z + F.v.z};
            // orient so that cross(p1-p0, p2-p0) follows the inward normal
            Vec3 e1 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z}, e2 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
            Vec3 c = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
            int32_t mat = (int32_t)S->materials.size();
            S->materials.push_back(lambert(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f)));
            if (c.x * F.n.x + c.y * F.n.y + c.z * F.n.z > 0) addQuadModel(*S, p0, p1, p2, p3, F.n, mat);
            else addQuadModel(*S, p0, p3, p2, p1, F.n, mat);
        }
    }
    // 64 icospheres: one shared vertex set, one index range (model) per sphere
    Mesh ico = icosphere(2);
    std::vector<rt_mesh_vertex> iv;
    for (const Vec3& p : ico.pos) iv.push_back(vtx(p, p));
    uint32_t vbase = (uint32_t)S->vertices.size();
    S->vertices.insert(S->vertices.end(), iv.begin(), iv.end());
    for (int i = 0; i < 64; i++) {
        float r = rng.uniform(0.05f, 0.2f);
        Vec3 c = {rng.uniform(-1.0f + r, 1.0f - r), rng.uniform(-1.0f + r, 1.0f - r), rng.uniform(r, 2.0f - r)};
        int32_t mat = (int32_t)S->materials.size();
        S->materials.push_back(lambert(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f)));
        uint32_t ioff = addModelSharedVerts(*S, vbase, ico.tri, mat);
        float m[16]; translateScale(m, c, r);
        addInstance(*S, ioff, (uint32_t)ico.tri.size(), m);
    }
    // ceiling quad light, radiance 10, facing down
    float rad[3] = {10.0f, 10.0f, 10.0f};
    float a[3] = {-0.3f, -0.3f, 1.99f}, b[3] = {0.3f, -0.3f, 1.99f}, c[3] = {0.3f, 0.3f, 1.99f}, d[3] = {-0.3f, 0.3f, 1.99f};
    S->addLightTriangle(a, c, b, rad);
    S->addLightTriangle(a, d, c, rad);
    S->buildLightDataStructure();
    S->fov = 70.0f;
    Pcg32 cam(seed, 7);
    Scene::PathKey k{{0.0f, -0.95f, 1.0f}, {0.0f, 0.0f, 0.0f}, 0};
    for (int f = 0; f < 64; f++) { k.seed = cam.next(); S->cameraPath.push_back(k); }
    return S;
}

// fBm heightfield of N x N cells (4 triangles per cell) over 24 units, numSpheres
// instanced icospheres (8 unique models) and numLights emissive triangles;
// GGX roughness uniform in [roughMin, 1], metallic with p = 0.3.
std::unique_ptr<Scene> terrain(uint32_t seed, int N = 512, int numSpheres = 200, int numLights = 16,
                               float roughMin = 0.05f, bool displaced = false) {
    auto S = std::make_unique<Scene>();
    Pcg32 rng(seed, 1);
    const float extent = 24.0f, cell = extent / N, amp = 2.4f;
    Noise noise{seed * 7919u + 13u};
    auto height = [&](float x, float y) { return amp * noise.fbm(x * 0.125f + 100.0f, y * 0.125f + 100.0f, 6); };
    auto normalAt = [&](float x, float y) {
        float e = cell * 0.5f;
        float dx = (height(x + e, y) - height(x - e, y)) / (2 * e);
        float dy = (height(x, y + e) - height(x, y - e)) / (2 * e);
        Vec3 n = {-dx, -dy, 1.0f};
        float l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
        return Vec3{n.x / l, n.y / l, n.z / l};
    };
    // 64 patch materials (8 x 8 blocks of N/8 x N/8 cells)
    int32_t matBase = (int32_t)S->materials.size();
    for (int i = 0; i < 64; i++) {
        float metallic = rng.uniform() < 0.3f ? 1.0f : 0.0f;
        S->materials.push_back(ggx(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f),
                                   rng.uniform(roughMin, 1.0f), metallic));
    }
    std::vector<rt_mesh_vertex> v;
    v.reserve((size_t)(N + 1) * (N + 1) + (size_t)N * N);
    const float x0 = -extent * 0.5f;
    for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++) {
            float x = x0 + i * cell, y = x0 + j * cell;
            v.push_back(vtx({x, y, height(x, y)}, normalAt(x, y)));
        }
    uint32_t centreBase = (uint32_t)v.size();
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            float x = x0 + (i + 0.5f) * cell, y = x0 + (j + 0.5f) * cell;
            v.push_back(vtx({x, y, height(x, y)}, normalAt(x, y)));
        }
    std::vector<uint32_t> tris;
    std::vector<int32_t> triMat;
    tris.reserve((size_t)N * N * 12);
    triMat.reserve((size_t)N * N * 4);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            uint32_t a = j * (N + 1) + i, b = a + 1, c = a + (N + 1) + 1, d = a + (N + 1);
            uint32_t m = centreBase + j * N + i;
            uint32_t q[12] = {a, b, m, b, c, m, c, d, m, d, a, m};
            tris.insert(tris.end(), q, q + 12);
            int32_t mat = matBase + (j * 8 / N) * 8 + (i * 8 / N);
            for (int t = 0; t < 4; t++) triMat.push_back(mat);
        }
    uint32_t ioff = addModel(*S, v, tris, triMat);
    float id[16]; identityM(id);
    addInstance(*S, ioff, (uint32_t)tris.size(), id);

    // instanced icospheres: 8 unique models (own material; with `displaced`, own displaced
    // shape), instanced round robin
    Mesh ico = icosphere(2);
    std::vector<rt_mesh_vertex> iv;
    for (const Vec3& p : ico.pos) iv.push_back(vtx(p, p));
    uint32_t vbase = (uint32_t)S->vertices.size();
    if (!displaced) S->vertices.insert(S->vertices.end(), iv.begin(), iv.end());
    uint32_t modelOff[8];
    for (int k = 0; k < 8; k++) {
        int32_t mat = (int32_t)S->materials.size();
        float metallic = rng.uniform() < 0.3f ? 1.0f : 0.0f;
        S->materials.push_back(ggx(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f),
                                   rng.uniform(roughMin, 1.0f), metallic));
        if (displaced) {
            std::vector<int32_t> triMat(ico.tri.size() / 3, mat);
            modelOff[k] = addModel(*S, displacedIcosphere(ico, seed * 31u + (uint32_t)k), ico.tri, triMat);
        } else {
            modelOff[k] = addModelSharedVerts(*S, vbase, ico.tri, mat);
        }
    }
    for (int i = 0; i < numSpheres; i++) {
        float r = rng.uniform(0.12f, 0.6f);
        float x = rng.uniform(-10.8f, 10.8f), y = rng.uniform(-10.8f, 10.8f);
        float z = height(x, y) + r * rng.uniform(0.2f, 1.0f);
        float m[16]; translateScale(m, {x, y, z}, r);
        addInstance(*S, modelOff[i % 8], (uint32_t)ico.tri.size(), m);
    }
    // emissive triangles above the terrain, facing down, radiance log-uniform in [1, 100]
    for (int i = 0; i < numLights; i++) {
        float cx = rng.uniform(-9.6f, 9.6f), cy = rng.uniform(-9.6f, 9.6f);
        float cz = height(cx, cy) + rng.uniform(2.0f, 5.0f);
        float s = rng.uniform(0.25f, 0.75f), ang = rng.uniform(0.0f, 6.2831853f);
        float L = std::exp(rng.uniform(0.0f, std::log(100.0f)));
        float p[3][3];
        for (int k = 0; k < 3; k++) {
            float a = ang + k * 2.0943951f;
            p[k][0] = cx + s * std::cos(a); p[k][1] = cy + s * std::sin(a); p[k][2] = cz + rng.uniform(-0.07f, 0.07f);
        }
        float rad[3] = {L, L * rng.uniform(0.8f, 1.0f), L * rng.uniform(0.7f, 1.0f)};
        // counter-clockwise seen from above => normal up; swap to face down
        S->addLightTriangle(p[0], p[2], p[1], rad);
    }
    S->buildLightDataStructure();
    S->fov = 60.0f;
    // 16-frame seeded small-rotation random walk
    Pcg32 cam(seed, 7);
    Scene::PathKey k{{0.0f, -12.5f, 5.5f}, {0.0f, -22.0f, 0.0f}, 0};
    for (int f = 0; f < 64; f++) {
        k.seed = cam.next();
        S->cameraPath.push_back(k);
        k.angle.x += cam.uniform(-0.5f, 0.5f);
        k.angle.y += cam.uniform(-0.25f, 0.25f);
        k.pos.x += cam.uniform(-0.01f, 0.01f);
        k.pos.y += cam.uniform(-0.01f, 0.01f);
        k.pos.z += cam.uniform(-0.01f, 0.01f);
    }
    return S;
}

std::unique_ptr<Scene> floorUnderLight(uint32_t seed) {
    auto S = std::make_unique<Scene>();
    S->materials.push_back(lambert(0.5f, 0.5f, 0.5f));
    addQuadModel(*S, {-50, -50, 0}, {50, -50, 0}, {50, 50, 0}, {-50, 50, 0}, {0, 0, 1}, 0);
    float rad[3] = {4.0f, 4.0f, 4.0f};
    float a[3] = {-0.5f, -0.5f, 1.0f}, b[3] = {0.5f, -0.5f, 1.0f}, c[3] = {0.5f, 0.5f, 1.0f}, d[3] = {-0.5f, 0.5f, 1.0f};
    S->addLightTriangle(a, c, b, rad);
    S->addLightTriangle(a, d, c, rad);
    S->buildLightDataStructure();
    S->fov = 40.0f;
    Pcg32 cam(seed, 7);
    Scene::PathKey k{{0.0f, -3.0f, 3.0f}, {0.0f, -45.0f, 0.0f}, 0};
    for (int f = 0; f < 64; f++) { k.seed = cam.next(); S->cameraPath.push_back(k); }
    return S;
}

}  // namespace

std::unique_ptr<Scene> generateScene(uint32_t config, uint32_t seed) {
    switch (config) {
    case 0: return boxRoom(seed);
    case 1: return terrain(seed);
    case 2: return floorUnderLight(seed);
    case 3: return terrain(seed, 200);                        // ~225k triangles
    case 4: return terrain(seed, 860);                        // ~3.0M triangles
    case 5: return terrain(seed, 1024, 2000, 64, 0.02f, true); // ~4.8M triangles, 64 lights, displaced spheres
    default: return nullptr;
    }
}
