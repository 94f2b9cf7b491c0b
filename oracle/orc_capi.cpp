// ORACLE — test infrastructure only. Only tests/, the smoke check and the CPU
// baseline leg of the benchmark load this library; the product never does.
//
// C entry points over the scalar restatement, so that Python tests can drive it
// with ctypes. Frames are split into row bands over std::thread workers.
#include "orc_passes.h"
#include "orc_guide.h"
#include <atomic>
#include <thread>

using namespace orc;

namespace {

struct OScene {
    SceneArrays S;
    // learned-proposal bindings of the GRIS passes (orc_gris_bind_guide)
    const float* guideLobes = nullptr;
    const float* guideLobesPrev = nullptr;
    float* guideSamples = nullptr;
    // private copies so the caller's arrays may be freed
    std::vector<Material> materials;
    std::vector<int32_t> materialIndices;
    std::vector<MeshVertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<ObjectInstance> instances;
    std::vector<TriangleLight> lights;
    std::vector<LightSampleTableElement> table;
    std::vector<Texture> textures;
    std::vector<std::vector<uint8_t>> texels;
    BVH bvh;
};

template <class F>
void parallelRows(uint32_t H, int nthreads, F&& f) {
    if (nthreads <= 1) { for (uint32_t y = 0; y < H; y++) f(y, 0); return; }
    std::atomic<uint32_t> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        th.emplace_back([&, t] {
            for (;;) {
                uint32_t y = next.fetch_add(1);
                if (y >= H) break;
                f(y, t);
            }
        });
    }
    for (auto& x : th) x.join();
}

void sumStats(const std::vector<TraceStats>& st, uint64_t* out) {
    if (!out) return;
    uint64_t r = 0, n = 0, t = 0;
    for (auto& s : st) { r += s.rays; n += s.nodes; t += s.tris; }
    out[0] += r; out[1] += n; out[2] += t;
}

inline void accumulate(float* out, vec3 c, uint32_t frameIndex) {
    float k = (float)(frameIndex & CameraFrameIndexMask);
    vec3 acc = V3(out[0], out[1], out[2]);
    acc = (acc * k + c) / (k + 1.0f);
    out[0] = acc.x; out[1] = acc.y; out[2] = acc.z; out[3] = 1.0f;
}

}  // namespace

extern "C" {

// ---- scene ---------------------------------------------------------------------
void* orc_scene_create(const SceneArrays* in) {
    OScene* s = new OScene();
    s->materials.assign(in->materials, in->materials + in->numMaterials);
    s->materialIndices.assign(in->materialIndices, in->materialIndices + in->numMaterialIndices);
    s->vertices.assign(in->vertices, in->vertices + in->numVertices);
    s->indices.assign(in->indices, in->indices + in->numIndices);
    s->instances.assign(in->instances, in->instances + in->numInstances);
    s->lights.assign(in->lights, in->lights + in->numLights);
    s->table.assign(in->lightTable, in->lightTable + in->numLightTable);
    s->texels.resize(in->numTextures);
    s->textures.resize(in->numTextures);
    for (uint32_t i = 0; i < in->numTextures; i++) {
        const Texture& t = in->textures[i];
        s->texels[i].assign(t.rgba8, t.rgba8 + 4 * (size_t)t.width * t.height);
        s->textures[i] = t;
        s->textures[i].rgba8 = s->texels[i].data();
    }
    s->S = {s->materials.data(), (uint32_t)s->materials.size(), s->materialIndices.data(),
            (uint32_t)s->materialIndices.size(), s->vertices.data(), (uint32_t)s->vertices.size(),
            s->indices.data(), (uint32_t)s->indices.size(), s->instances.data(), (uint32_t)s->instances.size(),
            s->lights.data(), (uint32_t)s->lights.size(), s->table.data(), (uint32_t)s->table.size(),
            s->textures.data(), (uint32_t)s->textures.size()};
    buildWorldTris(s->S, s->bvh.tris);
    buildBVH(s->bvh);
    collapseBVH4(s->bvh, s->bvh.nodes4);
    return s;
}

// 4-wide nodes as 28 words each: lo[3][4], hi[3][4] (float bits), child[4].
uint32_t orc_bvh4_num_nodes(void* h) { return (uint32_t)((OScene*)h)->bvh.nodes4.size(); }
void orc_bvh4_export(void* h, uint32_t* nodes28) {
    const BVH& B = ((OScene*)h)->bvh;
    for (size_t i = 0; i < B.nodes4.size(); i++) std::memcpy(nodes28 + 28 * i, &B.nodes4[i], 112);
}
void orc_scene_destroy(void* h) { delete (OScene*)h; }
uint32_t orc_scene_num_prims(void* h) { return (uint32_t)((OScene*)h)->bvh.tris.size(); }

// World-space triangles in primitive order: 9 floats + (inst, prim) per primitive.
void orc_scene_world_tris(void* h, float* pos9, uint32_t* ref2) {
    OScene* s = (OScene*)h;
    for (size_t i = 0; i < s->bvh.tris.size(); i++) {
        const WorldTri& t = s->bvh.tris[i];
        float* p = pos9 + 9 * i;
        p[0] = t.v0.x; p[1] = t.v0.y; p[2] = t.v0.z; p[3] = t.v1.x; p[4] = t.v1.y; p[5] = t.v1.z;
        p[6] = t.v2.x; p[7] = t.v2.y; p[8] = t.v2.z;
        ref2[2 * i] = t.inst; ref2[2 * i + 1] = t.prim;
    }
}

// LBVH arrays: sorted codes [N], leaf order [N], internal nodes [N-1] (14 words each:
// lmin3 lmax3 rmin3 rmax3 as float bits, left, right), scene box [6].
void orc_bvh_export(void* h, uint64_t* codes, uint32_t* order, uint32_t* nodes14, float* box6) {
    OScene* s = (OScene*)h;
    const BVH& B = s->bvh;
    for (size_t i = 0; i < B.codes.size(); i++) { codes[i] = B.codes[i]; order[i] = B.order[i]; }
    for (size_t i = 0; i < B.nodes.size(); i++) std::memcpy(nodes14 + 14 * i, &B.nodes[i], 56);
    for (int a = 0; a < 3; a++) { box6[a] = B.sceneLo[a]; box6[3 + a] = B.sceneHi[a]; }
}
uint32_t orc_bvh_num_nodes(void* h) { return (uint32_t)((OScene*)h)->bvh.nodes.size(); }

// rays8: ox oy oz dx dy dz tmin tmax. mode 0 closest, 1 any (4-wide tree), 2 brute force
// closest, 3 closest and 4 any over the binary tree.
// out: t, bary(2), inst, prim per ray (inst = 0xffffffff on miss; any-hit: inst = hit?1:0)
void orc_trace_rays(void* h, const float* rays8, uint32_t n, int mode, int skipLights, float* outT, float* outBary,
                    uint32_t* outInst, uint32_t* outPrim, uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    const uint32_t rowsPer = 256;
    uint32_t rows = (n + rowsPer - 1) / rowsPer;
    parallelRows(rows, nthreads, [&](uint32_t r, int t) {
        for (uint32_t i = r * rowsPer; i < n && i < (r + 1) * rowsPer; i++) {
            const float* q = rays8 + 8 * (size_t)i;
            RayIn ray = makeRayIn(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), q[6], q[7]);
            if (mode == 1 || mode == 4) {
                bool hit = mode == 1 ? traceAny4(s->bvh, ray, &st[t]) : traceAny(s->bvh, ray, &st[t]);
                outInst[i] = hit ? 1u : 0u; outT[i] = 0; outBary[2 * i] = outBary[2 * i + 1] = 0; outPrim[i] = 0;
                continue;
            }
            Hit hh = (mode == 2)   ? traceBrute(s->bvh, ray, skipLights != 0)
                     : (mode == 3) ? traceClosest(s->bvh, ray, skipLights != 0, &st[t])
                                   : traceClosest4(s->bvh, ray, skipLights != 0, &st[t]);
            outT[i] = hh.hit ? hh.t : 0.0f;
            outBary[2 * i] = hh.isec.bary[0]; outBary[2 * i + 1] = hh.isec.bary[1];
            outInst[i] = hh.isec.instanceIdx; outPrim[i] = hh.isec.triangleIdx;
        }
    });
    sumStats(st, stats3);
}

// ---- G-buffer -----------------------------------------------------------------------
void orc_gbuffer(void* h, const Camera* cam, float* depthNormal, uint32_t* albedoMatId, uint16_t* motion,
                 uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    const uint32_t W = cam->filmSize[0], H = cam->filmSize[1];
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            gbufferPixel(c, *cam, x, y, depthNormal + 4 * p, albedoMatId + 2 * p, motion + 2 * p);
        }
    });
    sumStats(st, stats3);
}

// ---- plain PT (di_naive + gi_naive, with their .comp accumulation) --------------------------
void orc_pt_plain(void* h, const Camera* cam, const float* depthNormal, const uint32_t* albedoMatId, uint32_t maxDepth,
                  float* directOut, float* indirectOut, uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    const uint32_t W = cam->filmSize[0], H = cam->filmSize[1];
    GBufView g{depthNormal, albedoMatId, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            if (directOut) accumulate(directOut + 4 * p, directIllumination2(c, *cam, g, x, y), cam->frameIndex);
            if (indirectOut) accumulate(indirectOut + 4 * p, indirectIllumination(c, *cam, g, x, y, maxDepth), cam->frameIndex);
        }
    });
    sumStats(st, stats3);
}

// ---- ReSTIR GI temporal -----------------------------------------------------------------------
void orc_gi_temporal(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                     const uint16_t* motion, const float* depthNormalPrev, const uint32_t* albedoMatIdPrev,
                     GIReservoir* resvThis, const GIReservoir* resvPrev, float* indirectOut, uint64_t* stats3,
                     int nthreads) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, motion, W, H};
    GBufView gp{depthNormalPrev, albedoMatIdPrev, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            vec3 col = clampColor(giTemporal(c, cam, cams[1], g, gp, resvThis, resvPrev, x, y));
            accumulate(indirectOut + 4 * ((size_t)y * W + x), col, cam.frameIndex);
        }
    });
    sumStats(st, stats3);
}

// ---- ReSTIR DI passes ------------------------------------------------------------------------
// cams[0] = uCamera, cams[1] = uPrevCamera; reservoirs are the reference's 64-B DIReservoir.
void orc_di_path_gen(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                     const DISettings* set, DIReservoir* resvThis, uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            DIReservoir r;
            if (diPathGen(c, cam, g, *set, x, y, r)) resvThis[(size_t)y * W + x] = r;
        }
    });
    sumStats(st, stats3);
}

void orc_di_temporal(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                     const uint16_t* motion, const float* depthNormalPrev, const uint32_t* albedoMatIdPrev,
                     const DISettings* set, const DIReservoir* resvThis, const DIReservoir* resvPrev,
                     DIReservoir* resvTemp, uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, motion, W, H};
    GBufView gp{depthNormalPrev, albedoMatIdPrev, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            DIReservoir r;
            if (diTemporal(c, cam, cams[1], g, gp, *set, resvThis, resvPrev, x, y, r)) resvTemp[(size_t)y * W + x] = r;
        }
    });
    sumStats(st, stats3);
}

void orc_di_spatial(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                    const DISettings* set, const DIReservoir* resvTemp, DIReservoir* resvThis, float* directOut,
                    uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            const size_t p = (size_t)y * W + x;
            DIReservoir r; bool written;
            vec3 col = diSpatial(c, cam, g, *set, resvTemp, x, y, r, written);
            if (written) resvThis[p] = r;
            accumulate(directOut + 4 * p, col, cam.frameIndex);
        }
    });
    sumStats(st, stats3);
}

// ---- GRIS passes ----------------------------------------------------------------------------
// Learned-proposal planes for the following GRIS calls with set->guidedProposal = 1 (float4
// planes of W*H pixels as the kernels lay them out: 16 lobe planes for this and the previous
// frame, one record plane written by the path tracer); null unbinds.
void orc_gris_bind_guide(void* h, const float* lobes, const float* lobesPrev, float* samples) {
    OScene* s = (OScene*)h;
    s->guideLobes = lobes; s->guideLobesPrev = lobesPrev; s->guideSamples = samples;
}

// Mixture pdf and samples of one pixel's lobes (lobes: 16 float4 of that pixel, plane-major as
// 16 x 4 floats): pdf at n directions, and n samples from (uLobe, u1, u2) triples.
void orc_guide_mixture(const float* lobes16x4, const float* dirs, uint32_t n, float* pdf, const float* u3, uint32_t m,
                       float* out) {
    GuidePix g = guideAt(lobes16x4, 1, 0);
    for (uint32_t i = 0; i < n; i++) pdf[i] = mixturePdf(g, V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    for (uint32_t i = 0; i < m; i++) {
        vec3 w = mixtureSample(g, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]);
        out[3 * i] = w.x; out[3 * i + 1] = w.y; out[3 * i + 2] = w.z;
    }
}

// The oracle's exp and log on an array (mode 0 exp, 1 log).
void orc_explog(const float* x, uint32_t n, int mode, float* y) {
    for (uint32_t i = 0; i < n; i++) y[i] = mode ? rt_log(x[i]) : rt_exp(x[i]);
}
// cams[0] = uCamera, cams[1] = uPrevCamera
void orc_gris_path_trace(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                         const GRISSettings* set, GRISReservoir* resvThis, uint64_t* stats3, int nthreads) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t], s->guideLobes, s->guideLobesPrev, s->guideSamples};
        for (uint32_t x = 0; x < W; x++) {
            GRISReservoir r; bool written;
            tracePath(c, cam, g, *set, x, y, r, written);
            if (written) resvThis[(size_t)y * W + x] = r;
        }
    });
    sumStats(st, stats3);
}

// learn* (float4 per pixel, may be null): learned-target planes of the three
// buffers, used when set->targetFunction == 1.
void orc_gris_temporal(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                       const uint16_t* motion, const float* depthNormalPrev, const uint32_t* albedoMatIdPrev,
                       const GRISSettings* set, const GRISReservoir* resvThis, const GRISReservoir* resvPrev,
                       GRISReservoir* resvTemp, uint64_t* stats3, int nthreads, const float* learnThis,
                       const float* learnPrev, float* learnTemp) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, motion, W, H};
    GBufView gp{depthNormalPrev, albedoMatIdPrev, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t], s->guideLobes, s->guideLobesPrev, s->guideSamples};
        for (uint32_t x = 0; x < W; x++) {
            GRISReservoir r;
            LearnedData l{};
            if (temporalReuse(c, cam, cams[1], g, gp, *set, resvThis, resvPrev, x, y, r, (const LearnedData*)learnThis,
                              (const LearnedData*)learnPrev, &l)) {
                resvTemp[(size_t)y * W + x] = r;
                if (learnTemp) ((LearnedData*)learnTemp)[(size_t)y * W + x] = l;
            }
        }
    });
    sumStats(st, stats3);
}

void orc_gris_spatial(void* h, const Camera* cams, const float* depthNormal, const uint32_t* albedoMatId,
                      const GRISSettings* set, const GRISReservoir* resvTemp, GRISReservoir* resvThis, float* indirectOut,
                      uint64_t* stats3, int nthreads, const float* learnTemp, float* learnThis) {
    OScene* s = (OScene*)h;
    const Camera& cam = cams[0];
    const uint32_t W = cam.filmSize[0], H = cam.filmSize[1];
    GBufView g{depthNormal, albedoMatId, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t], s->guideLobes, s->guideLobesPrev, s->guideSamples};
        for (uint32_t x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            GRISReservoir r; bool written;
            LearnedData l{};
            vec3 col = spatialReuse(c, cam, g, *set, resvTemp, x, y, r, written, (const LearnedData*)learnTemp, &l);
            if (written) {
                resvThis[p] = r;
                if (learnThis) ((LearnedData*)learnThis)[p] = l;
            }
            accumulate(indirectOut + 4 * p, col, cam.frameIndex);
        }
    });
    sumStats(st, stats3);
}

// Learned-target refresh over a frame, given the network's predictions lhat4
// (float4 per pixel; xyz used). Writes the refreshed reservoirs in place, the
// learned planes to learnOut4 and the training batch (p % stride == offset).
void orc_gris_learned_refresh(void* h, uint32_t W, uint32_t H, const float* depthNormal, GRISReservoir* resv,
                              const float* lhat4, float* learnOut4, float* batchPos, float* batchTarget, uint32_t stride,
                              uint32_t offset, int nthreads) {
    OScene* s = (OScene*)h;
    GBufView g{depthNormal, nullptr, nullptr, W, H};
    std::vector<TraceStats> st(nthreads > 1 ? nthreads : 1);
    parallelRows(H, nthreads, [&](uint32_t y, int t) {
        Ctx c{&s->S, &s->bvh, &st[t]};
        for (uint32_t x = 0; x < W; x++) {
            const size_t p = (size_t)y * W + x;
            LearnedData l{};
            bool need = false;
            vec3 pos = V3s(0.0f);
            if (learnedRefresh(c, g, x, y, resv[p], lhat4 + 4 * p, l, need, pos)) ((LearnedData*)learnOut4)[p] = l;
            if (p % stride == offset) {
                const size_t slot = p / stride;
                const GRISPathSample& ps = resv[p].pathSample;
                batchPos[3 * slot] = pos.x; batchPos[3 * slot + 1] = pos.y; batchPos[3 * slot + 2] = pos.z;
                batchTarget[3 * slot] = need ? ps.rcLi[0] : -1.0f;
                batchTarget[3 * slot + 1] = need ? ps.rcLi[1] : 0.0f;
                batchTarget[3 * slot + 2] = need ? ps.rcLi[2] : 0.0f;
            }
        }
    });
}

// ---- component entry points for pinning tests ------------------------------------------------
uint32_t orc_hash2(uint32_t s) { return hash2(s); }
uint32_t orc_make_seed(uint32_t seed, uint32_t x, uint32_t y) { return makeSeedPix(seed, x, y); }
// texture(uTextures[i], uv) for n coordinates (for the sampler's pins)
void orc_sample_texture(const Texture* tex, const float* uv2, uint32_t n, float* rgb3) {
    for (uint32_t i = 0; i < n; i++) {
        vec3 c = sampleTexture(*tex, uv2[2 * i], uv2[2 * i + 1]);
        rgb3[3 * i] = c.x; rgb3[3 * i + 1] = c.y; rgb3[3 * i + 2] = c.z;
    }
}
void orc_srgb_table(float* out256) { std::memcpy(out256, srgbTable(), 256 * sizeof(float)); }
// n draws of sample1f from state rng
void orc_sample1f(uint32_t rng, uint32_t n, float* out) { for (uint32_t i = 0; i < n; i++) out[i] = sample1f(rng); }
void orc_sincos(const float* x, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; i++) { s[i] = rt_sin(x[i]); c[i] = rt_cos(x[i]); }
}
void orc_concentric(const float* uv, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) { vec2 r = toConcentricDisk(V2(uv[2 * i], uv[2 * i + 1])); out[2 * i] = r.x; out[2 * i + 1] = r.y; }
}
void orc_camera_rays(const Camera* cam, float* dirs3) {
    const uint32_t W = cam->filmSize[0], H = cam->filmSize[1];
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            vec2 uv = V2(((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H);
            Ray r = pinholeCameraSampleRay(*cam, V2(uv.x, 1.0f - uv.y), V2(0, 0));
            float* d = dirs3 + 3 * ((size_t)y * W + x);
            d[0] = r.dir.x; d[1] = r.dir.y; d[2] = r.dir.z;
        }
}
// BSDF evaluation: per query n(3) wo(3) wi(3) -> f(3), pdf
void orc_eval_bsdf(const Material* mat, const float* albedo3, const float* q9, uint32_t n, float* f3, float* pdf) {
    vec3 a = L3(albedo3);
    for (uint32_t i = 0; i < n; i++) {
        const float* q = q9 + 9 * i;
        vec3 nn = V3(q[0], q[1], q[2]), wo = V3(q[3], q[4], q[5]), wi = V3(q[6], q[7], q[8]);
        vec3 f = evalBSDF(*mat, a, nn, wo, wi);
        f3[3 * i] = f.x; f3[3 * i + 1] = f.y; f3[3 * i + 2] = f.z;
        pdf[i] = evalPdf(*mat, nn, wo, wi);
    }
}
// BSDF sampling: per query n(3) wo(3) r(3) -> wi(3), bsdf(3), pdf, type, ok
void orc_sample_bsdf(const Material* mat, const float* albedo3, const float* q9, uint32_t n, float* wi3, float* f3,
                     float* pdf, uint32_t* type, uint32_t* ok) {
    vec3 a = L3(albedo3);
    for (uint32_t i = 0; i < n; i++) {
        const float* q = q9 + 9 * i;
        BSDFSample s{};
        ok[i] = sampleBSDF(*mat, a, V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]), s) ? 1u : 0u;
        wi3[3 * i] = s.wi.x; wi3[3 * i + 1] = s.wi.y; wi3[3 * i + 2] = s.wi.z;
        f3[3 * i] = s.bsdf.x; f3[3 * i + 1] = s.bsdf.y; f3[3 * i + 2] = s.bsdf.z;
        pdf[i] = s.pdf; type[i] = s.type;
    }
}
// GGX visible-normal sampler (GTR2Sample) alone: n(3) wo(3) r(2), alpha -> wh(3)
void orc_gtr2_sample(const float* q8, uint32_t n, float alpha, float* wh3) {
    for (uint32_t i = 0; i < n; i++) {
        const float* q = q8 + 8 * i;
        vec3 h = GTR2Sample(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), alpha, V2(q[6], q[7]));
        wh3[3 * i] = h.x; wh3[3 * i + 1] = h.y; wh3[3 * i + 2] = h.z;
    }
}
// Light sampling by power: ref(3) r(4) -> wi(3) dist pdf jacobian bary(2) id radiance(3)
void orc_sample_light(const SceneArrays* S, const float* q7, uint32_t n, float* out11, uint32_t* id) {
    for (uint32_t i = 0; i < n; i++) {
        const float* q = q7 + 7 * i;
        vec3 wi; float dist, pdf, jac; vec2 bary; uint32_t lid;
        vec3 L = sampleLightByPower(*S, V3(q[0], q[1], q[2]), wi, dist, pdf, jac, bary, lid, {q[3], q[4], q[5], q[6]});
        float* o = out11 + 11 * i;
        o[0] = wi.x; o[1] = wi.y; o[2] = wi.z; o[3] = dist; o[4] = pdf; o[5] = jac; o[6] = bary.x; o[7] = bary.y;
        o[8] = L.x; o[9] = L.y; o[10] = L.z;
        id[i] = lid;
    }
}

// src/util/AliasTable.h:26-71 (DiscreteSampler1D<float>::build), output N+1 entries.
void orc_alias_build(const float* distrib, uint32_t n, LightSampleTableElement* out) {
    std::vector<float> d(distrib, distrib + n);
    float sumAll = 0.0f;
    for (float v : d) sumAll += v;
    float sumInv = (float)n / sumAll;
    for (float& v : d) v *= sumInv;
    std::vector<LightSampleTableElement> gt(n * 2 + 1), ls(n * 2 + 1);
    int topG = 0, topL = 0;
    for (uint32_t i = 0; i < n; i++) {
        float v = d[i];
        if (v > 1.0f) gt[topG++] = {v, i + 1}; else ls[topL++] = {v, i + 1};
    }
    while (topG && topL) {
        LightSampleTableElement g = gt[--topG];
        LightSampleTableElement l = ls[--topL];
        out[l.failId] = {l.prob, g.failId};
        g.prob -= (1.0f - l.prob);
        if (g.prob > 1.0f) gt[topG++] = g; else ls[topL++] = g;
    }
    for (int i = topG - 1; i >= 0; i--) out[gt[i].failId] = gt[i];
    for (int i = topL - 1; i >= 0; i--) out[ls[i].failId] = ls[i];
    out[0] = {sumAll, n};
}

uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }

}  // extern "C"
