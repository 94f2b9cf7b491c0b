// Host scene: light triangles, light table and device upload.
#include "Scene.h"
#include <cmath>
#include <cstring>

// src/util/AliasTable.h:26-71
void DiscreteSampler1D::build(std::vector<float> distribution) {
    const uint32_t n = (uint32_t)distribution.size();
    binomDistribs.assign(n + 1, rt_light_table_entry{0.0f, 0u});
    float sumAll = 0.0f;
    for (float v : distribution) sumAll += v;
    float sumInv = (float)n / sumAll;
    for (float& v : distribution) v *= sumInv;
    std::vector<rt_light_table_entry> gtOne(2 * (size_t)n + 1), lsOne(2 * (size_t)n + 1);
    int topG = 0, topL = 0;
    for (uint32_t i = 0; i < n; i++) {
        rt_light_table_entry e{distribution[i], i + 1};
        if (e.prob > 1.0f) gtOne[topG++] = e; else lsOne[topL++] = e;
    }
    while (topG && topL) {
        rt_light_table_entry gt = gtOne[--topG];
        rt_light_table_entry ls = lsOne[--topL];
        binomDistribs[ls.failId] = rt_light_table_entry{ls.prob, gt.failId};
        gt.prob -= (1.0f - ls.prob);
        if (gt.prob > 1.0f) gtOne[topG++] = gt; else lsOne[topL++] = gt;
    }
    for (int i = topG - 1; i >= 0; i--) binomDistribs[gtOne[i].failId] = gtOne[i];
    for (int i = topL - 1; i >= 0; i--) binomDistribs[lsOne[i].failId] = lsOne[i];
    binomDistribs[0] = rt_light_table_entry{sumAll, n};
}

uint32_t DiscreteSampler1D::sample(float r1, float r2) const {
    uint32_t passId = (uint32_t)((float)binomDistribs.size() * r1);
    const rt_light_table_entry& d = binomDistribs[passId];
    return (r2 < d.prob) ? passId : d.failId;
}

void Scene::addLightTriangle(const float v0[3], const float v1[3], const float v2[3], const float radiance[3]) {
    rt_triangle_light t{};
    for (int a = 0; a < 3; a++) { t.v0[a] = v0[a]; t.v1[a] = v1[a]; t.v2[a] = v2[a]; t.radiance[a] = radiance[a]; }
    float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    t.area = 0.5f * len;
    t.nx = n[0] / len; t.ny = n[1] / len; t.nz = n[2] / len;
    triangleLights.push_back(t);
}

void Scene::buildLightDataStructure() {
    std::vector<float> power(triangleLights.size());
    for (size_t i = 0; i < triangleLights.size(); i++) {
        const rt_triangle_light& t = triangleLights[i];
        power[i] = t.radiance[0] * t.area * 0.299f + t.radiance[1] * t.area * 0.587f + t.radiance[2] * t.area * 0.114f;
    }
    lightSampleTable.build(power);
}

rt_scene_arrays Scene::arrays() const {
    rt_scene_arrays a{};
    a.materials = materials.data(); a.numMaterials = (uint32_t)materials.size();
    a.materialIndices = materialIndices.data(); a.numMaterialIndices = (uint32_t)materialIndices.size();
    a.vertices = vertices.data(); a.numVertices = (uint32_t)vertices.size();
    a.indices = indices.data(); a.numIndices = (uint32_t)indices.size();
    a.instances = objectInstances.data(); a.numInstances = (uint32_t)objectInstances.size();
    a.lights = triangleLights.data(); a.numLights = (uint32_t)triangleLights.size();
    a.lightTable = lightSampleTable.binomDistribs.data();
    a.numLightTable = (uint32_t)lightSampleTable.binomDistribs.size();
    textureViews.clear();
    for (const Texture& t : textures) {
        rt_texture v{};
        v.width = t.width; v.height = t.height; v.filter = t.filter; v.rgba8 = t.rgba.data();
        textureViews.push_back(v);
    }
    a.textures = textureViews.data();
    a.numTextures = (uint32_t)textureViews.size();
    return a;
}

// General 4x4 inverse (cofactors), as glm::inverse.
static void inverse4(const float* m, float* inv) {
    float t[16];
    t[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    t[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    t[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    t[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    t[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    t[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    t[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    t[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    t[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    t[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    t[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    t[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    t[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    t[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    t[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    t[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * t[0] + m[1] * t[4] + m[2] * t[8] + m[3] * t[12];
    float id = 1.0f / det;
    for (int i = 0; i < 16; i++) inv[i] = t[i] * id;
}

void makeInstanceTransform(rt_object_instance& inst, const float m[16]) {
    std::memcpy(inst.transform, m, 64);
    inverse4(m, inst.transformInv);
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) inst.transformInvT[c * 4 + r] = inst.transformInv[r * 4 + c];
}

DeviceScene::DeviceScene(const Scene& scene, hipStream_t stream) {
    rt_scene_arrays a = scene.arrays();
    rtThrow(rt_device_scene_create(&a, stream, &mScene), "rt_device_scene_create");
}
DeviceScene::~DeviceScene() {
    if (mScene && mOwned) rt_device_scene_destroy(mScene);
}
rt_scene_info DeviceScene::info() const {
    rt_scene_info i{};
    rtThrow(rt_device_scene_info(mScene, &i), "rt_device_scene_info");
    return i;
}
