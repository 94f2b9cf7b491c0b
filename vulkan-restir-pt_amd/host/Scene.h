// Host scene and device scene of the C++ layer.
//
// Scene keeps the reference's public members (src/Scene.h:40-66): camera,
// object instances, triangle lights and the light sample table, plus the flat
// resource arrays the reference keeps in Resource (vertices, indices,
// materials, per-triangle material indices). DeviceScene is the RAII owner of
// an rt_device_scene (src/Scene.h:68-105).
#pragma once
#include "Camera.h"
#include "restirpt.h"
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

// DiscreteSampler1D<float> (src/util/AliasTable.h:22-90): entry 0 holds
// {sum, count}; entries 1..N are 1-based binomial buckets.
struct DiscreteSampler1D {
    std::vector<rt_light_table_entry> binomDistribs;
    void build(std::vector<float> distribution);
    uint32_t sample(float r1, float r2) const;
};

// An albedo texture: RGBA8 texels, row 0 first; filter 0 nearest, 1 linear.
struct Texture {
    uint32_t width = 0, height = 0, filter = 1;
    std::vector<uint8_t> rgba;
};
// Reads an 8-bit PNG file (grey, grey + alpha, RGB, RGBA or palette, not interlaced) as RGBA8.
Texture readPng(const std::string& path);

class Scene {
public:
    // Adds a light triangle with world-space vertices and per-triangle radiance;
    // area and normal are derived as in src/Scene.cpp:242-252.
    void addLightTriangle(const float v0[3], const float v1[3], const float v2[3], const float radiance[3]);
    // src/Scene.cpp:309-322: power = luminance(radiance * area)
    void buildLightDataStructure();
    rt_scene_arrays arrays() const;

    Camera camera;
    std::vector<rt_mesh_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<rt_material> materials;
    std::vector<int32_t> materialIndices;
    std::vector<rt_object_instance> objectInstances;
    std::vector<rt_triangle_light> triangleLights;
    DiscreteSampler1D lightSampleTable;
    std::vector<Texture> textures;              // Resource::imagePool (src/Resource.h)
    mutable std::vector<rt_texture> textureViews;  // what arrays() hands out

    // camera path of the procedural workloads: per frame position, angle, seed
    struct PathKey { Vec3 pos; Vec3 angle; uint32_t seed; };
    std::vector<PathKey> cameraPath;
    float fov = 60.0f;
    float lensRadius = 0.0f, focalDist = 1.0f;
    uint32_t filmSize[2] = {0, 0};  // the scene file's film size; 0 for procedural scenes
};

// Fills transform / transformInv / transformInvT from a column-major affine matrix.
void makeInstanceTransform(rt_object_instance& inst, const float m[16]);

class DeviceScene {
public:
    DeviceScene(const Scene& scene, hipStream_t stream);
    // Wraps a scene created through the C ABI without taking ownership.
    explicit DeviceScene(rt_device_scene* borrowed) : mScene(borrowed), mOwned(false) {}
    ~DeviceScene();
    DeviceScene(const DeviceScene&) = delete;
    DeviceScene& operator=(const DeviceScene&) = delete;
    rt_device_scene* handle() const { return mScene; }
    rt_scene_info info() const;

private:
    rt_device_scene* mScene = nullptr;
    bool mOwned = true;
};

std::unique_ptr<Scene> generateScene(uint32_t config, uint32_t seed);
// Scene::load (src/Scene.cpp:104-303): a scene file in the reference's XML format.
std::unique_ptr<Scene> loadSceneFile(const std::string& path);

inline void rtThrow(rt_status st, const char* what) {
    if (st != RT_OK) throw std::runtime_error(std::string(what) + ": " + rt_last_error());
}
