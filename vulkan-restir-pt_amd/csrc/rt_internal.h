// Host-side internals shared by the C ABI implementation and the kernel
// launchers (not part of the public interface).
#pragma once
#include "rt_dtrace.h"
#include <atomic>
#include <string>
#include <vector>

struct rt_device_scene {
    rtd::DevScene dev;            // device view passed to kernels by value
    // owned device allocations
    float4* nodes = nullptr;      // binary LBVH: 4 x float4 per internal node
    float4* nodes4 = nullptr;     // collapsed 4-wide tree: 8 x float4 per node (traversed)
    uint32_t numNodes4 = 0, bvh4Depth = 0;
    float4* leafTris = nullptr;   // 3 x float4 per leaf
    float4* primTris = nullptr;   // 3 x float4 per primitive (primitive order)
    unsigned long long* codes = nullptr;  // sorted Morton codes
    uint32_t* order = nullptr;    // leaf -> primitive
    rt_material* materials = nullptr;
    int32_t* materialIndices = nullptr;
    rt_mesh_vertex* vertices = nullptr;
    uint32_t* indices = nullptr;
    rtd::DevInstance* instances = nullptr;
    rt_triangle_light* lights = nullptr;
    rt_light_table_entry* lightTable = nullptr;
    uint32_t numPrims = 0, numLights = 0, numInstances = 0;
    uint32_t bvhDepth = 0;
    float sceneLo[3] = {0, 0, 0}, sceneHi[3] = {0, 0, 0};
    size_t deviceBytes = 0;
    // Pixel queues of the persistent passes, handed out round robin: up to
    // kQueueSlots such launches on this scene may be in flight at once.
    static constexpr uint32_t kQueueSlots = 64;
    rtd::WorkQueue* queues = nullptr;
    mutable std::atomic<uint32_t> queueTick{0};
    rtd::WorkQueue* nextQueue() const { return queues + (queueTick.fetch_add(1u) % kQueueSlots); }
};

namespace rti {

void setError(const std::string& msg);
bool hipCheck(hipError_t e, const char* what);

// LBVH build over dev scene arrays already uploaded; fills nodes/leafTris/codes/order.
bool buildLBVH(rt_device_scene* s, const std::vector<uint32_t>& instPrimOffset, hipStream_t stream);

// Counters of the counting kernel builds.
extern bool gCountersEnabled;
rtd::TraceCounters* counterBuffer();

}  // namespace rti

#define RT_HIP(call) do { if (!rti::hipCheck((call), #call)) return RT_ERR_HIP; } while (0)
