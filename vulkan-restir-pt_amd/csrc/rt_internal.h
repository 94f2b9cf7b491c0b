// Host-side internals shared by the C ABI implementation and the kernel
// launchers (not part of the public interface).
#pragma once
#include "rt_dtrace.h"
#include "rt_mlp.h"
#include <atomic>
#include <memory>
#include <string>
#include <tuple>
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
    uint4* texDesc = nullptr;     // per texture {first texel, width, height, filter}
    uint32_t* texels = nullptr;   // RGBA8 texels of all textures
    float* srgbLut = nullptr;     // sRGB byte -> linear
    uint32_t numPrims = 0, numLights = 0, numInstances = 0;
    uint32_t bvhDepth = 0;
    float bvhBinaryMs = 0.0f, bvhBuildMs = 0.0f;  // GPU time of the binary build / of the whole build
    float sceneLo[3] = {0, 0, 0}, sceneHi[3] = {0, 0, 0};
    size_t deviceBytes = 0;
    // Pixel queues of the eager persistent path traces, handed out round robin.
    // Each slot's event marks the end of its last launch; a launch that takes a
    // slot whose previous launch may still run waits for that event on its own
    // stream, so launches in flight on other streams never share a queue.
    static constexpr uint32_t kQueueSlots = 64;
    rtd::WorkQueue* queues = nullptr;
    hipEvent_t queueEvents[kQueueSlots] = {};
    mutable std::atomic<uint32_t> queueTick{0};
    uint32_t nextQueueSlot() const { return queueTick.fetch_add(1u) % kQueueSlots; }
};

// Learned target function: parameters, optimiser state and the frame's
// training batch, all on the device.
struct rt_mlp {
    rt_mlp_config cfg{};
    // 3: the radiance network of the learned target; 32: the vMF-mixture proposal network
    uint32_t outputs = rtm::kOut;
    uint32_t weights = rtm::kWeights;   // FP16/FP32 weight count for `outputs`
    uint32_t targetStride = 3;          // floats per training target (RGB, or direction + weight)
    float* wmean = nullptr;             // proposal network: mean sample weight of the last batch
    rtm::MlpView view{};          // kernels' view of the FP16 copies
    uint64_t gridParams = 0;      // levels * T * 2
    float* gridF32 = nullptr;     // FP32 master grid
    _Float16* gridF16 = nullptr;  // FP16 copy read by the kernels
    long long* gridGrad = nullptr;  // summed per-sample gradients, 32.32 fixed point (order-independent sums)
    float* gridM = nullptr;
    float* gridV = nullptr;
    float* wF32 = nullptr;
    _Float16* wF16 = nullptr;
    float* wGrad = nullptr;       // reduced weight gradients (unscaled), then loss sum, sample count
    float* wM = nullptr;
    float* wV = nullptr;
    float* partial = nullptr;     // per-wave weight-gradient partial sums
    uint32_t partialWaves = 0;
    float* batchPos = nullptr;    // training batch gathered from the frame
    float* batchTarget = nullptr;
    uint32_t batchSize = 0;       // samples gathered by the last refresh
    uint32_t step = 0;            // Adam steps taken
};

namespace rti {

void setError(const std::string& msg);
bool hipCheck(hipError_t e, const char* what);

// LBVH build over dev scene arrays already uploaded; fills nodes/leafTris/codes/order.
bool buildLBVH(rt_device_scene* s, const std::vector<uint32_t>& instPrimOffset, hipStream_t stream);

// Scratch of one binary build, carved from one buffer (256-B aligned pieces).
struct LbvhScratch {
    unsigned long long* keysB = nullptr;
    uint32_t* valsB = nullptr;
    uint32_t* blockHist = nullptr;   // 256 x tiles
    uint32_t* digitTotal = nullptr;  // 256
    float* partial = nullptr;        // 6 x bounds blocks
    float* box = nullptr;            // 6
    uint32_t* parentI = nullptr;     // internal nodes
    uint32_t* parentL = nullptr;     // leaves
    uint2* range = nullptr;          // internal nodes
    uint32_t* height = nullptr;      // internal nodes
    uint32_t* topList = nullptr;     // internal nodes
};

// Bytes of the scratch of an n-primitive binary build; with base non-null also
// carves it into *out.
size_t carveLbvhScratch(uint32_t n, char* base, LbvhScratch* out);
// Binary LBVH of n world-space triangles (3 float4 per primitive) into codes,
// order (n) and nodes (max(n - 1, 1) x 4 float4): kernel launches only, no
// allocation and no host synchronisation. With leafTris non-null the
// triangles are also gathered into Morton order. The centroid box lands in
// sc.box, the tree's height in sc.height[0].
hipError_t launchBinaryBuild(const float4* primTris, uint32_t n, unsigned long long* codes, uint32_t* order,
                             float4* nodes, float4* leafTris, const LbvhScratch& sc, hipStream_t stream);

// Counters of the counting kernel builds.
extern bool gCountersEnabled;
rtd::TraceCounters* counterBuffer();

// A kernel launch with its arguments held by value: launched at once, or
// recorded while a recorder is installed on this thread (frame graphs build
// and update their kernel nodes from recorded launches).
struct KernelLaunch {
    const void* func = nullptr;
    dim3 grid, block;
    std::shared_ptr<void> store;   // owns the argument values
    std::vector<void*> args;       // pointers into store, in parameter order
};
extern thread_local std::vector<KernelLaunch>* tlsRecorder;

template <class... P, class... A>
hipError_t launch(void (*kernel)(P...), dim3 grid, dim3 block, hipStream_t stream, A&&... a) {
    using Args = std::tuple<std::decay_t<P>...>;
    if (tlsRecorder) {
        auto store = std::make_shared<Args>(std::forward<A>(a)...);
        KernelLaunch k;
        k.func = reinterpret_cast<const void*>(kernel);
        k.grid = grid;
        k.block = block;
        std::apply([&](auto&... x) { (k.args.push_back(static_cast<void*>(&x)), ...); }, *store);
        k.store = store;
        tlsRecorder->push_back(std::move(k));
        return hipSuccess;
    }
    Args values(std::forward<A>(a)...);
    void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
    size_t i = 0;
    std::apply([&](auto&... x) { ((ptrs[i++] = static_cast<void*>(&x)), ...); }, values);
    return hipLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, ptrs, 0, stream);
}

// A chain of kernel nodes replayed as one HIP graph. update() takes a frame's
// recorded launches: when the kernels and launch shapes are those of the
// instantiated graph it writes the new arguments into its nodes
// (hipGraphExecKernelNodeSetParams), otherwise it builds and instantiates the
// graph anew. The graph holds kernel nodes only (no memset or copy nodes).
struct KernelChain {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> nodes;
    std::vector<const void*> funcs;
    std::vector<dim3> grids, blocks;
    hipError_t update(std::vector<KernelLaunch>& launches);
    hipError_t launch(hipStream_t stream) const;
    void reset();
    ~KernelChain() { reset(); }
};

// Path state of the wavefront (material-sorted) variant of the batched path
// trace, for up to `capacity` paths (one per pixel of the batch).
struct WavefrontBuffers {
    float* words = nullptr;       // per-path state planes
    uint32_t* lists = nullptr;    // 2 x capacity live-path lists
    uint32_t* sorted = nullptr;   // capacity: a bounce's live paths in bucket order
    uint32_t* keys = nullptr;     // capacity: buckets
    uint32_t* counters = nullptr; // list lengths, bucket counts and cursors
    uint32_t capacity = 0;
};

// Register budgets (waves per SIMD) of the batched path-trace variants that a
// batch's tuning chooses from.
constexpr uint32_t kBatchWaveVariants = 3;
constexpr int kBatchWaves[kBatchWaveVariants] = {3, 4, 5};

}  // namespace rti

#define RT_HIP(call) do { if (!rti::hipCheck((call), #call)) return RT_ERR_HIP; } while (0)
