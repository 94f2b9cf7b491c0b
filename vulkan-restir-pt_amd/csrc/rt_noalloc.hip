// Allocation-free LBVH build: the binary build of rt_device_scene_create
// (Morton codes, radix sort, Karras topology, refit) over caller-owned device
// buffers, with its scratch in a file-scope device array instead of hipMalloc.
// For callers that cannot allocate (code under HIP graph capture, serving
// loops that rebuild dynamic geometry every frame).
//
// The array is linked into the library's code object, so every GPU that the
// library runs on holds its own copy for the life of the process. Host code
// cannot pass the array's name to a kernel (it is a host-side stand-in, not a
// device address): its address on each device is looked up once with
// hipGetSymbolAddress, with that device current, and that pointer is used
// everywhere. All calls on one device share the array: calls that overlap on
// one GPU (different streams, or threads without ordering) race.
#include "rt_internal.h"
#include <mutex>

namespace {

// The scratch of a build of kMaxPrims primitives fits (checked per call).
constexpr uint32_t kMaxPrims = 1u << 17;
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex gMutex;
void* gScratch[kMaxDevices] = {};

// The device address of this device's copy of `scratch`.
rt_status scratchOnCurrentDevice(void** out) {
    int dev = 0;
    RT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        rti::setError("device ordinal beyond the supported range");
        return RT_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(gMutex);
    if (!gScratch[dev]) {
        void* p = nullptr;
        RT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        gScratch[dev] = p;
    }
    *out = gScratch[dev];
    return RT_OK;
}

}  // namespace

extern "C" {

rt_status rt_bvh_build_noalloc(const float* primTris, uint32_t n, uint64_t* codes, uint32_t* order, float* nodes,
                               hipStream_t stream) {
    if (!primTris || !codes || !order || !nodes) {
        rti::setError("null argument");
        return RT_ERR_INVALID_ARGUMENT;
    }
    if (n == 0 || n > kMaxPrims) {
        rti::setError("rt_bvh_build_noalloc builds 1 to " + std::to_string(kMaxPrims) + " primitives, got " +
                      std::to_string(n));
        return RT_ERR_CAPACITY;
    }
    const size_t need = rti::carveLbvhScratch(n, nullptr, nullptr);
    if (need > kScratchBytes) {
        rti::setError("build scratch exceeds the static array");
        return RT_ERR_CAPACITY;
    }
    void* base = nullptr;
    rt_status st = scratchOnCurrentDevice(&base);
    if (st != RT_OK) return st;
    rti::LbvhScratch sc;
    rti::carveLbvhScratch(n, static_cast<char*>(base), &sc);
    RT_HIP(rti::launchBinaryBuild(reinterpret_cast<const float4*>(primTris), n,
                                  reinterpret_cast<unsigned long long*>(codes), order, reinterpret_cast<float4*>(nodes),
                                  nullptr, sc, stream));
    return RT_OK;
}

}  // extern "C"
