// Exact direct-sum gravity on caller device arrays, allocating nothing: the no-allocation variant of
// isph_gravity_accelerations(mode ISPH_GRAVITY_DIRECT), for callers that cannot allocate device memory
// (work under HIP graph capture, tight serving loops).
//
// The f32 bodies the sum reads and its f32 result live in `scratch`, a file-scope device array of fixed
// size that every GPU the library runs on holds until the process ends. Host code never uses the array's
// name as an address: on a device's first call the entry point asks the runtime for that device's copy
// (hipGetSymbolAddress, with the device current) and keeps the address per device. Every call on one GPU
// shares that GPU's array, so calls that overlap on one GPU (two streams, two threads) race.
#include <mutex>

#include "../../include/isph.h"
#include "state.h"

constexpr uint32_t kNoAllocMax = 1u << 20;  // bodies per call: 2 x 16 MiB of scratch

static __device__ __align__(256) float4 scratch[2 * kNoAllocMax];  // [bodies | result]

namespace isph {
namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_lock;
float4* g_scratch_ptr[kMaxDevices] = {};

// gpu_gravity.rs:340-345: the bodies as f32 (pos.x, pos.y, pos.z, mass).
__global__ void k_stage_bodies(const double* __restrict__ pos3, const double* __restrict__ mass, uint32_t n,
                               float4* __restrict__ bodies) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bodies[i] = make_float4((float)pos3[3 * i], (float)pos3[3 * i + 1], (float)pos3[3 * i + 2], (float)mass[i]);
}

__global__ void k_widen_acc(const float4* __restrict__ g, uint32_t n, double* __restrict__ acc3) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 a = g[i];
        acc3[3 * i] = (double)a.x;
        acc3[3 * i + 1] = (double)a.y;
        acc3[3 * i + 2] = (double)a.z;
    }
}

int fail_thread(int code, const std::string& msg) {
    set_thread_error(msg);
    return code;
}

}  // namespace
}  // namespace isph

using namespace isph;

extern "C" int isph_gravity_direct_noalloc(const double* pos3, const double* mass, size_t n, double softening,
                                           double* acc3, void* stream) {
    if (n == 0) return ISPH_OK;
    if (!pos3 || !mass || !acc3) return fail_thread(ISPH_EINVAL, "isph_gravity_direct_noalloc: NULL array");
    if (n > kNoAllocMax)
        return fail_thread(ISPH_ERANGE, "isph_gravity_direct_noalloc: n = " + std::to_string(n) +
                                            " exceeds the scratch capacity of " + std::to_string(kNoAllocMax) +
                                            " bodies");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_thread(ISPH_EDEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail_thread(ISPH_ENODEV, "isph_gravity_direct_noalloc: device ordinal");
    float4* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_lock);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail_thread(ISPH_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            g_scratch_ptr[dev] = static_cast<float4*>(p);
        }
        base = g_scratch_ptr[dev];
    }
    float4* bodies = base;
    float4* result = base + kNoAllocMax;
    const uint32_t m = (uint32_t)n;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t blocks = (m + 255) / 256;
    hipLaunchKernelGGL(k_stage_bodies, dim3(blocks), dim3(256), 0, s, pos3, mass, m, bodies);
    // soft2 = (s*s) as f32, as the allocating entry point forms it (gpu_gravity.rs:340-345)
    launch_gravity_direct_raw(bodies, m, (float)(softening * softening), result, true, s);
    hipLaunchKernelGGL(k_widen_acc, dim3(blocks), dim3(256), 0, s, result, m, acc3);
    e = hipGetLastError();
    if (e != hipSuccess) return fail_thread(ISPH_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return ISPH_OK;
}
