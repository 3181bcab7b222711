// noalloc.hip -- the charge deposit for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// No reference counterpart.  aw_deposit_rho accumulates into the context's
// int64 grid and, on its first call, reserves the FP32 staging copy the solve
// reads; aw_deposit_rho_noalloc runs the same deposit kernels into a
// file-scope device array instead and needs no context at all.  The array
// holds at most kScratchCells cells (8 MiB; 1024 x 1024 grids and smaller);
// each GPU keeps its own copy until the process ends.  Its address is taken
// once per device with hipGetSymbolAddress: in host code the array's name is
// not a device address.  Every call on one GPU shares that GPU's array, so
// calls that overlap on one GPU (different streams, several host threads)
// race; calls on one stream are ordered and safe.
#include "aw_internal.hpp"

#include <mutex>

namespace aw {
namespace {

constexpr int64_t kScratchCells = 1 << 20;
constexpr int kMaxDevices = 64;
constexpr int kThreads = 256;

static __device__ __align__(256) unsigned long long scratch[kScratchCells];

struct DeviceScratch {
    unsigned long long* ptr = nullptr;  // device address of this device's scratch
    int num_cu = 0;
    size_t lds_per_block = 0;
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// The current device's scratch address and attributes, looked up on the first
// call on that device (with it current) and kept for the process.
const DeviceScratch& device_scratch() {
    int dev = 0;
    AW_HIP(hipGetDevice(&dev));
    AW_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        AW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        AW_REQUIRE(p != nullptr, "hipGetSymbolAddress returned NULL for the scratch array");
        int cu = 0, lds = 0;
        AW_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        AW_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        d.num_cu = cu;
        d.lds_per_block = (size_t)lds;
        d.ptr = static_cast<unsigned long long*>(p);
    }
    return d;
}

// acc[0..n) = 0 (a kernel, not hipMemsetAsync, so the call can be captured)
__global__ void zero_acc_kernel(unsigned long long* __restrict__ acc, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = 0ull;
}

}  // namespace
}  // namespace aw

using namespace aw;

extern "C" {

aw_status aw_deposit_rho_noalloc(const aw_grid* g, const aw_particles* p, double qm,
                                 int32_t policy, int32_t scale_log2, float* rho, aw_stream s) {
    return guard("aw_deposit_rho_noalloc", [&] {
        check_grid(g);
        check_particles(p);
        AW_REQUIRE(rho, "rho is NULL");
        const int64_t nc = (int64_t)g->nx * g->ny;
        if (nc > kScratchCells)
            fail(AW_ERR_INVALID,
                 "grid of " + std::to_string(nc) + " cells exceeds the static scratch of " +
                     std::to_string(kScratchCells) + " cells (aw_deposit_rho has no such limit)",
                 __FILE__, __LINE__);
        AW_REQUIRE(policy == AW_DEPOSIT_GLOBAL || policy == AW_DEPOSIT_LDS_GRID ||
                       policy == AW_DEPOSIT_AUTO,
                   "aw_deposit_rho_noalloc takes AW_DEPOSIT_GLOBAL, AW_DEPOSIT_LDS_GRID or "
                   "AW_DEPOSIT_AUTO (binning needs scratch beyond the static array)");
        AW_REQUIRE(scale_log2 >= 0 && scale_log2 <= 62, "scale_log2 must be in [0, 62]");
        const DeviceScratch& d = device_scratch();
        const hipStream_t st = S(s);
        zero_acc_kernel<<<(int)((nc + kThreads - 1) / kThreads), kThreads, 0, st>>>(d.ptr, nc);
        AW_LAUNCH_CHECK();
        deposit_rho_acc(*g, *p, qm, policy, scale_log2, d.ptr, d.num_cu, d.lds_per_block, st);
        fix_to_float_add(d.ptr, rho, nc, scale_log2, false, st, false);
    });
}

}  // extern "C"
