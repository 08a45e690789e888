// The database pass that allocates nothing (ish_hip_align_db_noalloc), for
// callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// Its kernels' scratch (the work-queue counter and the per-wave boundary
// rows between query chunks, which ish_hip_align_db keeps in context
// buffers it grows on demand) lives in the file-scope device array below,
// held by each device the library runs on until the process ends. Host code
// never passes the array's name to the device: its device address is looked
// up once per device, when a context is created on that device with the
// device current, and that address is what the kernels receive.
//
// Every call on one device shares that device's array and leaves the last
// call's contents in it, so calls that overlap on one device race (two
// contexts of one device used from two threads at once). Nothing in it has
// to start at zero: each launch resets its counter with a kernel first, and
// every boundary row is written before it is read within a launch.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ish_hip.h"
#include "host.h"
#include "pieces.h"

// 512 MiB: the boundary rows of 2,048 resident waves (full occupancy) for
// targets up to 512 residues; longer targets run on fewer workgroups, and a
// target whose single workgroup's rows do not fit (over 262,143 residues) is
// rejected with ISH_E_RANGE.
constexpr size_t kNoallocScratchBytes = (size_t)512 << 20;
static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};
} // namespace

hipError_t ish::fixed_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

extern "C" int32_t ish_hip_align_db_noalloc(ish_hip_ctx *ctx, ish_hip_db *db, int32_t *d_score,
                                            int32_t *d_end_ref, int32_t *d_end_query,
                                            uint8_t *d_status) {
    const int32_t dev = ish::ctx_device(ctx);
    if (dev < 0 || dev >= kMaxDevices) return ISH_E_INVALID;
    void *ptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        ptr = g_scratch_ptr[dev];
    }
    return ish::align_db_fixed(ctx, db, ptr, kNoallocScratchBytes, d_score, d_end_ref, d_end_query,
                               d_status);
}
