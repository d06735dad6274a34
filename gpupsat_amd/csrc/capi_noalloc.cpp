// capi_noalloc.cpp -- gpsat_solve_noalloc: gpsat_solve with every device buffer carved
// from a static device array instead of allocated, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"
#include "solve_core.hpp"
#include "solver.hpp"

using namespace gpsat;

namespace {

// Device memory of one call carved from a fixed region (the static scratch array of
// gpsat_solve_noalloc): a bump pointer, 256-byte aligned pieces, nothing allocated.
struct ScratchMem {
    char* base = nullptr;
    size_t size = 0, used = 0;
    template <class T>
    hipError_t alloc(T*& p, size_t count)
    {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (bytes > size - used) {
            p = nullptr;
            return hipErrorOutOfMemory;
        }
        p = (T*)(base + used);
        used += bytes;
        return hipSuccess;
    }
    void mark_synced() {}
    static constexpr bool kNoAlloc = true;
    size_t remaining() const { return size - used; }
};

// ---- the allocation-free variant's device memory.  scratch is the region every
// gpsat_solve_noalloc call on a GPU carves its buffers from; each GPU holds its own copy.
// Host code never passes the array's name to the device: its per-device address comes
// from hipGetSymbolAddress once, at the first call on that device.
constexpr size_t kScratchMiB = GPSAT_NOALLOC_SCRATCH_MIB;
constexpr size_t kScratchBytes = kScratchMiB << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoallocDevice {
    bool ready = false;
    void* ptr = nullptr;       // device address of this GPU's scratch
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[kMaxDevices];

// Set-up on `dev` (made current for the call): the scratch address and three timing
// events, once per device for the life of the process.
int noalloc_device(int dev, NoallocDevice*& out)
{
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index out of range");
        return GPSAT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoallocDevice& nd = g_noalloc[dev];
    if (!nd.ready) {
        int cur = 0;
        HIP_OR(hipGetDevice(&cur));
        if (cur != dev) HIP_OR(hipSetDevice(dev));
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        for (int k = 0; k < 3 && e == hipSuccess; k++) e = hipEventCreate(&nd.ev[k]);
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return hip_fail("gpsat_solve_noalloc set-up (hipGetSymbolAddress / events)", e);
        if (!ptr) {
            set_error("gpsat_solve_noalloc set-up: no device address for the scratch region");
            return GPSAT_ERR_HIP;
        }
        nd.ptr = ptr;
        nd.ready = true;
    }
    out = &nd;
    return GPSAT_OK;
}

}  // namespace

extern "C" int gpsat_solve_noalloc(gpsat_batch* b, const gpsat_solve_opts* opts, const int32_t* job_inst, const int32_t* cube_off,
                        const int32_t* cube_lits, int32_t n_jobs, const float* act_init, const int8_t* phase_init,
                        int32_t* status, uint32_t* model_rows, int32_t* verified, int64_t* job_stats,
                        int32_t* var_bumps, gpsat_solve_info* info)
{
    if (!b || !status) {
        set_error("null batch or status");
        return GPSAT_ERR_ARG;
    }
    gpsat_solve_opts o;
    if (opts) o = *opts;
    else gpsat_solve_opts_default(&o);
    if (b->d.n_inst == 0) return GPSAT_OK;
    NoallocDevice* nd = nullptr;
    if (int rc = noalloc_device(b->d.device, nd)) return rc;
    ScratchMem mem;
    mem.base = (char*)nd->ptr;
    mem.size = kScratchBytes;
    const int rc = solve_core(b, o, job_inst, cube_off, cube_lits, n_jobs, act_init, phase_init, status, model_rows,
                              verified, job_stats, var_bumps, info, mem, nd->ev);
    if (rc == GPSAT_ERR_NOMEM) {
        set_error(std::string("gpsat_solve_noalloc: the job exceeds the ") + std::to_string(kScratchMiB) +
                  " MiB scratch region (" + gpsat_last_error() + ")");
        return GPSAT_ERR_CAPACITY;
    }
    return rc;
}
