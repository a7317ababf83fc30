// ise_noalloc.hip -- device scratch of the allocation-free entry point.
//
// ise_check_batch_device_noalloc keeps the work-queue words and the schedule
// list (batch size + SCHED_EXTRA_WORDS words) in the file-scope array below
// instead of caller or context buffers.  Each device holds its own copy for the
// life of the process.  Host code never passes the array's name to a kernel:
// it looks up the array's device address once per device (with that device
// current) and uses that address for every launch.
#include <hip/hip_runtime.h>
#include <mutex>
#include "ise_internal.h"
#include "ise_noalloc.h"

namespace ise {

constexpr size_t NOALLOC_WORDS = ISE_QUEUE_WORDS + ISE_NOALLOC_MAX_CANDIDATES + SCHED_EXTRA_WORDS;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static uint32_t *scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

hipError_t noalloc_scratch_init(int device, uint32_t **out) {
    *out = nullptr;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;   // the caller makes `device` current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[device] = (uint32_t *)p;
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}

size_t noalloc_scratch_words() { return NOALLOC_WORDS; }

} // namespace ise
