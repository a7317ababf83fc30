// ise_noalloc.h -- the scratch array of the allocation-free entry point
// (ise_check_batch_device_noalloc, include/ise.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "ise.h"

namespace ise {

// Device address of the scratch array on `device`, which must be the current
// device (looked up once per device with hipGetSymbolAddress).  Layout:
// ISE_QUEUE_WORDS queue words, then the schedule list of up to
// ISE_NOALLOC_MAX_CANDIDATES entries and its SCHED_EXTRA_WORDS counters.
hipError_t noalloc_scratch_init(int device, uint32_t **out);
size_t noalloc_scratch_words();

} // namespace ise
