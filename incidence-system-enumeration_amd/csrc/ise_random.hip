// ise_random.hip -- seeded random valid extensions generated on the GPU: one
// thread runs the walk of ise_random.h for one candidate, so the batch equals
// the host threads' one for the same seed.  Rejection sampling of the depth
// (many walks for a near-complete depth) makes the host version slow; here the
// walks of a whole batch run side by side.
#include <hip/hip_runtime.h>
#include "ise_internal.h"
#include "ise_random.h"

namespace ise {

__global__ __launch_bounds__(64) void random_extensions_kernel(ise_rw::Params p, uint64_t n, int dmin, int dmax,
                                                               uint64_t seed, uint32_t attempts, uint32_t stride,
                                                               uint64_t *rows, uint32_t *meta) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ise_rw::batch_candidate(p, i, dmin, dmax, seed, attempts, stride, rows, meta);
}

int random_extensions_device(ise_ctx *c, const ise_rw::Params &p, uint64_t n, int dmin, int dmax, uint64_t seed,
                             uint32_t attempts, uint32_t stride, uint64_t *rows_out, uint32_t *meta_out) {
    if (n == 0) return ISE_OK;
    if (n > (1ull << 31)) return ISE_ERR_ARG;
    uint64_t *d_rows = nullptr;
    uint32_t *d_meta = nullptr;
    hipStream_t s = ctx_stream(c);
    hipError_t e = hipSetDevice(ctx_device(c));
    if (e == hipSuccess) e = hipMalloc(&d_rows, n * stride * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&d_meta, n * sizeof(uint32_t));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(random_extensions_kernel, dim3((uint32_t)((n + 63) / 64)), dim3(64), 0, s, p, n, dmin,
                           dmax, seed, attempts, stride, d_rows, d_meta);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(rows_out, d_rows, n * stride * sizeof(uint64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(meta_out, d_meta, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_rows) (void)hipFree(d_rows);
    if (d_meta) (void)hipFree(d_meta);
    if (e != hipSuccess) {
        ctx_set_error(c, hipGetErrorString(e));
        return ISE_ERR_HIP;
    }
    return ISE_OK;
}

} // namespace ise
