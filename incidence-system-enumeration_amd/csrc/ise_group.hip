// ise_group.hip -- the automorphism generators returned by the canonicity
// test, after the test: closing them to the group, re-indexing them to a
// subset of column orbits, and mapping row solutions to the representative
// of their orbit under the group.
//
// Reference: CPermutStorage::constructGroup / adjustGenerators /
// findSolutionIndex (Source/PermutStorage.h:155-335), used by
// CRowSolution::sortSolutionsByGroup (Source/RowSolution.h:325-383) and
// CInsSysEnumerator (Source/InsSysEnumerator.h:297).
//
// The closure runs on the host as a breadth-first product with the
// generators (each element is multiplied by every generator once, a hash set
// removes repeats), O(|G| x generators x len) instead of the reference's
// pairwise products; the resulting set is the same group, listed identity
// first.  The orbit mapping is data-parallel -- every solution against every
// group element -- and runs on the GPU: one wave per solution, each lane a
// strided share of the group, images packed big-endian into 32-bit words so
// that lexicographic order is word order, and a cross-lane max.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include "ise_internal.h"

namespace ise {

constexpr int SOL_WORDS = ISE_MAX_COLS / 4;   // up to 64 bytes per solution

// lexicographic a < b over SOL_WORDS packed words
__device__ __forceinline__ bool lex_less(const uint32_t *a, const uint32_t *b) {
    for (int w = 0; w < SOL_WORDS; w++)
        if (a[w] != b[w]) return a[w] < b[w];
    return false;
}

__global__ __launch_bounds__(64) void solution_canon_kernel(const uint8_t *__restrict__ group, uint64_t order,
                                                            uint32_t len, const uint8_t *__restrict__ sols,
                                                            uint32_t n_sol, uint8_t *__restrict__ canon) {
    __shared__ uint8_t s_sol[ISE_MAX_COLS];
    const uint32_t s = blockIdx.x;
    const int l = threadIdx.x;
    if (s >= n_sol) return;
    const uint8_t *sol = sols + (uint64_t)s * len;
    s_sol[l] = l < (int)len ? sol[l] : 0;
    __syncthreads();
    // best image of this lane's share; the solution itself (the identity) to start
    uint32_t best[SOL_WORDS], img[SOL_WORDS];
    for (int w = 0; w < SOL_WORDS; w++) {
        uint32_t x = 0;
        for (int t = 0; t < 4; t++) x = (x << 8) | s_sol[4 * w + t];
        best[w] = x;
    }
    for (uint64_t e = l; e < order; e += 64) {
        const uint8_t *p = group + e * len;
        for (int w = 0; w < SOL_WORDS; w++) {
            uint32_t x = 0;
            for (int t = 0; t < 4; t++) {
                const int i = 4 * w + t;
                x = (x << 8) | (i < (int)len ? s_sol[p[i]] : 0u);
            }
            img[w] = x;
        }
        if (lex_less(best, img))
            for (int w = 0; w < SOL_WORDS; w++) best[w] = img[w];
    }
    // cross-lane maximum
    for (int off = 32; off > 0; off >>= 1) {
        for (int w = 0; w < SOL_WORDS; w++) img[w] = (uint32_t)__shfl_xor((int)best[w], off);
        if (lex_less(best, img))
            for (int w = 0; w < SOL_WORDS; w++) best[w] = img[w];
    }
    uint32_t mine = 0;
    for (int w = 0; w < SOL_WORDS; w++)
        if ((l >> 2) == w) mine = best[w];
    if (l < (int)len) canon[(uint64_t)s * len + l] = (uint8_t)(mine >> (8 * (3 - (l & 3))));
}

} // namespace ise

using namespace ise;

extern "C" {

int ise_group_closure(const uint8_t *gens, uint32_t n_gens, uint32_t len, uint8_t *elements, uint64_t cap,
                      uint64_t *order_out) {
    if (!order_out || len == 0 || len > 255 || (n_gens && !gens) || (cap && !elements)) return ISE_ERR_ARG;
    *order_out = 0;
    for (uint32_t g = 0; g < n_gens; g++) {   // every generator must be a permutation of 0..len-1
        std::vector<uint8_t> seen(len, 0);
        for (uint32_t i = 0; i < len; i++) {
            const uint8_t x = gens[(size_t)g * len + i];
            if (x >= len || seen[x]) return ISE_ERR_ARG;
            seen[x] = 1;
        }
    }
    std::vector<uint8_t> list(len);
    for (uint32_t i = 0; i < len; i++) list[i] = (uint8_t)i;
    std::unordered_set<std::string> seen;
    seen.emplace((const char *)list.data(), len);
    std::vector<uint8_t> y(len);
    for (size_t at = 0; at < list.size() / len; at++) {
        for (uint32_t g = 0; g < n_gens; g++) {
            const uint8_t *x = &list[at * len];
            const uint8_t *gp = gens + (size_t)g * len;
            for (uint32_t k = 0; k < len; k++) y[k] = x[gp[k]];   // (x * g)[k] = x[g[k]]
            if (seen.emplace((const char *)y.data(), len).second) {
                list.insert(list.end(), y.begin(), y.end());
                if (list.size() / len > cap) {
                    *order_out = list.size() / len;
                    return ISE_ERR_CAPACITY;
                }
            }
        }
    }
    const uint64_t order = list.size() / len;
    if (order > cap) {
        *order_out = order;
        return ISE_ERR_CAPACITY;
    }
    memcpy(elements, list.data(), list.size());
    *order_out = order;
    return ISE_OK;
}

int ise_adjust_generators(const uint8_t *gens, uint32_t n_gens, uint32_t len, const uint32_t *idx,
                          uint32_t len_idx, uint8_t *out) {
    if ((n_gens && (!gens || !out)) || !idx || len == 0 || len > 255 || len_idx == 0 || len_idx > len)
        return ISE_ERR_ARG;
    std::vector<int> from(len, -1);
    for (uint32_t i = 0; i < len_idx; i++) {
        if (idx[i] >= len || from[idx[i]] >= 0) return ISE_ERR_ARG;
        from[idx[i]] = (int)i;
    }
    for (uint32_t g = 0; g < n_gens; g++)
        for (uint32_t j = 0; j < len_idx; j++) {
            const int v = from[gens[(size_t)g * len + idx[j]]];
            if (v < 0) return ISE_ERR_ARG;   // the subset is not invariant under this generator
            out[(size_t)g * len_idx + j] = (uint8_t)v;
        }
    return ISE_OK;
}

int ise_solution_orbits(ise_ctx *ctx, const uint8_t *group, uint64_t order, uint32_t len, const uint8_t *sols,
                        uint32_t n_sol, uint8_t *canon_out, uint32_t *rep_out) {
    if (!ctx || !group || order == 0 || len == 0 || len > ISE_MAX_COLS || (n_sol && (!sols || !canon_out)))
        return ISE_ERR_ARG;
    if (n_sol == 0) return ISE_OK;
    for (uint64_t e = 0; e < order * len; e++)
        if (group[e] >= len) return ISE_ERR_ARG;
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ISE_ERR_HIP;
    hipStream_t s = ctx_stream(ctx);
    uint8_t *d_group = nullptr, *d_sols = nullptr, *d_canon = nullptr;
    const size_t gb = (size_t)order * len, sb = (size_t)n_sol * len;
    hipError_t e = hipMalloc((void **)&d_group, gb);
    if (e == hipSuccess) e = hipMalloc((void **)&d_sols, sb);
    if (e == hipSuccess) e = hipMalloc((void **)&d_canon, sb);
    if (e == hipSuccess) e = hipMemcpyAsync(d_group, group, gb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sols, sols, sb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(solution_canon_kernel, dim3(n_sol), dim3(64), 0, s, d_group, order, len, d_sols, n_sol,
                           d_canon);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(canon_out, d_canon, sb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_group) (void)hipFree(d_group);
    if (d_sols) (void)hipFree(d_sols);
    if (d_canon) (void)hipFree(d_canon);
    if (e != hipSuccess) {
        ctx_set_error(ctx, "solution orbit kernel failed");
        return ISE_ERR_HIP;
    }
    if (rep_out) {
        // representative: the first listed solution equal to the canonical image
        std::unordered_map<std::string, uint32_t> first;
        for (uint32_t t = 0; t < n_sol; t++) first.emplace(std::string((const char *)sols + (size_t)t * len, len), t);
        for (uint32_t t = 0; t < n_sol; t++) {
            auto it = first.find(std::string((const char *)canon_out + (size_t)t * len, len));
            rep_out[t] = it == first.end() ? 0xffffffffu : it->second;
        }
    }
    return ISE_OK;
}

} // extern "C"
