// ise_kfact.hip -- canonicity of partial 1-factorizations of K_n on gfx950.
//
// One wavefront (one 64-thread workgroup) per candidate; the 64 lanes take 64
// consecutive relabellings of the plain order at a time.  The plain order is
// the reference's (alldata::canonizator / cnvCheckTgNew / cnvCheckKm1,
// k-Factorization/TripleSys/Source/cnv.cpp:65-351): arrangements of the n/2
// pair slots in QuickPerm order (outer), pair orientations as an odometer with
// the last slot fastest (middle), days 0..nrounds-1 (inner).  Relabelling r of
// day d sends the pair in slot i of round d to labels (a_i, a_i+1) in the
// slot's orientation; the relabelled matrix is normalised (pairs sorted, pairs
// by first element, rounds by the partner of player 0) and compared round by
// round with the original (kmProcessMatrix, kmSupport.cpp:275-313).
//
// Per chunk a ballot of "smaller" decides: its lowest set lane is the first
// witness in plain order (the reference stops there); a ballot of "equal"
// counts automorphisms.  Everything is exact integer work on 4-bit labels
// packed into 64-bit words (n <= 16).
#include <hip/hip_runtime.h>
#include "ise_kfact.h"

namespace ise {

__device__ __forceinline__ uint32_t nib(uint64_t w, uint32_t i) { return (uint32_t)(w >> (4 * i)) & 0xfu; }

__global__ void kf_queue_reset_kernel(uint32_t *head) {
    if (threadIdx.x == 0) head[0] = 0u;
}

__global__ __launch_bounds__(64) void kf_check_kernel(uint32_t ncand, uint32_t stride,
                                                      const uint64_t *__restrict__ rounds,
                                                      const uint32_t *__restrict__ meta, KfOut o,
                                                      uint32_t *head, KfTables qp) {
    __shared__ uint64_t s_part[KF_MAX_ROUNDS];   // partner of each player, per round
    __shared__ uint64_t s_row[KF_MAX_ROUNDS];    // the rounds as given (player per position)
    __shared__ uint64_t s_enc[KF_MAX_ROUNDS];    // the rounds as comparison keys
    __shared__ uint32_t s_ci;
    const uint32_t l = threadIdx.x;
    for (;;) {
        if (l == 0) s_ci = atomicAdd(head, 1u);
        __syncthreads();
        const uint32_t ci = s_ci;
        __syncthreads();
        if (ci >= ncand) break;
        const uint32_t m = meta[ci];
        const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff;
        const bool ok = n >= 4 && n <= KF_MAX_PLAYERS && (n & 1) == 0 && nr >= 1 && nr <= n - 1 &&
                        nr <= stride && nr <= KF_MAX_ROUNDS;
        if (!ok) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            continue;
        }
        const uint32_t np = n / 2;
        if (l < nr) {
            const uint64_t row = rounds[(uint64_t)ci * stride + l];
            uint64_t part = 0, enc = 0;
            for (uint32_t p = 0; p < n; p += 2) {
                const uint64_t x = nib(row, p), y = nib(row, p + 1);
                part |= (y << (4 * x)) | (x << (4 * y));
            }
            for (uint32_t p = 0; p < n; p++) enc = (enc << 4) | nib(row, p);
            s_row[l] = row;
            s_part[l] = part;
            s_enc[l] = enc;
        }
        __syncthreads();
        const uint64_t *tab = qp.base + qp.offset[np];
        const uint32_t per_perm = (1u << np) * nr;
        const uint32_t total = qp.count[np] * per_perm;
        uint64_t auts = 0;
        uint32_t witness = 0xffffffffu;
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t r = base + l;
            int icmp = 1;
            if (r < total) {
                const uint32_t day = r % nr, q = r / nr;
                const uint32_t t = q & ((1u << np) - 1u), h = q >> np;
                const uint64_t A = tab[h];
                const uint64_t rd = s_row[day];
                // labels of this relabelling, and the player that becomes 0
                uint64_t ttr = 0;
                uint32_t p0 = 0;
                for (uint32_t i = 0; i < np; i++) {
                    const uint32_t a = (uint32_t)(A >> (8 * i)) & 0xffu;
                    const uint32_t f = (t >> (np - 1 - i)) & 1u;
                    const uint32_t x0 = nib(rd, 2 * i), x1 = nib(rd, 2 * i + 1);
                    ttr |= ((uint64_t)(a + f) << (4 * x0)) | ((uint64_t)(a + 1 - f) << (4 * x1));
                    if (a == 0) p0 = f ? x1 : x0;
                }
                // rounds ordered by the new partner of player 0
                uint32_t present = 0;
                uint64_t keyrow = 0;
                for (uint32_t k = 0; k < nr; k++) {
                    const uint32_t key = nib(ttr, nib(s_part[k], p0));
                    present |= 1u << key;
                    keyrow |= (uint64_t)k << (4 * key);
                }
                icmp = 0;
                for (uint32_t i = 0; present; i++) {
                    const uint32_t key = __builtin_ctz(present);
                    present &= present - 1;
                    const uint64_t pk = s_part[nib(keyrow, key)];
                    uint64_t pp = 0;   // partner array after relabelling
                    for (uint32_t x = 0; x < n; x++)
                        pp |= (uint64_t)nib(ttr, nib(pk, x)) << (4 * nib(ttr, x));
                    uint64_t enc = 0;  // pairs (v, partner) for v ascending, v < partner
                    for (uint32_t v = 0; v < n; v++) {
                        const uint32_t w = nib(pp, v);
                        if (w > v) enc = (enc << 8) | (v << 4) | w;
                    }
                    const uint64_t e = s_enc[i];
                    if (enc != e) {
                        icmp = enc < e ? -1 : 1;
                        break;
                    }
                }
            }
            const uint64_t smaller = __builtin_amdgcn_ballot_w64(r < total && icmp < 0);
            const uint64_t equal = __builtin_amdgcn_ballot_w64(r < total && icmp == 0);
            if (smaller) {
                const uint32_t first = __builtin_ctzll(smaller);
                witness = base + first;
                auts += __popcll(equal & ((1ull << first) - 1ull));
                break;
            }
            auts += __popcll(equal);
        }
        if (l == 0) {
            const bool canon = witness == 0xffffffffu;
            o.verdict[ci] = canon ? ISE_VERDICT_CANONICAL : ISE_VERDICT_NONCANONICAL;
            if (o.witness) o.witness[ci] = canon ? ~0ull : (uint64_t)witness;
            if (o.n_auts) o.n_auts[ci] = auts;
            if (o.n_tried) o.n_tried[ci] = canon ? (uint64_t)total : (uint64_t)witness + 1;
        }
        __syncthreads();
    }
}

hipError_t kf_launch(uint32_t blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                     const uint32_t *meta, const KfOut &o, uint32_t *head, const KfTables &qp,
                     hipStream_t stream) {
    hipLaunchKernelGGL(kf_queue_reset_kernel, dim3(1), dim3(64), 0, stream, head);
    hipLaunchKernelGGL(kf_check_kernel, dim3(blocks), dim3(64), 0, stream, ncand, stride, rounds, meta, o,
                       head, qp);
    return hipGetLastError();
}

hipError_t kf_occupancy(int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)&kf_check_kernel, 64, 0);
}

} // namespace ise
