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
//
// Split launch.  Most candidates are decided within a chunk or two, but a
// canonical one must try every relabelling (46080 x rounds for K_12), which
// on one wave is the whole batch's tail.  With a first-pass limit the batch
// runs in three kernels:
//   1. kf_check_kernel: every candidate up to `limit` relabellings; those
//      still undecided are listed with their automorphism count so far;
//   2. kf_range_kernel: the rest of each listed candidate's index space is
//      cut into G contiguous ranges taken by independent waves; each range
//      records its first smaller index and the automorphisms before it, and
//      publishes the index (atomicMin) so that ranges lying wholly after it
//      stop early;
//   3. kf_combine_kernel: per listed candidate, the ranges in index order up
//      to the first that holds a smaller relabelling.
// The results are those of the single scan: the first range with a witness
// holds the least one, every range before it was scanned in full, and a
// range that stopped early lies after a range with a witness.
#include <hip/hip_runtime.h>
#include "ise_kfact.h"

namespace ise {

constexpr uint32_t KF_NONE = 0xffffffffu;     // no smaller relabelling in the scanned range
constexpr uint32_t KF_PRUNED = 0xfffffffeu;   // range abandoned: a smaller one lies before it
constexpr uint32_t KF_STOP_EVERY = 4;         // chunks between reads of the published witness

__device__ __forceinline__ uint32_t nib(uint64_t w, uint32_t i) { return (uint32_t)(w >> (4 * i)) & 0xfu; }

// Ranges per listed candidate: as many as the range words allow, at most
// max_ranges.  Both the range and the combine kernel derive it from the list
// length, so they agree.
__device__ __forceinline__ uint32_t kf_ranges(uint32_t listed, uint32_t range_words, uint32_t max_ranges) {
    if (listed == 0) return 1u;
    const uint32_t g = range_words / listed;
    return g < 1u ? 1u : (g > max_ranges ? max_ranges : g);
}

__device__ __forceinline__ uint32_t kf_total(const KfTables &qp, uint32_t np, uint32_t nr) {
    return qp.count[np] * ((1u << np) * nr);
}

struct KfLds {
    uint32_t plo[KF_MAX_ROUNDS], phi[KF_MAX_ROUNDS];   // partner of each player, per round
    uint32_t rlo[KF_MAX_ROUNDS], rhi[KF_MAX_ROUNDS];   // the rounds as given (player per position)
    uint64_t key[KF_MAX_ROUNDS];                        // the rounds as comparison keys
    uint32_t fast, skip0;
};

// A 16-entry table of 4-bit values held in two 32-bit words, entry i at bits
// 4i of lo:hi.  tget reads entry i; tset sets entry i (still zero) to v; rset
// does the same in the reversed layout, entry 0 the most significant nibble
// of hi.
__device__ __forceinline__ uint32_t tget(uint32_t lo, uint32_t hi, uint32_t i) {
    return __builtin_amdgcn_ubfe((i & 8u) ? hi : lo, i << 2, 4);
}
__device__ __forceinline__ void tset(uint32_t &lo, uint32_t &hi, uint32_t i, uint32_t v) {
    const uint32_t m = v << ((i & 7u) << 2);
    if (i & 8u) hi |= m; else lo |= m;
}
__device__ __forceinline__ void rset(uint32_t &lo, uint32_t &hi, uint32_t i, uint32_t v) {
    const uint32_t m = v << ((~i & 7u) << 2);
    if (i & 8u) lo |= m; else hi |= m;
}

// Comparison keys.  Two rounds in normal form (pairs (a, b) with a < b,
// pairs by first element) compare exactly as their partner arrays p[0], p[1],
// ... compare lexicographically: the positions before the next pair's first
// element are covered by equal earlier pairs, and at that element the arrays
// hold the partners the pair comparison looks at.  When every given round is
// in this form (`fast`) a round's key is its partner array, p[0] most
// significant; otherwise the key is the round's players in given order and the
// relabelled round is spelled out pair by pair after normalising.  Round 0 of
// every relabelled matrix is its day round, relabelled to (0 1)(2 3)...; when
// the given round 0 is that, it compares equal and is skipped (`skip0`).

// Lanes < nr decode their round (`row`, loaded by the caller) into the LDS
// tables.  Ends with a barrier.
__device__ __forceinline__ void kf_stage(KfLds &s, uint64_t row, uint32_t nr, uint32_t n, uint32_t l) {
    bool normal = true;
    uint32_t plo = 0, phi = 0, klo = 0, khi = 0;
    uint64_t raw = 0;
    if (l < nr) {
        for (uint32_t p = 0; p < n; p += 2) {
            const uint32_t x = nib(row, p), y = nib(row, p + 1);
            tset(plo, phi, x, y);
            tset(plo, phi, y, x);
            rset(klo, khi, x, y);
            rset(klo, khi, y, x);
            normal = normal && x < y && (p == 0 || x > nib(row, p - 2));
        }
        for (uint32_t p = 0; p < n; p++) raw = (raw << 4) | nib(row, p);
    }
    const bool fast = __builtin_amdgcn_ballot_w64(l < nr && !normal) == 0;
    if (l < nr) {
        s.plo[l] = plo;
        s.phi[l] = phi;
        s.rlo[l] = (uint32_t)row;
        s.rhi[l] = (uint32_t)(row >> 32);
        s.key[l] = fast ? (((uint64_t)khi << 32) | klo) : raw;
    }
    if (l == 0) {
        uint64_t id = 0;
        for (uint32_t p = 0; p < n; p++) id = (id << 4) | p;
        s.fast = fast;
        s.skip0 = raw == id;
    }
    __syncthreads();
}

// Relabelling r against the matrix: -1 smaller, 0 equal, 1 larger.  A is the
// slot arrangement of r (tab[r / nr >> np]).
__device__ __forceinline__ int kf_compare(const KfLds &s, uint64_t A, uint32_t r, uint32_t nr, uint32_t n,
                                          uint32_t np, bool fast, uint32_t skip0) {
    const uint32_t day = r % nr, q = r / nr;
    const uint32_t t = q & ((1u << np) - 1u);
    const uint32_t dlo = s.rlo[day], dhi = s.rhi[day];
    // labels of this relabelling, and the player that becomes 0
    uint32_t tlo = 0, thi = 0, p0 = 0;
    for (uint32_t i = 0; i < np; i++) {
        const uint32_t a = (uint32_t)(A >> (8 * i)) & 0xffu;
        const uint32_t f = (t >> (np - 1 - i)) & 1u;
        const uint32_t w = i < 4 ? dlo : dhi;
        const uint32_t x0 = __builtin_amdgcn_ubfe(w, i << 3, 4), x1 = __builtin_amdgcn_ubfe(w, (i << 3) + 4, 4);
        tset(tlo, thi, x0, a + f);
        tset(tlo, thi, x1, a + 1 - f);
        if (a == 0) p0 = f ? x1 : x0;
    }
    // rounds ordered by the new partner of player 0
    uint32_t present = 0, kylo = 0, kyhi = 0;
    for (uint32_t k = 0; k < nr; k++) {
        const uint32_t key = tget(tlo, thi, tget(s.plo[k], s.phi[k], p0));
        present |= 1u << key;
        tset(kylo, kyhi, key, k);
    }
    uint32_t i = 0;
    if (skip0) {
        present &= present - 1;
        i = 1;
    }
    for (; present; i++) {
        const uint32_t key = __builtin_ctz(present);
        present &= present - 1;
        const uint32_t k = tget(kylo, kyhi, key);
        uint64_t enc;
        if (fast) {
            const uint32_t rlo = s.rlo[k], rhi = s.rhi[k];
            uint32_t elo = 0, ehi = 0;   // relabelled partner array, reversed
            for (uint32_t j = 0; j < np; j++) {
                const uint32_t w = j < 4 ? rlo : rhi;
                const uint32_t x = tget(tlo, thi, __builtin_amdgcn_ubfe(w, j << 3, 4));
                const uint32_t y = tget(tlo, thi, __builtin_amdgcn_ubfe(w, (j << 3) + 4, 4));
                rset(elo, ehi, x, y);
                rset(elo, ehi, y, x);
            }
            enc = ((uint64_t)ehi << 32) | elo;
        } else {
            const uint32_t klo = s.plo[k], khi = s.phi[k];
            uint32_t qlo = 0, qhi = 0;   // relabelled partner array
            for (uint32_t x = 0; x < n; x++)
                tset(qlo, qhi, tget(tlo, thi, x), tget(tlo, thi, tget(klo, khi, x)));
            enc = 0;   // pairs (v, partner) for v ascending, v < partner
            for (uint32_t v = 0; v < n; v++) {
                const uint32_t w = tget(qlo, qhi, v);
                if (w > v) enc = (enc << 8) | (v << 4) | w;
            }
        }
        const uint64_t e = s.key[i];
        if (enc != e) return enc < e ? -1 : 1;
    }
    return 0;
}

// Relabellings [lo, hi) in chunks of 64.  Returns the first smaller index
// (KF_NONE if none) and adds the automorphisms met before it to `auts`.  With
// `stop`, every few chunks the scan reads the least witness published for the
// candidate and returns KF_PRUNED once its next chunk starts past it.  The
// arrangement of the next chunk is loaded while the current one is compared.
__device__ __forceinline__ uint32_t kf_scan(const KfLds &s, const uint64_t *tab, uint32_t lo, uint32_t hi,
                                            uint32_t nr, uint32_t n, uint32_t np, const uint32_t *stop,
                                            uint64_t &auts, uint32_t l) {
    const bool fast = s.fast != 0;
    const uint32_t skip0 = s.skip0;
    const uint32_t shift = np;
    uint32_t since = 0;
    uint64_t A = lo + l < hi ? tab[(lo + l) / nr >> shift] : 0;
    for (uint32_t base = lo; base < hi; base += 64) {
        if (stop && ++since == KF_STOP_EVERY) {
            since = 0;
            const uint32_t best = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (best < base) return KF_PRUNED;
        }
        const uint32_t r = base + l;
        const uint32_t rn = r + 64;
        const uint64_t An = rn < hi ? tab[rn / nr >> shift] : 0;
        const int icmp = r < hi ? kf_compare(s, A, r, nr, n, np, fast, skip0) : 1;
        const uint64_t smaller = __builtin_amdgcn_ballot_w64(icmp < 0);
        const uint64_t equal = __builtin_amdgcn_ballot_w64(r < hi && icmp == 0);
        if (smaller) {
            const uint32_t first = __builtin_ctzll(smaller);
            auts += __popcll(equal & ((1ull << first) - 1ull));
            return base + first;
        }
        auts += __popcll(equal);
        A = An;
    }
    return KF_NONE;
}

__device__ __forceinline__ void kf_write(const KfOut &o, uint32_t ci, uint32_t witness, uint64_t auts,
                                         uint32_t total) {
    const bool canon = witness == KF_NONE;
    o.verdict[ci] = canon ? ISE_VERDICT_CANONICAL : ISE_VERDICT_NONCANONICAL;
    if (o.witness) o.witness[ci] = canon ? ~0ull : (uint64_t)witness;
    if (o.n_auts) o.n_auts[ci] = auts;
    if (o.n_tried) o.n_tried[ci] = canon ? (uint64_t)total : (uint64_t)witness + 1;
}

__global__ void kf_queue_reset_kernel(uint32_t *head) {
    if (threadIdx.x < 3) head[KF_HEAD_WORD * threadIdx.x] = 0u;
}

__device__ __forceinline__ bool kf_meta_ok(uint32_t m, uint32_t stride) {
    const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff;
    return n >= 4 && n <= KF_MAX_PLAYERS && (n & 1) == 0 && nr >= 1 && nr <= n - 1 && nr <= stride &&
           nr <= KF_MAX_ROUNDS;
}

// Pass 1 (or the whole test when limit == 0).
__global__ __launch_bounds__(64) void kf_check_kernel(uint32_t ncand, uint32_t stride,
                                                      const uint64_t *__restrict__ rounds,
                                                      const uint32_t *__restrict__ meta, KfOut o,
                                                      uint32_t *head, KfTables qp, uint32_t limit, KfSplit sp) {
    __shared__ KfLds s;
    __shared__ uint32_t s_ci;
    const uint32_t l = threadIdx.x;
    const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;   // round words a lane may load
    // With a limit every candidate's time is bounded, so the candidates go by
    // a fixed grid stride and the next one is loaded while this one runs (one
    // queue word taken by every wave of the chip would serialise the short
    // tests).  Without one, the first candidate by position, the later ones
    // from the queue.
    uint32_t ci = blockIdx.x;
    uint32_t m = ci < ncand ? meta[ci] : 0u;
    uint64_t row = (ci < ncand && l < rl) ? rounds[(uint64_t)ci * stride + l] : 0ull;
    while (ci < ncand) {
        uint32_t cn = ci + gridDim.x, mn = 0;
        uint64_t rown = 0;
        if (limit && cn < ncand) {
            mn = meta[cn];
            if (l < rl) rown = rounds[(uint64_t)cn * stride + l];
        }
        if (!kf_meta_ok(m, stride)) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
        } else {
            const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff, np = n / 2;
            kf_stage(s, row, nr, n, l);
            const uint32_t total = kf_total(qp, np, nr);
            const uint32_t hi = (limit && limit < total) ? limit : total;
            uint64_t auts = 0;
            const uint32_t witness = kf_scan(s, qp.base + qp.offset[np], 0, hi, nr, n, np, nullptr, auts, l);
            if (l == 0) {
                if (witness == KF_NONE && hi < total) {
                    const uint32_t j = atomicAdd(&head[KF_HEAD_WORD], 1u);
                    sp.list[j] = ci;
                    sp.part[j] = auts;
                    sp.best[j] = KF_NONE;
                } else {
                    kf_write(o, ci, witness, auts, total);
                }
            }
        }
        __syncthreads();
        if (!limit) {
            if (l == 0) s_ci = atomicAdd(head, 1u) + gridDim.x;
            __syncthreads();
            cn = s_ci;
            __syncthreads();
            if (cn < ncand) {
                mn = meta[cn];
                if (l < rl) rown = rounds[(uint64_t)cn * stride + l];
            }
        }
        ci = cn;
        m = mn;
        row = rown;
    }
}

// Pass 2: ranges of the listed candidates' remaining relabellings.
__global__ __launch_bounds__(64) void kf_range_kernel(uint32_t stride, const uint64_t *__restrict__ rounds,
                                                      const uint32_t *__restrict__ meta, uint32_t *head,
                                                      KfTables qp, uint32_t limit, KfSplit sp) {
    __shared__ KfLds s;
    __shared__ uint32_t s_item;
    const uint32_t l = threadIdx.x;
    const uint32_t listed = head[KF_HEAD_WORD];
    const uint32_t G = kf_ranges(listed, sp.range_words, sp.max_ranges);
    const uint32_t items = listed * G;
    for (uint32_t it = 0;; it++) {
        uint32_t item = blockIdx.x;
        if (it) {
            if (l == 0) s_item = atomicAdd(&head[2 * KF_HEAD_WORD], 1u) + gridDim.x;
            __syncthreads();
            item = s_item;
            __syncthreads();
        }
        if (item >= items) break;
        const uint32_t j = item / G, g = item - j * G;
        const uint32_t ci = sp.list[j];
        const uint32_t m = meta[ci];
        const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff, np = n / 2;
        const uint32_t total = kf_total(qp, np, nr);
        const uint32_t chunks = (total - limit + 63u) / 64u;
        const uint32_t per = (chunks + G - 1u) / G;
        const uint64_t lo64 = (uint64_t)limit + (uint64_t)g * per * 64u;
        uint32_t witness = KF_NONE;
        uint64_t auts = 0;
        if (lo64 < total) {
            const uint32_t lo = (uint32_t)lo64;
            const uint32_t hi = total - lo > per * 64u ? lo + per * 64u : total;
            const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;
            kf_stage(s, l < rl ? rounds[(uint64_t)ci * stride + l] : 0ull, nr, n, l);
            witness = kf_scan(s, qp.base + qp.offset[np], lo, hi, nr, n, np, &sp.best[j], auts, l);
            if (l == 0 && witness < KF_PRUNED) atomicMin(&sp.best[j], witness);
        }
        if (l == 0) sp.ranges[(uint64_t)j * G + g] = (auts << 32) | witness;
        __syncthreads();
    }
}

// Pass 3: per listed candidate, the ranges in index order.
__global__ void kf_combine_kernel(const uint32_t *__restrict__ meta, const uint32_t *head, KfTables qp,
                                  KfSplit sp, KfOut o) {
    const uint32_t listed = head[KF_HEAD_WORD];
    const uint32_t G = kf_ranges(listed, sp.range_words, sp.max_ranges);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < listed; j += gridDim.x * blockDim.x) {
        const uint32_t ci = sp.list[j];
        const uint32_t m = meta[ci];
        const uint32_t nr = m & 0xff, np = ((m >> 8) & 0xff) / 2;
        uint64_t auts = sp.part[j];
        uint32_t witness = KF_NONE;
        for (uint32_t g = 0; g < G; g++) {
            const uint64_t e = sp.ranges[(uint64_t)j * G + g];
            auts += e >> 32;
            if ((uint32_t)e != KF_NONE) {
                witness = (uint32_t)e;   // a real index: see the ordering argument above
                break;
            }
        }
        kf_write(o, ci, witness, auts, kf_total(qp, np, nr));
    }
}

hipError_t kf_launch(uint32_t blocks, uint32_t range_blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                     const uint32_t *meta, const KfOut &o, uint32_t *head, const KfTables &qp,
                     uint32_t limit, const KfSplit &sp, hipStream_t stream) {
    hipLaunchKernelGGL(kf_queue_reset_kernel, dim3(1), dim3(64), 0, stream, head);
    hipLaunchKernelGGL(kf_check_kernel, dim3(blocks), dim3(64), 0, stream, ncand, stride, rounds, meta, o,
                       head, qp, limit, sp);
    if (limit) {
        hipLaunchKernelGGL(kf_range_kernel, dim3(range_blocks), dim3(64), 0, stream, stride, rounds, meta, head, qp,
                           limit, sp);
        hipLaunchKernelGGL(kf_combine_kernel, dim3(64), dim3(256), 0, stream, meta, head, qp, sp, o);
    }
    return hipGetLastError();
}

hipError_t kf_occupancy(int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)&kf_check_kernel, 64, 0);
}

} // namespace ise
