// ise_kf_backjump.hip -- the k-Sys backjump position of a 1-factorization
// test (alldata::setPlayerIndex / setPlayerIndexByPos,
// k-Factorization/TripleSys/Source/kmSupport.cpp:303-356), after the
// canonizer kernels have written each candidate's first witness index.
//
// The k-Sys enumerator builds a matrix round by round and player by player
// (position = round * n + slot); when a test finds a smaller relabelling it
// backtracks straight to the position this names: every matrix that agrees on
// the rounds before it and on its round up to that slot is not canonical
// either.  It starts at nrounds * n - 3 (cnv.cpp:213, nrows * players -
// groupSize - 1) and is lowered at the first smaller row of the witness's
// normal form when no earlier row of that form came from a later round.
//
// One lane per candidate: the witness relabelling is rebuilt from its index
// in the order of the canonizer that produced it (general: pair-slot
// arrangement x orientation x day; two-row P1F: ordered pair of rounds x start
// player; two-row U1F: ordered pair of rounds x cycle mapping), the rounds are
// relabelled and normalised (pairs by their smaller label, rows by the
// partner of label 0) and compared with the rounds as given.  Per-lane state
// is nibble-packed words (a round = one uint64 of 16 four-bit entries).
#include <hip/hip_runtime.h>
#include "ise_kfact.h"

namespace ise {

namespace {

constexpr uint32_t BJ_MAX = 16;

__device__ __forceinline__ uint32_t nb(uint64_t w, uint32_t i) { return (uint32_t)(w >> (4 * i)) & 0xfu; }
__device__ __forceinline__ uint64_t nb_set(uint64_t w, uint32_t i, uint32_t v) {
    return (w & ~(0xfull << (4 * i))) | ((uint64_t)v << (4 * i));
}

// Partner word of a round: nibble x = partner of player x.
__device__ uint64_t partners(uint64_t round, uint32_t n) {
    uint64_t p = 0;
    for (uint32_t q = 0; q < n; q += 2) {
        const uint32_t a = nb(round, q), b = nb(round, q + 1);
        p = nb_set(nb_set(p, a, b), b, a);
    }
    return p;
}

// Round relabelled by tr (nibble x = label of player x) in normal form.
__device__ uint64_t relabel_normal(uint64_t round, uint64_t tr, uint32_t n) {
    uint64_t lp = 0;   // nibble L = partner label of label L
    for (uint32_t q = 0; q < n; q += 2) {
        const uint32_t a = nb(tr, nb(round, q)), b = nb(tr, nb(round, q + 1));
        lp = nb_set(nb_set(lp, a, b), b, a);
    }
    uint64_t out = 0;
    uint32_t used = 0, k = 0;
    for (uint32_t L = 0; L < n; L++) {
        if (used >> L & 1u) continue;
        const uint32_t M = nb(lp, L);
        used |= (1u << L) | (1u << M);
        out = nb_set(nb_set(out, k, L), k + 1, M);
        k += 2;
    }
    return out;
}

// Cycles of the union of rounds a, b (partner words): from the smallest
// player not yet met, alternately the a- and b-partner; path entries as the
// pairs (x, a(x)), (a(x), b(a(x))), ...; cycles stably ordered by length.
struct Cycles {
    uint32_t n, len[4], start[4];
    uint8_t path[2 * BJ_MAX];
};

__device__ bool pair_cycles(uint64_t pa, uint64_t pb, uint32_t np, Cycles &c) {
    uint32_t used = 0, ip = 0;
    c.n = 0;
    for (uint32_t k0 = 0; k0 < np; k0++) {
        if (used >> k0 & 1u) continue;
        if (c.n == 4) return false;
        c.start[c.n] = ip / 2;
        uint32_t k = k0, len = 0;
        for (uint32_t guard = 0; guard < np; guard++) {
            const uint32_t k1 = nb(pa, k), k2 = nb(pb, k1);
            c.path[ip++] = (uint8_t)k;
            c.path[ip++] = (uint8_t)k1;
            c.path[ip++] = (uint8_t)k1;
            c.path[ip++] = (uint8_t)k2;
            used |= (1u << k) | (1u << k1);
            len += 2;
            if (k2 == k0) break;
            k = k2;
        }
        c.len[c.n++] = len;
    }
    for (uint32_t j = 1; j < c.n; j++)
        for (uint32_t i = 0; i + j < c.n; i++)
            if (c.len[i] > c.len[i + 1]) {
                uint32_t t = c.len[i]; c.len[i] = c.len[i + 1]; c.len[i + 1] = t;
                t = c.start[i]; c.start[i] = c.start[i + 1]; c.start[i + 1] = t;
            }
    return true;
}

// Ordered pair of rounds number `pair_idx` in the two-row canonizers' order:
// last = 1.., ind < last, then (ind, last) before (last, ind).
__device__ void ordered_pair(uint64_t pair_idx, uint32_t &r0, uint32_t &r1) {
    const uint64_t swap = pair_idx & 1u, pair = pair_idx >> 1;
    uint32_t last = 1;
    while ((uint64_t)(last + 1) * last / 2 <= pair) last++;
    const uint32_t ind = (uint32_t)(pair - (uint64_t)last * (last - 1) / 2);
    r0 = swap ? last : ind;
    r1 = swap ? ind : last;
}

// Relabelling of witness `w` (nibble x = label of player x); false if none.
template <int MODE>
__device__ bool witness_relabelling(uint64_t w, const uint64_t *rounds, const uint64_t *part, uint32_t nr,
                                    uint32_t n, const KfTables &qp, uint64_t &tr) {
    tr = 0;
    if (MODE == 0) {
        const uint32_t np = n / 2;
        const uint64_t day = w % nr, rest = w / nr;
        const uint64_t t = rest & ((1ull << np) - 1), h = rest >> np;
        if (h >= qp.count[np]) return false;
        const uint64_t a = qp.base[qp.offset[np] + h];
        for (uint32_t i = 0; i < np; i++) {
            const uint32_t base = (uint32_t)(a >> (8 * i)) & 0xffu;
            const uint32_t flip = (uint32_t)(t >> (np - 1 - i)) & 1u;
            tr = nb_set(tr, nb(rounds[day], 2 * i), base + flip);
            tr = nb_set(tr, nb(rounds[day], 2 * i + 1), base + 1 - flip);
        }
        return true;
    } else if (MODE == 1) {
        const uint64_t k = w % n;
        uint32_t r0, r1;
        ordered_pair(w / n, r0, r1);
        if (r0 >= nr || r1 >= nr) return false;
        uint32_t v1 = (uint32_t)k, it = 0;
        for (uint32_t s = 0; s < n; s += 2) {
            tr = nb_set(tr, v1, it);
            const uint32_t v2 = nb(part[r0], v1);
            it = nb(part[0], it);
            tr = nb_set(tr, v2, it);
            it = nb(part[1], it);
            v1 = nb(part[r1], v2);
        }
        return true;
    } else {
        Cycles ref, cy;
        if (!pair_cycles(part[0], part[1], n, ref)) return false;
        uint64_t T = 1;
        for (uint32_t c = 0; c < ref.n; c++) T *= ref.len[c];
        uint32_t g0 = 0, g1 = 0;   // the group of equal lengths that permutes (at most one of size > 1)
        for (uint32_t c = 0, e; c < ref.n; c = e) {
            for (e = c; e < ref.n && ref.len[e] == ref.len[c]; e++) {}
            for (uint32_t f = 2; f <= e - c; f++) T *= f;
            if (e - c > 1) { g0 = c; g1 = e; }
        }
        uint32_t r0, r1;
        ordered_pair(w / T, r0, r1);
        if (r0 >= nr || r1 >= nr || !pair_cycles(part[r0], part[r1], n, cy) || cy.n != ref.n) return false;
        uint64_t t = w % T;
        uint32_t dir[4], off[4], perm[4], sout[4];
        for (uint32_t c = 0; c < cy.n; c++) dir[c] = (uint32_t)(t >> (cy.n - 1 - c)) & 1u;
        t >>= cy.n;
        for (uint32_t c = 0; c < cy.n; c++) {
            off[c] = (uint32_t)(t % (cy.len[c] / 2));
            t /= cy.len[c] / 2;
            perm[c] = c;
        }
        // t = lexicographic rank of the permuting group's arrangement (factorial base)
        if (g1 > g0 + 1) {
            uint32_t avail = 0;
            for (uint32_t c = g0; c < g1; c++) avail |= 1u << c;
            for (uint32_t j = g0; j < g1; j++) {
                uint64_t fact = 1;
                for (uint32_t f = 2; f <= g1 - 1 - j; f++) fact *= f;
                uint64_t q = t / fact;
                t %= fact;
                for (uint32_t c = g0; c < g1; c++)
                    if (avail >> c & 1u) {
                        if (q == 0) {
                            perm[j] = c;
                            avail &= ~(1u << c);
                            break;
                        }
                        q--;
                    }
            }
        }
        for (uint32_t c = 0; c < cy.n; c++) sout[c] = cy.start[perm[c]];
        for (uint32_t c = 0; c < cy.n; c++) {
            const int L2 = 2 * (int)cy.len[c];
            int ip = dir[c] ? 4 * (int)off[c] + 1 : 4 * (int)off[c];
            const int step = dir[c] ? -1 : 1;
            for (int j = 0; j < L2; j++, ip += step) {
                if (ip >= L2) ip -= L2;
                else if (ip < 0) ip += L2;
                tr = nb_set(tr, cy.path[2 * sout[c] + ip], ref.path[2 * ref.start[c] + j]);
            }
        }
        return true;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void kf_backjump_kernel(uint32_t ncand, uint32_t stride,
                                                          const uint64_t *__restrict__ rounds,
                                                          const uint32_t *__restrict__ meta, KfOut o, KfTables qp) {
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < ncand; ci += gridDim.x * blockDim.x) {
        const uint32_t m = meta[ci], nr = m & 0xffu, n = (m >> 8) & 0xffu;
        const uint8_t v = o.verdict[ci];
        if (v > ISE_VERDICT_NONCANONICAL + 1 || nr < 1 || n < 4 || n > BJ_MAX || (n & 1u) || nr > n - 1 ||
            nr > stride) {
            o.player_index[ci] = 0xffffu;   // invalid record: no position
            continue;
        }
        const uint32_t start = nr * n - 3;
        uint16_t result = (uint16_t)start;
        const uint64_t w = o.witness[ci];
        uint64_t rw[KF_MAX_ROUNDS], part[KF_MAX_ROUNDS];
        for (uint32_t r = 0; r < nr; r++) {
            rw[r] = rounds[(size_t)ci * stride + r];
            part[r] = partners(rw[r], n);
        }
        uint64_t tr = 0;
        if (v == ISE_VERDICT_NONCANONICAL && w != ~0ull &&
            witness_relabelling<MODE>(w, rw, part, nr, n, qp, tr)) {
            // the witness's rounds in normal form, ordered by the partner of label 0
            uint64_t rel[KF_MAX_ROUNDS];
            int src_of_key[BJ_MAX];
            for (uint32_t q = 0; q < BJ_MAX; q++) src_of_key[q] = -1;
            for (uint32_t r = 0; r < nr; r++) {
                rel[r] = relabel_normal(rw[r], tr, n);
                src_of_key[nb(rel[r], 1)] = (int)r;
            }
            int day_max = -1;
            uint32_t row = 0;
            for (uint32_t key = 0; key < n; key++) {
                const int r = src_of_key[key];
                if (r < 0) continue;
                const uint64_t co = rel[r], ci_row = rw[row];
                uint32_t q = 0;
                while (q < n && nb(co, q) == nb(ci_row, q)) q++;
                if (day_max < 0) day_max = r;
                if (q == n) {                      // equal row
                    if (day_max < r) day_max = r;
                    row++;
                    continue;
                }
                if (nb(co, q) > nb(ci_row, q) || day_max > r) break;   // not smaller / kept
                const uint32_t ip = q < n - 3 ? q : n - 3;
                uint64_t label_to_player = 0, player_to_slot = 0;
                for (uint32_t x = 0; x < n; x++) {
                    label_to_player = nb_set(label_to_player, nb(tr, x), x);
                    player_to_slot = nb_set(player_to_slot, nb(rw[r], x), x);
                }
                uint32_t j = 0;
                for (uint32_t p = 0; p <= ip; p++) {
                    const uint32_t pos = nb(player_to_slot, nb(label_to_player, nb(co, p)));
                    j = pos > j ? pos : j;
                }
                const uint32_t idx = (uint32_t)r * n + j;
                result = (uint16_t)(idx < start ? idx : start);
                break;
            }
        }
        o.player_index[ci] = result;
    }
}

} // namespace

hipError_t kf_backjump_launch(int mode, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                              const uint32_t *meta, const KfOut &o, const KfTables &qp, hipStream_t stream) {
    if (!ncand || !o.player_index) return hipSuccess;
    if (!o.witness || !o.verdict) return hipErrorInvalidValue;
    const uint32_t blocks = (ncand + 255) / 256 < 4096 ? (ncand + 255) / 256 : 4096;
    switch (mode) {
    case 0: hipLaunchKernelGGL(kf_backjump_kernel<0>, dim3(blocks), dim3(256), 0, stream, ncand, stride, rounds, meta, o, qp); break;
    case 1: hipLaunchKernelGGL(kf_backjump_kernel<1>, dim3(blocks), dim3(256), 0, stream, ncand, stride, rounds, meta, o, qp); break;
    default: hipLaunchKernelGGL(kf_backjump_kernel<2>, dim3(blocks), dim3(256), 0, stream, ncand, stride, rounds, meta, o, qp); break;
    }
    return hipGetLastError();
}

} // namespace ise
