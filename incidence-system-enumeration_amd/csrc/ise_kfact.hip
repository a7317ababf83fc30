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
// runs as:
//   1. kf_check_kernel: every candidate up to `limit` relabellings; those
//      still undecided are listed with their automorphism count so far;
//   2. per window [W_p, W_p+1) of relabelling indices (W_0 = limit, each
//      window 16 times further out than the last):
//      kf_range_kernel cuts the window of each listed candidate into G
//      contiguous ranges taken by independent waves; each range records its
//      first smaller index and the automorphisms before it, and publishes the
//      index (atomicMin) so that ranges lying wholly after it stop early;
//      kf_combine_kernel, one wave per listed candidate, takes the ranges in
//      index order up to the first that holds a smaller relabelling, writes
//      the results of the candidates decided and lists the others for the
//      next window.
// The results are those of the single scan: the first range with a witness
// holds the least one, every range before it was scanned in full, and a
// range that stopped early lies after a range with a witness.  Windows keep
// the work of a candidate whose witness lies just past the limit near that
// witness instead of spreading it over the whole index space.
#include <hip/hip_runtime.h>
#include "ise_kfact.h"

namespace ise {

constexpr uint32_t KF_NONE = 0xffffffffu;     // no smaller relabelling in the scanned range
constexpr uint32_t KF_PRUNED = 0xfffffffeu;   // range abandoned: a smaller one lies before it
constexpr uint32_t KF_STOP_EVERY = 4;         // chunks between reads of the published witness

__device__ __forceinline__ uint32_t nib(uint64_t w, uint32_t i) { return (uint32_t)(w >> (4 * i)) & 0xfu; }

// Ranges per listed candidate in a window of `span` relabellings: as many as
// the range words allow, at most max_ranges, and at least KF_MIN_RANGE_CHUNKS
// chunks each (a range stages its candidate once, which costs about a chunk).
// Both the range and the combine kernel derive it from the same values, so
// they agree.
constexpr uint32_t KF_MIN_RANGE_CHUNKS = 3;
__device__ __forceinline__ uint32_t kf_ranges(uint32_t listed, uint32_t range_words, uint32_t max_ranges,
                                              uint32_t span) {
    if (listed == 0) return 1u;
    uint32_t g = range_words / listed;
    const uint32_t by_size = span / (64u * KF_MIN_RANGE_CHUNKS);
    if (g > max_ranges) g = max_ranges;
    if (g > by_size) g = by_size;
    return g < 1u ? 1u : g;
}

__device__ __forceinline__ uint32_t kf_total(const KfTables &qp, uint32_t np, uint32_t nr) {
    return qp.count[np] * ((1u << np) * nr);
}

struct KfLds {
    uint32_t plo[KF_MAX_ROUNDS], phi[KF_MAX_ROUNDS];   // partner of each player, per round
    uint32_t rlo[KF_MAX_ROUNDS], rhi[KF_MAX_ROUNDS];   // the rounds as given (player per position)
    uint64_t key[KF_MAX_ROUNDS];                        // the rounds as comparison keys (general form)
    uint64_t meet[KF_MAX_PLAYERS];   // meet[x] entry y: the round pairing x with y, 15 if none
};

// A 16-entry table of 4-bit values held in two 32-bit words, entry i at bits
// 4i of lo:hi.  tget reads entry i; tset sets entry i (still zero) to v; tfind
// returns the least i whose entry is z (z must occur among entries < 16).
__device__ __forceinline__ uint32_t tget(uint32_t lo, uint32_t hi, uint32_t i) {
    return __builtin_amdgcn_ubfe((i & 8u) ? hi : lo, i << 2, 4);
}
__device__ __forceinline__ void tset(uint32_t &lo, uint32_t &hi, uint32_t i, uint32_t v) {
    const uint32_t m = v << ((i & 7u) << 2);
    if (i & 8u) hi |= m; else lo |= m;
}
__device__ __forceinline__ uint32_t tfind(uint32_t lo, uint32_t hi, uint32_t z) {
    // a zero nibble of w ^ z*0x11111111 marks z; the lowest flagged nibble is
    // exact (a borrow only flags nibbles above a true zero)
    const uint32_t rep = z * 0x11111111u;
    uint32_t x = lo ^ rep;
    uint32_t m = (x - 0x11111111u) & ~x & 0x88888888u;
    if (m) return (uint32_t)__builtin_ctz(m) >> 2;
    x = hi ^ rep;
    m = (x - 0x11111111u) & ~x & 0x88888888u;
    return 8u + ((uint32_t)__builtin_ctz(m) >> 2);
}

// Two forms of the comparison.
//
// Label order ("fast", every given round in normal form: pairs (a, b) with
// a < b, pairs by first element).  Two rounds in normal form compare exactly
// as their partner arrays p[0], p[1], ... compare lexicographically: the
// positions before the next pair's first element are covered by equal earlier
// pairs, and at that element the arrays hold the partners the pair comparison
// looks at.  The relabelled matrix's rounds are ordered by the label of
// player 0's partner, p[0]; walking the labels L = 1, 2, ... with y the player
// relabelled to L, the round meet[p0][y] (p0 the player relabelled to 0) is
// the next round of the relabelled matrix and its key is L.  So a round is
// decided by comparing L with the given round's p[0], and only on equality by
// its partner array from position 1; the relabelling is needed only as the
// inverse map label -> player, which is a byte gather of the day round's
// pairs by the inverse slot arrangement (v_perm_b32) with the pairs of
// flipped slots nibble-swapped.
//
// General form (some given round not in normal form): the key is the round's
// players in given order, and the relabelled round is spelled out pair by
// pair after normalising -- the comparison the reference makes byte by byte.
//
// Round 0 of every relabelled matrix is its day round, relabelled to
// (0 1)(2 3)...; when the given round 0 is that, it compares equal and is
// skipped (`skip0`).  When in addition the rounds are in normal form and in
// order of player 0's partner (`ident`), relabelling 0 (day 0, identity
// arrangement, no pair flipped) is the identity and reproduces the matrix: it
// is counted as an automorphism without the comparison, which would otherwise
// walk every round of the first chunk of every candidate.

struct KfUni {   // per-candidate values held in scalar registers
    uint32_t nr, n, np;
    uint32_t fast, skip0, ident;
    uint32_t valid;                  // every round a perfect matching, no pair in two rounds
    uint32_t glo, ghi;               // entry i: partner of player 0 in round i as given
};

__device__ __forceinline__ uint32_t kf_uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }

// Lanes < nr decode their round (`row`, loaded by the caller) into the LDS
// tables, lanes < n build their player's meet row; returns the candidate's
// uniform values, among them `valid`: every round a perfect matching of
// players 0..n-1 and each player's partners over the rounds all different
// (no pair in two rounds).  Ends with a barrier.
__device__ __forceinline__ KfUni kf_stage(KfLds &s, uint64_t row, uint32_t nr, uint32_t n, uint32_t l) {
    const uint32_t rlo = (uint32_t)row, rhi = (uint32_t)(row >> 32);
    bool normal = true, matching = true;
    if (l < nr) {
        uint32_t plo = 0, phi = 0, prev = 0, seen = 0;
        for (uint32_t p = 0; p < n; p += 2) {
            const uint32_t w = p < 8 ? rlo : rhi;
            const uint32_t x = __builtin_amdgcn_ubfe(w, p << 2, 4), y = __builtin_amdgcn_ubfe(w, (p << 2) + 4, 4);
            tset(plo, phi, x, y);
            tset(plo, phi, y, x);
            normal = normal && x < y && (p == 0 || x > prev);
            prev = x;
            seen |= (1u << x) | (1u << y);
        }
        matching = seen == (1u << n) - 1u;
        s.plo[l] = plo;
        s.phi[l] = phi;
        s.rlo[l] = rlo;
        s.rhi[l] = rhi;
    }
    const uint32_t fast = __builtin_amdgcn_ballot_w64(l < nr && !normal) == 0;
    // round 0 as given is (0 1)(2 3)...: player p at position p
    const uint64_t id = 0xfedcba9876543210ull & (n >= 16 ? ~0ull : ((1ull << (4 * n)) - 1ull));
    const uint32_t skip0 = (uint32_t)(__builtin_amdgcn_ballot_w64(l == 0 && row == id) != 0);
    if (!fast && l < nr) {
        uint64_t raw = 0;   // the players in given order, first most significant
        for (uint32_t p = 0; p < n; p++) raw = (raw << 4) | nib(row, p);
        s.key[l] = raw;
    }
    __syncthreads();
    KfUni u{nr, n, n / 2, fast, skip0, 0u, 1u, 0u, 0u};
    // player l's partners over the rounds: nr different players, not l
    bool partners = true;
    if (l < n) {
        uint32_t mlo = 0, mhi = 0, pm = 0;   // 15 - round, so that absent pairs read 15 after inverting
        for (uint32_t k = 0; k < nr; k++) {
            const uint32_t y = tget(s.plo[k], s.phi[k], l);
            pm |= 1u << y;
            tset(mlo, mhi, y, 15u - k);
        }
        partners = (uint32_t)__builtin_popcount(pm) == nr && !((pm >> l) & 1u);
        if (fast) s.meet[l] = ~(((uint64_t)mhi << 32) | mlo);
    }
    u.valid = __builtin_amdgcn_ballot_w64(!matching || !partners) == 0;
    if (fast) {
        uint32_t glo = 0, ghi = 0;
        bool sorted = true;   // rounds in order of player 0's partner
        uint32_t prev = 0;
        for (uint32_t k = 0; k < nr; k++) {
            const uint32_t g = kf_uni(s.plo[k]) & 0xfu;
            tset(glo, ghi, k, g);
            sorted = sorted && (k == 0 || g > prev);
            prev = g;
        }
        u.glo = glo;
        u.ghi = ghi;
        u.ident = skip0 && sorted;
    }
    __syncthreads();
    return u;
}

constexpr int KF_OPEN = 2;   // kf_compare_labels: the first `lim` compared rounds are equal

__device__ __forceinline__ int kf_compare_labels(const KfLds &s, const KfUni &u, uint32_t vlo, uint32_t vhi,
                                                 uint32_t same0 = 15u, uint32_t same1 = 15u,
                                                 uint32_t lim = 0xffffffffu);

// Label-order comparison of relabelling r (I = inverse slot arrangement).
__device__ __forceinline__ int kf_compare_fast(const KfLds &s, const KfUni &u, uint64_t I, uint32_t r) {
    const uint32_t day = r % u.nr, q = r / u.nr;
    const uint32_t t = q & ((1u << u.np) - 1u);
    const uint32_t dlo = s.rlo[day], dhi = s.rhi[day];
    const uint32_t slo = (uint32_t)I, shi = (uint32_t)(I >> 32);
    // orientation of slot i (bit np-1-i of t) as byte i of f: 0xff = flipped
    const uint32_t tr = __builtin_bitreverse32(t) >> (32u - u.np);
    const uint32_t flo = ((tr & 0xfu) * 0x00204081u & 0x01010101u) * 0xffu;
    const uint32_t fhi = (((tr >> 4) & 0xfu) * 0x00204081u & 0x01010101u) * 0xffu;
    // byte j: the pair that receives labels 2j, 2j+1, and whether it is flipped
    const uint32_t plo = __builtin_amdgcn_perm(dhi, dlo, slo), phi = __builtin_amdgcn_perm(dhi, dlo, shi);
    const uint32_t mlo = __builtin_amdgcn_perm(fhi, flo, slo), mhi = __builtin_amdgcn_perm(fhi, flo, shi);
    const uint32_t wlo = ((plo >> 4) & 0x0f0f0f0fu) | ((plo << 4) & 0xf0f0f0f0u);
    const uint32_t whi = ((phi >> 4) & 0x0f0f0f0fu) | ((phi << 4) & 0xf0f0f0f0u);
    const uint32_t vlo = (wlo & mlo) | (plo & ~mlo), vhi = (whi & mhi) | (phi & ~mhi);   // label -> player
    return kf_compare_labels(s, u, vlo, vhi);
}

// The label-order walk for a relabelling given as its inverse map label ->
// player (vlo:vhi): -1 smaller, 0 equal, 1 larger, KF_OPEN once `lim` rounds
// were compared and found equal.  Rounds same0 / same1 are known to become
// the given rounds 0 / 1 exactly (two-row relabellings): they are not
// compared when they land in those positions.
__device__ __forceinline__ int kf_compare_labels(const KfLds &s, const KfUni &u, uint32_t vlo, uint32_t vhi,
                                                 uint32_t same0, uint32_t same1, uint32_t lim) {
    const uint64_t mt = s.meet[vlo & 0xfu];
    const uint32_t elo = (uint32_t)mt, ehi = (uint32_t)(mt >> 32);
    uint32_t i = 0;
    for (uint32_t L = 1; L < u.n; L++) {
        const uint32_t k = tget(elo, ehi, tget(vlo, vhi, L));
        if (k == 15u) continue;
        if ((i == 0 && (u.skip0 || k == same0)) || (i == 1 && k == same1)) {
            i++;
            continue;
        }
        const uint32_t g = tget(u.glo, u.ghi, i);
        if (L != g) return L < g ? -1 : 1;
        // The partner arrays can first differ only at the first element of a
        // given pair (at a second element b of pair (a, b), a < b, position
        // a already differs), so the walk visits the given pairs (a, gv) in
        // order; pair 0 is (0, L), equal here.
        const uint32_t klo = s.plo[k], khi = s.phi[k];
        const uint32_t qlo = s.rlo[i], qhi = s.rhi[i];
        for (uint32_t j = 1; j < u.np; j++) {
            const uint32_t pr = __builtin_amdgcn_ubfe(j < 4 ? qlo : qhi, (j & 3u) << 3, 8);
            const uint32_t a = pr & 0xfu, gv = pr >> 4;                   // given pair (a, gv)
            const uint32_t z = tget(klo, khi, tget(vlo, vhi, a));        // player whose label is a's partner
            if (z != tget(vlo, vhi, gv)) {
                const uint32_t lz = tfind(vlo, vhi, z);
                return lz < gv ? -1 : 1;
            }
        }
        i++;
        if (--lim == 0) return KF_OPEN;
    }
    return 0;
}

// General-form comparison of relabelling r (A = slot arrangement).
__device__ __forceinline__ int kf_compare_general(const KfLds &s, const KfUni &u, uint64_t A, uint32_t r) {
    const uint32_t nr = u.nr, n = u.n, np = u.np;
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
    if (u.skip0) {
        present &= present - 1;
        i = 1;
    }
    for (; present; i++) {
        const uint32_t key = __builtin_ctz(present);
        present &= present - 1;
        const uint32_t k = tget(kylo, kyhi, key);
        const uint32_t klo = s.plo[k], khi = s.phi[k];
        uint32_t qlo = 0, qhi = 0;   // relabelled partner array
        for (uint32_t x = 0; x < n; x++)
            tset(qlo, qhi, tget(tlo, thi, x), tget(tlo, thi, tget(klo, khi, x)));
        uint64_t enc = 0;   // pairs (v, partner) for v ascending, v < partner
        for (uint32_t v = 0; v < n; v++) {
            const uint32_t w = tget(qlo, qhi, v);
            if (w > v) enc = (enc << 8) | (v << 4) | w;
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
// A0: this lane's entry of the inverse table for relabelling lo + l, loaded
// by the caller before staging (reloaded from the forward table when the
// candidate takes the general form).
__device__ __forceinline__ uint32_t kf_scan(const KfLds &s, const KfUni &u, const KfTables &qp, uint32_t lo,
                                            uint32_t hi, const uint32_t *stop, uint64_t &auts, uint32_t l,
                                            uint64_t A0) {
    const uint32_t nr = u.nr, np = u.np;
    const bool fast = u.fast != 0;
    const uint32_t ident = u.ident;
    const uint64_t *tab = (fast ? qp.inv : qp.base) + qp.offset[np];
    uint32_t since = 0;
    uint64_t A = fast ? A0 : (lo + l < hi ? tab[(lo + l) / nr >> np] : 0);
    for (uint32_t base = lo; base < hi; base += 64) {
        if (stop && ++since == KF_STOP_EVERY) {
            since = 0;
            const uint32_t best = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (best < base) return KF_PRUNED;
        }
        const uint32_t r = base + l;
        const uint32_t rn = r + 64;
        const uint64_t An = rn < hi ? tab[rn / nr >> np] : 0;
        int icmp = 1;
        if (r == 0 && ident) icmp = 0;
        else if (r < hi) icmp = fast ? kf_compare_fast(s, u, A, r) : kf_compare_general(s, u, A, r);
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
    if (threadIdx.x <= KF_MAX_WINDOWS + 1) head[KF_HEAD_WORD * threadIdx.x] = 0u;   // in bounds: see KF_HEAD_WORD
    if (threadIdx.x == 0) head[KF_QUAD_WORD] = 0u;
}

__device__ __forceinline__ bool kf_meta_ok(uint32_t m, uint32_t stride) {
    const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff;
    return n >= 4 && n <= KF_MAX_PLAYERS && (n & 1) == 0 && nr >= 1 && nr <= n - 1 && nr <= stride &&
           nr <= KF_MAX_ROUNDS;
}

// Lanes of a first-look group: each candidate gets W lanes (64 / W candidates
// per wave), lane gl of a group taking rounds and players gl, gl + W, ...
constexpr uint32_t KF_LOOK_LANES = 4;

// Group staging for kf_quad_kernel: the W lanes of group q stage candidate
// q's tables into s (lane gl takes rounds and players gl, gl + W, ...; rows[j]
// is round gl + j W as loaded by the caller); the per-candidate values stay
// per lane (uniform within the group).  Every lane of the wave calls it (it
// holds the barriers); `ok` marks groups with a valid candidate.
template <uint32_t W>
__device__ __forceinline__ KfUni kf_stage_q(KfLds &s, const uint64_t *rows, uint32_t nr, uint32_t n, uint32_t gl,
                                            uint32_t q, bool ok) {
    constexpr uint32_t KR = 16 / W;
    constexpr uint64_t GM = W == 64 ? ~0ull : (1ull << W) - 1ull;
    const uint32_t sh = W * q;
    bool normal = true, matching = true;
    uint32_t glo = 0, ghi = 0;   // the group's table of player 0's partner by round
#pragma unroll
    for (uint32_t j = 0; j < KR; j++) {
        const uint32_t r = gl + j * W;
        if (ok && r < nr) {
            const uint32_t rlo = (uint32_t)rows[j], rhi = (uint32_t)(rows[j] >> 32);
            uint32_t plo = 0, phi = 0, prev = 0, seen = 0;
            for (uint32_t p = 0; p < n; p += 2) {
                const uint32_t w = p < 8 ? rlo : rhi;
                const uint32_t x = __builtin_amdgcn_ubfe(w, p << 2, 4), y = __builtin_amdgcn_ubfe(w, (p << 2) + 4, 4);
                tset(plo, phi, x, y);
                tset(plo, phi, y, x);
                normal = normal && x < y && (p == 0 || x > prev);
                prev = x;
                seen |= (1u << x) | (1u << y);
            }
            matching = matching && seen == (1u << n) - 1u;
            s.plo[r] = plo;
            s.phi[r] = phi;
            s.rlo[r] = rlo;
            s.rhi[r] = rhi;
            tset(glo, ghi, r, plo & 0xfu);
        }
    }
    for (int d = 1; d < (int)W; d <<= 1) {   // OR over the group's lanes
        glo |= (uint32_t)__shfl_xor((int)glo, d, W);
        ghi |= (uint32_t)__shfl_xor((int)ghi, d, W);
    }
    // rounds in order of player 0's partner
    bool unsorted = false;
#pragma unroll
    for (uint32_t j = 0; j < KR; j++) {
        const uint32_t r = gl + j * W;
        if (r > 0 && r < nr) unsorted = unsorted || tget(glo, ghi, r) <= tget(glo, ghi, r - 1);
    }
    const bool sorted = ((__builtin_amdgcn_ballot_w64(ok && unsorted) >> sh) & GM) == 0;
    const uint32_t fast = ((__builtin_amdgcn_ballot_w64(ok && !normal) >> sh) & GM) == 0;
    const uint64_t id = 0xfedcba9876543210ull & (n >= 16 ? ~0ull : ((1ull << (4 * n)) - 1ull));
    const uint32_t skip0 = (uint32_t)((__builtin_amdgcn_ballot_w64(ok && gl == 0 && rows[0] == id) >> sh) & 1ull);
    __syncthreads();
    KfUni u{nr, n, n / 2, fast, skip0, 0u, 1u, 0u, 0u};
    // each player's partners over the rounds: nr different players, not itself
    bool partners = true;
#pragma unroll
    for (uint32_t j = 0; j < KR; j++) {
        const uint32_t x = gl + j * W;
        if (ok && x < n) {
            uint32_t mlo = 0, mhi = 0, pm = 0;
            for (uint32_t k = 0; k < nr; k++) {
                const uint32_t y = tget(s.plo[k], s.phi[k], x);
                pm |= 1u << y;
                tset(mlo, mhi, y, 15u - k);
            }
            partners = partners && (uint32_t)__builtin_popcount(pm) == nr && !((pm >> x) & 1u);
            if (fast) s.meet[x] = ~(((uint64_t)mhi << 32) | mlo);
        }
    }
    u.valid = ((__builtin_amdgcn_ballot_w64(!matching || !partners) >> sh) & GM) == 0;
    if (ok && fast) {
        u.glo = glo;
        u.ghi = ghi;
        u.ident = skip0 && sorted;
    }
    __syncthreads();
    return u;
}

// First look, 64 / W candidates per wave: the W lanes of a group compare the
// first W relabellings of their candidate (one in a hundred of the random K_12
// partials needs more than two).  A candidate decided there (a smaller
// relabelling, or all of them tried) gets its results; the others, and those
// whose rounds are not in normal form, are listed for the first pass, which
// runs them from relabelling 0 again.
template <uint32_t W>
__global__ __launch_bounds__(64) void kf_quad_kernel(uint32_t ncand, uint32_t stride,
                                                     const uint64_t *__restrict__ rounds,
                                                     const uint32_t *__restrict__ meta, KfOut o, uint32_t *head,
                                                     KfTables qp, KfSplit sp) {
    constexpr uint32_t G = 64 / W, KR = 16 / W;
    constexpr uint64_t GM = W == 64 ? ~0ull : (1ull << W) - 1ull;
    __shared__ KfLds sg[G];
    const uint32_t l = threadIdx.x, q = l / W, gl = l % W;
    const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;
    KfLds &s = sg[q];
    // the next candidate's meta and round words are loaded while this one
    // runs (the words lie in the batch whatever the meta says)
    const uint32_t c0 = blockIdx.x * G + q;
    uint32_t mc = c0 < ncand ? meta[c0] : 0u;
    uint64_t rc[KR];
#pragma unroll
    for (uint32_t j = 0; j < KR; j++)
        rc[j] = c0 < ncand && gl + j * W < rl ? rounds[(uint64_t)c0 * stride + gl + j * W] : 0ull;
    for (uint32_t base = blockIdx.x * G; base < ncand; base += gridDim.x * G) {
        const uint32_t ci = base + q;
        const uint32_t cn = base + gridDim.x * G + q;
        const uint32_t mn = cn < ncand ? meta[cn] : 0u;
        uint64_t rn[KR];
#pragma unroll
        for (uint32_t j = 0; j < KR; j++)
            rn[j] = cn < ncand && gl + j * W < rl ? rounds[(uint64_t)cn * stride + gl + j * W] : 0ull;
        const bool active = ci < ncand;
        const uint32_t m = mc;
        const bool ok = active && kf_meta_ok(m, stride);
        if (active && !ok && gl == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
        const uint32_t nr = ok ? (m & 0xff) : 0u, n = ok ? ((m >> 8) & 0xff) : 0u, np = n / 2;
        const uint32_t total = ok ? kf_total(qp, np, nr) : 0u;
        uint64_t row[KR];
#pragma unroll
        for (uint32_t j = 0; j < KR; j++) {
            row[j] = ok ? rc[j] : 0ull;
            rc[j] = rn[j];
        }
        mc = mn;
        const uint64_t A = gl < total ? qp.inv[qp.offset[np] + (gl / nr >> np)] : 0ull;
        const KfUni u = kf_stage_q<W>(s, row, nr, n, gl, q, ok);
        int icmp = 1;
        if (gl < total && u.fast && u.valid) icmp = (gl == 0 && u.ident) ? 0 : kf_compare_fast(s, u, A, gl);
        const uint32_t sh = W * q;
        const uint64_t sm = (__builtin_amdgcn_ballot_w64(icmp < 0) >> sh) & GM;
        const uint64_t eq = (__builtin_amdgcn_ballot_w64(gl < total && icmp == 0) >> sh) & GM;
        if (ok && gl == 0) {
            if (!u.valid) {
                o.verdict[ci] = ISE_VERDICT_INVALID;
            } else if (u.fast && sm) {
                const uint32_t first = (uint32_t)__builtin_ctzll(sm);
                kf_write(o, ci, first, (uint64_t)__popcll(eq & ((1ull << first) - 1ull)), total);
            } else if (u.fast && total <= W) {
                kf_write(o, ci, KF_NONE, (uint64_t)__popcll(eq), total);
            } else {
                const uint32_t j = atomicAdd(&head[KF_QUAD_WORD], 1u);
                sp.list[1][j] = ci;
            }
        }
        __syncthreads();
    }
}

// Pass 1 (or the whole test when limit == 0).
// With in_list, the candidates are those the first look listed (head[KF_QUAD_WORD]).
__global__ __launch_bounds__(64) void kf_check_kernel(uint32_t ncand, uint32_t stride,
                                                      const uint64_t *__restrict__ rounds,
                                                      const uint32_t *__restrict__ meta, KfOut o,
                                                      uint32_t *head, KfTables qp, uint32_t limit, KfSplit sp,
                                                      const uint32_t *in_list) {
    __shared__ KfLds s;
    __shared__ uint32_t s_ci;
    const uint32_t l = threadIdx.x;
    const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;   // round words a lane may load
    const uint32_t N = in_list ? head[KF_QUAD_WORD] : ncand;
    // With a limit every candidate's time is bounded, so the candidates go by
    // a fixed grid stride and the next one is loaded while this one runs (one
    // queue word taken by every wave of the chip would serialise the short
    // tests).  Without one, the first candidate by position, the later ones
    // from the queue.
    uint32_t i = blockIdx.x;
    uint32_t ci = i < N ? (in_list ? in_list[i] : i) : 0u;
    uint32_t m = i < N ? meta[ci] : 0u;
    uint64_t row = (i < N && l < rl) ? rounds[(uint64_t)ci * stride + l] : 0ull;
    while (i < N) {
        uint32_t in = i + gridDim.x, cn = 0, mn = 0;
        uint64_t rown = 0;
        if (limit && in < N) {
            cn = in_list ? in_list[in] : in;
            mn = meta[cn];
            if (l < rl) rown = rounds[(uint64_t)cn * stride + l];
        }
        if (!kf_meta_ok(m, stride)) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
        } else {
            const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff, np = n / 2;
            const uint32_t total = kf_total(qp, np, nr);
            const uint32_t hi = (limit && limit < total) ? limit : total;
            const uint64_t A0 = l < hi ? qp.inv[qp.offset[np] + (l / nr >> np)] : 0;
            const KfUni u = kf_stage(s, row, nr, n, l);
            uint64_t auts = 0;
            const uint32_t witness = u.valid ? kf_scan(s, u, qp, 0, hi, nullptr, auts, l, A0) : KF_NONE;
            if (l == 0) {
                if (!u.valid) {
                    o.verdict[ci] = ISE_VERDICT_INVALID;
                } else if (witness == KF_NONE && hi < total) {
                    const uint32_t j = atomicAdd(&head[KF_HEAD_WORD], 1u);
                    sp.list[0][j] = ci;
                    sp.part[0][j] = auts;
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
            in = cn = s_ci;
            __syncthreads();
            if (cn < ncand) {
                mn = meta[cn];
                if (l < rl) rown = rounds[(uint64_t)cn * stride + l];
            }
        }
        i = in;
        ci = cn;
        m = mn;
        row = rown;
    }
}

// Window p: ranges of the listed candidates' relabellings in [wlo, whi).
__global__ __launch_bounds__(64) void kf_range_kernel(uint32_t stride, const uint64_t *__restrict__ rounds,
                                                      const uint32_t *__restrict__ meta, const uint32_t *head,
                                                      KfTables qp, KfSplit sp, uint32_t p, uint32_t wlo,
                                                      uint32_t whi) {
    __shared__ KfLds s;
    const uint32_t l = threadIdx.x;
    const uint32_t listed = head[KF_HEAD_WORD * (1 + p)];
    const uint32_t *list = sp.list[p & 1];
    const uint32_t G = kf_ranges(listed, sp.range_words, sp.max_ranges, whi - wlo);
    const uint32_t items = listed * G;
    // Items by a fixed grid stride: the ranges of one candidate are of one
    // size and lie on consecutive items, so they spread over the waves.  (One
    // queue word taken by every wave per item would serialise the items.)
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t j = item / G, g = item - j * G;
        const uint32_t ci = list[j];
        const uint32_t m = meta[ci];
        const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff, np = n / 2;
        const uint32_t total = kf_total(qp, np, nr);
        const uint32_t end = whi < total ? whi : total;   // listed: wlo < total
        const uint32_t per = ((end - wlo + 63u) / 64u + G - 1u) / G;
        const uint64_t lo64 = (uint64_t)wlo + (uint64_t)g * per * 64u;
        uint32_t witness = KF_NONE;
        uint64_t auts = 0;
        if (lo64 < end) {
            const uint32_t lo = (uint32_t)lo64;
            const uint32_t hi = end - lo > per * 64u ? lo + per * 64u : end;
            const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;
            const uint64_t A0 = lo + l < hi ? qp.inv[qp.offset[np] + ((lo + l) / nr >> np)] : 0;
            const KfUni u = kf_stage(s, l < rl ? rounds[(uint64_t)ci * stride + l] : 0ull, nr, n, l);
            witness = kf_scan(s, u, qp, lo, hi, &sp.best[j], auts, l, A0);
            if (l == 0 && witness < KF_PRUNED) atomicMin(&sp.best[j], witness);
        }
        if (l == 0) sp.ranges[(uint64_t)j * G + g] = (auts << 32) | witness;
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t kf_wave_sum(uint64_t x) {
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, d, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), d, 64);
        x += ((uint64_t)hi << 32) | lo;
    }
    return x;
}

constexpr uint32_t KF_COMBINE_PER_LANE = 8;

// Window p: per listed candidate (one wave each), the ranges in index order.
__global__ __launch_bounds__(64) void kf_combine_kernel(const uint32_t *__restrict__ meta, uint32_t *head,
                                                        KfTables qp, KfSplit sp, KfOut o, uint32_t p,
                                                        uint32_t wlo, uint32_t whi) {
    const uint32_t l = threadIdx.x;
    const uint32_t listed = head[KF_HEAD_WORD * (1 + p)];
    const uint32_t G = kf_ranges(listed, sp.range_words, sp.max_ranges, whi - wlo);
    const uint32_t *list = sp.list[p & 1];
    const uint64_t *part = sp.part[p & 1];
    for (uint32_t j = blockIdx.x; j < listed; j += gridDim.x) {
        uint64_t auts = 0;
        uint32_t witness = KF_NONE;
        // blocks of 64 x KF_COMBINE_PER_LANE ranges, each lane taking
        // consecutive ones (all loads of a block issued together)
        for (uint32_t g0 = 0; g0 < G; g0 += 64u * KF_COMBINE_PER_LANE) {
            uint64_t e[KF_COMBINE_PER_LANE];
#pragma unroll
            for (uint32_t i = 0; i < KF_COMBINE_PER_LANE; i++) {
                const uint32_t g = g0 + l * KF_COMBINE_PER_LANE + i;
                e[i] = g < G ? sp.ranges[(uint64_t)j * G + g] : (uint64_t)KF_NONE;
            }
            // this lane's automorphisms up to and including its first range
            // with an index, and that index
            uint64_t la = 0;
            uint32_t lw = KF_NONE;
#pragma unroll
            for (uint32_t i = 0; i < KF_COMBINE_PER_LANE; i++) {
                if (lw == KF_NONE) {
                    la += e[i] >> 32;
                    lw = (uint32_t)e[i];
                }
            }
            const uint64_t hit = __builtin_amdgcn_ballot_w64(lw != KF_NONE);
            const uint32_t upto = hit ? (uint32_t)__builtin_ctzll(hit) : 63u;
            auts += kf_wave_sum(l <= upto ? la : 0ull);
            if (hit) {
                // a real index: see the ordering argument at the top
                witness = (uint32_t)__builtin_amdgcn_readlane((int)lw, (int)upto);
                break;
            }
        }
        if (l == 0) {
            const uint32_t ci = list[j];
            const uint32_t m = meta[ci];
            const uint32_t total = kf_total(qp, ((m >> 8) & 0xff) / 2, m & 0xff);
            auts += part[j];
            if (witness != KF_NONE || whi >= total) {
                kf_write(o, ci, witness, auts, total);
            } else {
                const uint32_t k = atomicAdd(&head[KF_HEAD_WORD * (2 + p)], 1u);
                sp.list[(p + 1) & 1][k] = ci;
                sp.part[(p + 1) & 1][k] = auts;
                sp.best[k] = KF_NONE;
            }
        }
    }
}

// ---- deferred full comparisons (the U1F canonizer) -------------------------
//
// In a canonical candidate with many automorphisms (a uniform factorization
// of K_12 has 660 among 7920 two-row relabellings) nearly every chunk of 64
// relabellings holds one, and that lane compares every round while the wave waits.  So
// a first pass decides each relabelling on its first compared round only
// (smaller or larger); those equal so far are queued in index order in an LDS
// ring and compared in full 64 at a time.  Equal results come only from the
// ring, in index order, so the first smaller relabelling and the
// automorphisms before it are exactly those of the plain scan: a smaller one
// in the first pass ends the scan after the queued relabellings before it,
// and a smaller one in the ring ends it at once (everything queued after it
// has a larger index).

constexpr uint32_t KF_DEFER_SLOTS = 128;   // ring of deferred relabellings (two chunks)
constexpr uint32_t KF_FIRST_ROUNDS = 1;    // rounds compared before a relabelling is deferred
constexpr uint32_t KF_KNOWN_EQUAL = 0x80000000u;

struct KfRing {
    uint32_t r[KF_DEFER_SLOTS];     // relabelling index
    uint32_t aux[KF_DEFER_SLOTS];   // the kernel's decode of it (< 2^31); KF_KNOWN_EQUAL: equal in full
};

struct KfDeferScan {   // wave-uniform state
    uint32_t head = 0, tail = 0, witness = KF_NONE;
    uint64_t auts = 0;

    // Full comparisons of the queued relabellings while 64 are waiting
    // (all: until none is).  full(aux) -> -1 / 0 / 1.
    template <typename Full> __device__ __forceinline__ void drain(KfRing &q, uint32_t l, bool all, Full full) {
        __syncthreads();
        while (tail - head >= 64u || (all && tail != head)) {
            const uint32_t cnt = tail - head < 64u ? tail - head : 64u;
            const uint32_t slot = (head + l) % KF_DEFER_SLOTS;
            int c = 1;
            if (l < cnt) {
                const uint32_t a = q.aux[slot];
                c = (a & KF_KNOWN_EQUAL) ? 0 : full(a);
            }
            const uint64_t smaller = __builtin_amdgcn_ballot_w64(c < 0);
            const uint64_t equal = __builtin_amdgcn_ballot_w64(l < cnt && c == 0);
            if (smaller) {
                const uint32_t first = __builtin_ctzll(smaller);
                auts += __popcll(equal & ((1ull << first) - 1ull));
                witness = kf_uni(q.r[(head + first) % KF_DEFER_SLOTS]);
                return;
            }
            auts += __popcll(equal);
            head += cnt;
        }
    }

    // One chunk of the first pass: lane l holds relabelling base + l (live
    // below the end), its first-pass result icmp (-1, 0, 1 or KF_OPEN) and
    // decode aux.  Returns true when the scan has ended (witness set).
    template <typename Full>
    __device__ __forceinline__ bool chunk(KfRing &q, uint32_t l, uint32_t base, bool live, int icmp, uint32_t aux,
                                          Full full) {
        const uint64_t smaller = __builtin_amdgcn_ballot_w64(icmp < 0);
        uint64_t defer = __builtin_amdgcn_ballot_w64(live && (icmp == 0 || icmp == KF_OPEN));
        const uint32_t first = smaller ? (uint32_t)__builtin_ctzll(smaller) : 64u;
        if (smaller) defer &= (1ull << first) - 1ull;
        if ((defer >> l) & 1u) {
            const uint32_t slot = (tail + __popcll(defer & ((1ull << l) - 1ull))) % KF_DEFER_SLOTS;
            q.r[slot] = base + l;
            q.aux[slot] = (icmp == 0 ? KF_KNOWN_EQUAL : 0u) | aux;
        }
        tail += __popcll(defer);
        if (smaller) {
            drain(q, l, true, full);
            if (witness == KF_NONE) witness = base + first;
            return true;
        }
        if (tail - head >= 64u) drain(q, l, false, full);
        return witness != KF_NONE;
    }

    // After the last chunk: the queued rest.
    template <typename Full> __device__ __forceinline__ void finish(KfRing &q, uint32_t l, Full full) {
        if (witness == KF_NONE) drain(q, l, true, full);
    }
};

// ---- two-row canonizer of perfect 1-factorizations ------------------------
//
// alldata::cnvCheck2P1F (k-Factorization/TripleSys/Source/cnvSupport.cpp:3-95)
// with create2P1FTr (p1fSupport.cpp:3-38) and kmProcessMatrix2p1f
// (kmSupport.cpp:519-598): in a perfect 1-factorization every pair of rounds
// is one Hamiltonian cycle, so a relabelling that could give a matrix no
// larger than the given one must send some ordered pair of rounds (r0, r1)
// onto rounds 0, 1; walking the two cycles side by side from a start player k
// and from player 0 fixes it.  Relabelling index r = ((pair * 2 + swap) * n +
// k) with pair = last * (last - 1) / 2 + ind over last = 1.., ind < last, and
// (r0, r1) = (ind, last) or, swapped, (last, ind).  Rounds r0, r1 become rounds
// 0, 1 and are not compared; the relabelled round with partner-of-0 label v
// (v = 3..nr) is compared with round v - 1 as given, in order of v, a missing
// v making the relabelling not smaller.  Fewer than 3 rounds: every
// relabelling is equal.  Rounds must be in normal form and every pair of them
// one Hamiltonian cycle; otherwise the record is invalid.

// Relabelling (ordered pair po of rounds r0 | r1 << 4 = rr, start player k)
// against the matrix: -1 smaller, 0 equal, 1 not smaller.
__device__ __forceinline__ int kf2_compare(const KfLds &s, const KfUni &u, uint32_t rr, uint32_t k, uint32_t z0lo,
                                           uint32_t z0hi, uint32_t z1lo, uint32_t z1hi) {
    const uint32_t n = u.n, nr = u.nr;
    const uint32_t r0 = rr & 0xfu, r1 = rr >> 4;
    const uint32_t alo = s.plo[r0], ahi = s.phi[r0], blo = s.plo[r1], bhi = s.phi[r1];
    // label -> player: walk round r0 / r1 from k against round 0 / 1 from player 0
    uint32_t vlo = 0, vhi = 0, v1 = k, it = 0;
    for (uint32_t st = 0; st < n; st += 2) {
        tset(vlo, vhi, it, v1);
        const uint32_t v2 = tget(alo, ahi, v1);
        it = tget(z0lo, z0hi, it);
        tset(vlo, vhi, it, v2);
        it = tget(z1lo, z1hi, it);
        v1 = tget(blo, bhi, v2);
    }
    if (nr < 3) return 0;
    const uint64_t mt = s.meet[k];
    const uint32_t elo = (uint32_t)mt, ehi = (uint32_t)(mt >> 32);
    for (uint32_t p = 2; p < nr; p++) {
        const uint32_t L = p + 1;
        const uint32_t m = tget(elo, ehi, tget(vlo, vhi, L));
        if (m == 15u) return 1;
        const uint32_t g = tget(u.glo, u.ghi, p);
        if (L != g) return L < g ? -1 : 1;
        // given pairs (a, gv) after (0, L), as in kf_compare_labels
        const uint32_t klo = s.plo[m], khi = s.phi[m];
        const uint32_t qlo = s.rlo[p], qhi = s.rhi[p];
        for (uint32_t j = 1; j < u.np; j++) {
            const uint32_t pr = __builtin_amdgcn_ubfe(j < 4 ? qlo : qhi, (j & 3u) << 3, 8);
            const uint32_t a = pr & 0xfu, gv = pr >> 4;
            const uint32_t z = tget(klo, khi, tget(vlo, vhi, a));   // player whose label is a's partner
            if (z != tget(vlo, vhi, gv)) return tfind(vlo, vhi, z) < gv ? -1 : 1;
        }
    }
    return 0;
}

__global__ __launch_bounds__(64) void kf2_check_kernel(uint32_t ncand, uint32_t stride,
                                                       const uint64_t *__restrict__ rounds,
                                                       const uint32_t *__restrict__ meta, KfOut o) {
    __shared__ KfLds s;
    __shared__ uint32_t s_rr[KF_MAX_ROUNDS * (KF_MAX_ROUNDS - 1)];   // ordered pair: r0 | r1 << 4
    const uint32_t l = threadIdx.x;
    const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;
    // at most 15 * 14 * 16 relabellings per candidate: a fixed grid stride
    for (uint32_t ci = blockIdx.x; ci < ncand; ci += gridDim.x) {
        const uint32_t m = meta[ci];
        if (!kf_meta_ok(m, stride)) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            continue;
        }
        const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff;
        const uint64_t row = l < rl ? rounds[(uint64_t)ci * stride + l] : 0ull;
        const KfUni u = kf_stage(s, row, nr, n, l);
        // every round a perfect matching of players 0..n-1 (device records
        // come unchecked; staging checks it) and every pair of rounds one
        // Hamiltonian cycle
        bool bad = !u.fast || !u.valid;
        const uint32_t npairs = nr * (nr - 1) / 2;
        for (uint32_t b0 = 0; b0 < npairs && !bad; b0 += 64) {
            bool ham = true;
            const uint32_t pi = b0 + l;
            if (pi < npairs) {
                uint32_t b = 1;
                while ((b + 1) * b / 2 <= pi) b++;
                const uint32_t a = pi - b * (b - 1) / 2;
                s_rr[2 * pi] = a | (b << 4);
                s_rr[2 * pi + 1] = b | (a << 4);
                const uint32_t alo = s.plo[a], ahi = s.phi[a], blo = s.plo[b], bhi = s.phi[b];
                uint32_t v = 0, len = 0;
                do {
                    v = tget(blo, bhi, tget(alo, ahi, v));
                    len += 2;
                } while (v != 0 && len <= n);
                ham = len == n;
            }
            bad = __builtin_amdgcn_ballot_w64(!ham) != 0;
        }
        if (bad) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            __syncthreads();
            continue;
        }
        const uint32_t z0lo = kf_uni(s.plo[0]), z0hi = kf_uni(s.phi[0]);
        const uint32_t z1lo = nr > 1 ? kf_uni(s.plo[1]) : 0u, z1hi = nr > 1 ? kf_uni(s.phi[1]) : 0u;
        // (Few automorphisms here: the deferral of the U1F kernel does not pay.)
        __syncthreads();
        const uint32_t total = nr * (nr - 1) * n;
        uint64_t auts = 0;
        uint32_t witness = KF_NONE;
        // lane l starts at relabelling l: ordered pair l / n, start player l % n
        uint32_t po = l / n, k = l - po * n;
        const uint32_t sq = 64u / n, sr = 64u - sq * n;
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t r = base + l;
            const int icmp = r < total ? kf2_compare(s, u, s_rr[po], k, z0lo, z0hi, z1lo, z1hi) : 1;
            const uint64_t smaller = __builtin_amdgcn_ballot_w64(icmp < 0);
            const uint64_t equal = __builtin_amdgcn_ballot_w64(r < total && icmp == 0);
            if (smaller) {
                const uint32_t first = __builtin_ctzll(smaller);
                auts += __popcll(equal & ((1ull << first) - 1ull));
                witness = base + first;
                break;
            }
            auts += __popcll(equal);
            k += sr;
            po += sq;
            if (k >= n) {
                k -= n;
                po++;
            }
        }
        if (l == 0) kf_write(o, ci, witness, auts, total);
        __syncthreads();
    }
}

// ---- two-row canonizer of uniform 1-factorizations ------------------------
//
// alldata::cnvCheck2U1F (k-Factorization/TripleSys/Source/cnvSupport2.cpp:3-142)
// with getCyclesAndPathFromNeighbors / sortCycles (p1fCheck.cpp:65-76,
// 153-228), CycleMapping (cycles.cpp:111-211) and createU1FTr
// (TrCycles.cpp:3-50).  The cycles of a pair of rounds (r0, r1) start at the
// smallest player not yet met and go r0-partner first; they are stably sorted
// by length.  A relabelling maps the cycles of (r0, r1) onto those of rounds
// 0, 1: cycles of one length in some order (permutations of each group of
// equal lengths, lexicographic rank, first group fastest), a start offset d
// per cycle (L/2 choices: players 2d of the cycle), a direction per cycle;
// cycle c of rounds 0, 1 with players P[0..L) receives Q[(2d + m) mod L] or,
// reversed, Q[(2d + 1 - m) mod L] at label P[m].  Index within a pair:
// (permutation rank * V + offsets) * 2^C + directions, with directions as a
// binary number whose last cycle is the lowest bit and offsets as a mixed-radix
// number whose first cycle is the lowest digit (V = product of L/2).  (The
// reference's enumeration keeps the start order of a group that wraps while a
// later group advances; with at most four cycles of at least four players no
// two groups of n <= 16 both have two cycles or more, so ranks map one to one.)
// A pair of another cycle type than rounds 0, 1 ends the test: not canonical,
// no witness, the relabellings before it tried.  Rounds must be in normal form.
//
// Layout on the device.  Each ordered pair po (unordered pair po >> 1 in
// order, then the swap bit) has its cycles staged once per candidate as a
// path: the players of its sorted cycles, each from its start going r0-partner
// first, 4 bits each (path[po]).  Swapping r0 and r1 keeps the cycles, their
// order and their starts and reverses each walk: Q'[j] = Q[(L - j) mod L].
// So "reversed from 2d + 1" is the swapped pair's path rotated by L - 2d - 1,
// and "forward from 2d" the pair's own path rotated by 2d: the label -> player
// map of a relabelling is a rotation of one segment per cycle, scattered by
// the players of rounds 0, 1 in their path order.  Lanes step through the
// relabellings in place (pair po, position t) and decode t with
// multiply-high by per-candidate magic numbers (all digits are below 2^13).

struct KfU1f {
    uint32_t ncyc, T, cut;            // cycles, relabellings per pair, relabellings before a mismatched pair
    uint32_t len01, off01;            // rounds 0, 1: half length and first path position of each cycle (4 bits each)
    uint32_t g0, gm;                  // the one group of equal cycles that permutes: first cycle, size (1: none)
    uint32_t V, mV, mh[4];            // product of the half lengths; magic numbers for V and each half length
    uint32_t plo01, phi01;            // rounds 0, 1: players in path order
};

struct KfU1fLds {
    uint64_t path[KF_MAX_ROUNDS * (KF_MAX_ROUNDS - 1)];   // ordered pair: players of its cycles in path order
    uint32_t rr[KF_MAX_ROUNDS * (KF_MAX_ROUNDS - 1)];     // ordered pair: r0 | r1 << 4
    uint32_t perm[24];                                     // lexicographic orders of the permuting group
    uint32_t bad;                                          // least unordered pair of another cycle type
    KfRing q;                                              // deferred relabellings: aux = po << 16 | t
};

// ceil(2^32 / d) for 2 <= d < 2^13: __umulhi(x, kf_magic(d)) = x / d for x < 2^19
__device__ __forceinline__ uint32_t kf_magic(uint32_t d) { return 0xffffffffu / d + 1u; }

// Cycles of rounds a, b (partner tables alo:ahi, blo:bhi): starts (4 bits per
// cycle, sorted order) and half lengths likewise; returns the cycle count
// (at most four when the rounds share no pair).
__device__ __forceinline__ uint32_t kf_cycles(uint32_t alo, uint32_t ahi, uint32_t blo, uint32_t bhi, uint32_t n,
                                              uint32_t &starts, uint32_t &halves) {
    uint32_t seen = 0, nc = 0;
    starts = halves = 0;
    for (uint32_t k0 = 0; k0 < n; k0++) {
        if ((seen >> k0) & 1u) continue;
        uint32_t x = k0, len = 0;
        do {
            const uint32_t y = tget(alo, ahi, x);
            seen |= (1u << x) | (1u << y);
            len += 2;
            x = tget(blo, bhi, y);
        } while (x != k0 && len < 2 * KF_MAX_PLAYERS);   // bounded also for tables that are not matchings
        // insert after every cycle no longer than this one (stable sort by length)
        uint32_t pos = nc;
        while (pos > 0 && ((halves >> (4 * (pos - 1))) & 0xfu) > len / 2) pos--;
        const uint32_t lowm = (1u << (4 * pos)) - 1u;
        starts = (starts & lowm) | ((starts & ~lowm) << 4) | (k0 << (4 * pos));
        halves = (halves & lowm) | ((halves & ~lowm) << 4) | ((len / 2) << (4 * pos));
        nc++;
    }
    return nc;
}

// The players of the sorted cycles of rounds a, b in path order: each cycle
// from its start, a-partner first.
__device__ __forceinline__ uint64_t kf_path(uint32_t alo, uint32_t ahi, uint32_t blo, uint32_t bhi, uint32_t nc,
                                            uint32_t starts, uint32_t halves) {
    uint64_t path = 0;
    uint32_t pos = 0;
    for (uint32_t c = 0; c < nc; c++) {
        uint32_t q = (starts >> (4 * c)) & 0xfu;
        const uint32_t L = 2 * ((halves >> (4 * c)) & 0xfu);
        for (uint32_t j = 0; j < L; j++, pos++) {
            path |= (uint64_t)q << (4 * pos);
            q = (j & 1u) ? tget(blo, bhi, q) : tget(alo, ahi, q);
        }
    }
    return path;
}

// Relabelling (po, t): -1 smaller, 0 equal, 1 larger (KF_OPEN: equal in the
// first `lim` compared rounds).
__device__ __forceinline__ int kf_u1f_compare(const KfLds &s, const KfU1fLds &x, const KfUni &u, const KfU1f &w,
                                              uint32_t po, uint32_t t, uint32_t lim) {
    const uint32_t C = w.ncyc;
    const uint64_t qf = x.path[po], qb = x.path[po ^ 1u];
    const uint32_t rr = x.rr[po];
    const uint32_t dirs = t & ((1u << C) - 1u), rest = t >> C;
    const uint32_t rank = __umulhi(rest, w.mV);
    uint32_t var = rest - rank * w.V;
    const uint32_t pw = w.gm > 1 ? x.perm[rank] : 0u;
    uint64_t R = 0;   // players in the path order of rounds 0, 1
#pragma unroll
    for (uint32_t c = 0; c < 4; c++) {
        if (c >= C) break;
        const uint32_t h = (w.len01 >> (4 * c)) & 0xfu, L = 2 * h;
        const uint32_t q = __umulhi(var, w.mh[c]);
        const uint32_t d = var - q * h;
        var = q;
        const bool back = (dirs >> (C - 1 - c)) & 1u;
        const uint32_t e = c - w.g0 < w.gm ? w.g0 + ((pw >> (4 * (c - w.g0))) & 0xfu) : c;
        const uint32_t offe = (w.off01 >> (4 * e)) & 0xfu, offc = (w.off01 >> (4 * c)) & 0xfu;
        const uint64_t mask = L == 16 ? ~0ull : (1ull << (4 * L)) - 1ull;
        const uint64_t seg = ((back ? qb : qf) >> (4 * offe)) & mask;
        const uint32_t rot = back ? L - 2 * d - 1 : 2 * d;
        const uint64_t segr = rot ? ((seg >> (4 * rot)) | (seg << (4 * (L - rot)))) & mask : seg;
        R |= segr << (4 * offc);
    }
    const uint32_t Rlo = (uint32_t)R, Rhi = (uint32_t)(R >> 32);
    uint32_t vlo = 0, vhi = 0;   // label -> player
    for (uint32_t m = 0; m < u.n; m++)
        tset(vlo, vhi, tget(w.plo01, w.phi01, m), tget(Rlo, Rhi, m));
    return kf_compare_labels(s, u, vlo, vhi, rr & 0xfu, rr >> 4, lim);
}

__global__ __launch_bounds__(64) void kf_u1f_kernel(uint32_t ncand, uint32_t stride,
                                                    const uint64_t *__restrict__ rounds,
                                                    const uint32_t *__restrict__ meta, KfOut o, KfU1fMode md) {
    __shared__ KfLds s;
    __shared__ KfU1fLds x;
    const uint32_t l = threadIdx.x;
    const uint32_t rl = stride < KF_MAX_ROUNDS ? stride : KF_MAX_ROUNDS;
    for (uint32_t ci = blockIdx.x; ci < ncand; ci += gridDim.x) {
        const uint32_t m = meta[ci];
        if (!kf_meta_ok(m, stride)) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            continue;
        }
        const uint32_t nr = m & 0xff, n = (m >> 8) & 0xff;
        const uint64_t row = l < rl ? rounds[(uint64_t)ci * stride + l] : 0ull;
        const KfUni u = kf_stage(s, row, nr, n, l);
        // Every round a perfect matching of players 0..n-1 and no pair in two
        // rounds (what the host entry points check; device records come
        // unchecked, staging checks them): only then are the cycles below at
        // most four, of at least four players each.
        if (!u.fast || !u.valid) {
            if (l == 0) o.verdict[ci] = ISE_VERDICT_INVALID;
            __syncthreads();
            continue;
        }
        KfU1f w{};
        w.T = 1u;
        uint32_t total = 0;
        if (nr > 1) {
            const uint32_t z0lo = kf_uni(s.plo[0]), z0hi = kf_uni(s.phi[0]);
            const uint32_t z1lo = kf_uni(s.plo[1]), z1hi = kf_uni(s.phi[1]);
            uint32_t st01, h01;
            w.ncyc = kf_uni(kf_cycles(z0lo, z0hi, z1lo, z1hi, n, st01, h01));
            w.len01 = kf_uni(h01);
            const uint64_t p01 = kf_path(z0lo, z0hi, z1lo, z1hi, w.ncyc, kf_uni(st01), w.len01);
            w.plo01 = kf_uni((uint32_t)p01);
            w.phi01 = kf_uni((uint32_t)(p01 >> 32));
            uint32_t T = 1, V = 1, off = 0;
            w.g0 = 0;
            w.gm = 1;
            for (uint32_t c = 0, g = 0; c < w.ncyc; c = g) {
                const uint32_t h = (w.len01 >> (4 * c)) & 0xfu;
                for (g = c; g < w.ncyc && ((w.len01 >> (4 * g)) & 0xfu) == h; g++) {
                    T *= 2 * h;
                    V *= h;
                    w.off01 |= off << (4 * g);
                    off += 2 * h;
                }
                for (uint32_t f = 2; f <= g - c; f++) T *= f;
                if (g - c > 1) {
                    w.g0 = c;
                    w.gm = g - c;
                }
            }
            w.T = T;
            w.V = V;
            w.mV = kf_magic(V);
#pragma unroll
            for (uint32_t c = 0; c < 4; c++) w.mh[c] = c < w.ncyc ? kf_magic((w.len01 >> (4 * c)) & 0xfu) : 0u;
            // the orders of the permuting group, lexicographic
            uint32_t fact = 1;
            for (uint32_t f = 2; f <= w.gm; f++) fact *= f;
            if (w.gm > 1 && l < fact) {
                uint32_t rk = l, fr = fact, avail = (1u << w.gm) - 1u, word = 0;
                for (uint32_t j = 0; j < w.gm; j++) {
                    fr /= (w.gm - j);
                    uint32_t pick = rk / fr;
                    rk %= fr;
                    uint32_t bits = avail;
                    while (pick--) bits &= bits - 1u;
                    const uint32_t e = (uint32_t)__builtin_ctz(bits);
                    avail &= ~(1u << e);
                    word |= e << (4 * j);
                }
                x.perm[l] = word;
            }
            // every ordered pair's path, and the first unordered pair (in
            // order) of another cycle type
            if (l == 0) x.bad = 0xffffffffu;
            __syncthreads();
            const uint32_t nord = nr * (nr - 1);
            for (uint32_t b0 = 0; b0 < nord; b0 += 64) {
                const uint32_t po = b0 + l;
                if (po < nord) {
                    const uint32_t pi = po >> 1;
                    uint32_t last = 1;
                    while ((last + 1) * last / 2 <= pi) last++;
                    const uint32_t ind = pi - last * (last - 1) / 2;
                    const uint32_t r0 = (po & 1) ? last : ind, r1 = (po & 1) ? ind : last;
                    const uint32_t alo = s.plo[r0], ahi = s.phi[r0], blo = s.plo[r1], bhi = s.phi[r1];
                    uint32_t st, h;
                    const uint32_t nc = kf_cycles(alo, ahi, blo, bhi, n, st, h);
                    const bool same = nc == w.ncyc && h == w.len01;
                    if (!same) atomicMin(&x.bad, pi);
                    x.path[po] = same ? kf_path(alo, ahi, blo, bhi, nc, st, h) : 0ull;
                    x.rr[po] = r0 | (r1 << 4);
                }
            }
            __syncthreads();
            const uint32_t bad = kf_uni(x.bad);
            total = nr * (nr - 1) * T;
            w.cut = bad == 0xffffffffu ? total : 2u * bad * T;
        }
        if (md.reps > 1 && nr > 1) {
            // cnvCheck3U1F: rounds 0, 1 of a type the run does not allow -- not canonical
            // before any relabelling (cyclesFor2Rows collects no cycle set)
            bool allowed = md.ntypes == 0 && w.ncyc == 1 && 2u * (w.len01 & 0xfu) == n;
            for (uint32_t k = 0; k < md.ntypes && k < KF_U1F_MAX_TYPES; k++)
                allowed = allowed || md.types[k] == ((w.ncyc << 16) | w.len01);
            if (!allowed) {
                if (l == 0) {
                    o.verdict[ci] = ISE_VERDICT_NONCANONICAL;
                    if (o.witness) o.witness[ci] = ~0ull;
                    if (o.n_auts) o.n_auts[ci] = 0ull;
                    if (o.n_tried) o.n_tried[ci] = 0ull;
                }
                __syncthreads();
                continue;
            }
        }
        const auto full = [&](uint32_t a) {
            return kf_u1f_compare(s, x, u, w, a >> 16, a & 0xffffu, 0xffffffffu);
        };
        KfDeferScan sc;
        // lane l starts at relabelling l: ordered pair l / T, position l % T
        const uint32_t mT = w.T > 1 ? kf_magic(w.T) : 0u;
        uint32_t po = w.T > 1 ? __umulhi(l, mT) : l;
        uint32_t t = l - po * w.T;
        const uint32_t sq = w.T > 1 ? __umulhi(64u, mT) : 64u, sr = 64u - sq * w.T;
        for (uint32_t base = 0; base < w.cut; base += 64) {
            const uint32_t r = base + l;
            const int icmp = r < w.cut ? kf_u1f_compare(s, x, u, w, po, t, KF_FIRST_ROUNDS) : 1;
            if (sc.chunk(x.q, l, base, r < w.cut, icmp, (po << 16) | t, full)) break;
            t += sr;
            po += sq;
            if (t >= w.T) {
                t -= w.T;
                po++;
            }
        }
        sc.finish(x.q, l, full);
        const uint32_t witness = sc.witness;
        const uint64_t auts = sc.auts;
        if (l == 0) {
            // cnvCheck3U1F tries every pair's relabellings md.reps times over (its loop over
            // MAX_CYCLE_SETS cycle sets, one set for groups of 2): index ((ordered pair * reps) +
            // repetition) * T + position, a witness met in the first repetition
            const uint64_t R = md.reps;
            if (witness == KF_NONE && w.cut < total) {   // a pair of another cycle type
                o.verdict[ci] = ISE_VERDICT_NONCANONICAL;
                if (o.witness) o.witness[ci] = ~0ull;
                if (o.n_auts) o.n_auts[ci] = auts;
                if (o.n_tried) o.n_tried[ci] = (uint64_t)w.cut * R;
            } else if (R == 1) {
                kf_write(o, ci, witness, auts, total);
            } else if (witness == KF_NONE) {
                kf_write(o, ci, witness, auts, 0u);
                if (o.n_tried) o.n_tried[ci] = (uint64_t)total * R;
            } else {
                const uint32_t po = witness / w.T, t = witness - po * w.T;
                const uint64_t r3 = (uint64_t)po * R * w.T + t;
                o.verdict[ci] = ISE_VERDICT_NONCANONICAL;
                if (o.witness) o.witness[ci] = r3;
                if (o.n_auts) o.n_auts[ci] = auts;
                if (o.n_tried) o.n_tried[ci] = r3 + 1;
            }
        }
        __syncthreads();
    }
}

hipError_t kf_u1f_launch(uint32_t blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                         const uint32_t *meta, const KfOut &o, const KfU1fMode &md, hipStream_t stream) {
    hipLaunchKernelGGL(kf_u1f_kernel, dim3(blocks), dim3(64), 0, stream, ncand, stride, rounds, meta, o, md);
    return hipGetLastError();
}

hipError_t kf2_launch(uint32_t blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                      const uint32_t *meta, const KfOut &o, hipStream_t stream) {
    hipLaunchKernelGGL(kf2_check_kernel, dim3(blocks), dim3(64), 0, stream, ncand, stride, rounds, meta, o);
    return hipGetLastError();
}

hipError_t kf_launch(uint32_t blocks, uint32_t range_blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                     const uint32_t *meta, const KfOut &o, uint32_t *head, const KfTables &qp,
                     uint32_t limit, const KfSplit &sp, hipStream_t stream) {
    hipLaunchKernelGGL(kf_queue_reset_kernel, dim3(1), dim3(64), 0, stream, head);
    if (limit) {
        constexpr uint32_t G = 64 / KF_LOOK_LANES;
        const uint32_t need = (ncand + G - 1u) / G;
        const uint32_t qblocks = need < range_blocks ? need : range_blocks;
        hipLaunchKernelGGL(kf_quad_kernel<KF_LOOK_LANES>, dim3(qblocks ? qblocks : 1u), dim3(64), 0, stream, ncand,
                           stride, rounds, meta, o, head, qp, sp);
    }
    hipLaunchKernelGGL(kf_check_kernel, dim3(blocks), dim3(64), 0, stream, ncand, stride, rounds, meta, o,
                       head, qp, limit, sp, limit ? (const uint32_t *)sp.list[1] : nullptr);
    // Windows [wlo, 16 wlo) from the first-pass limit; the one that starts at
    // KF_LAST_WINDOW or later (or is the last the head words allow) runs to
    // the end of every candidate's index space, so every listed candidate is
    // decided, and no launch is spent on windows past the sizes that need
    // them (K_12: < 2^20 relabellings; K_16, 15 rounds: 40320 * 256 * 15).
    uint64_t wlo = limit;
    for (uint32_t p = 0; limit && p < KF_MAX_WINDOWS; p++) {
        const bool last = wlo >= KF_LAST_WINDOW || p + 1 == KF_MAX_WINDOWS;
        const uint64_t whi = last ? 0xffffffffull : wlo * KF_WINDOW_GROWTH;
        hipLaunchKernelGGL(kf_range_kernel, dim3(range_blocks), dim3(64), 0, stream, stride, rounds, meta,
                           (const uint32_t *)head, qp, sp, p, (uint32_t)wlo, (uint32_t)whi);
        hipLaunchKernelGGL(kf_combine_kernel, dim3(1024), dim3(64), 0, stream, meta, head, qp, sp, o, p,
                           (uint32_t)wlo, (uint32_t)whi);
        if (last) break;
        wlo = whi;
    }
    return hipGetLastError();
}

hipError_t kf_occupancy(int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)&kf_check_kernel, 64, 0);
}

} // namespace ise
