// ise_wave.h -- lane-parallel canonicity test of one bit-packed candidate.
//
// One group of G lanes (G = 16, 32 or 64; 64/G candidates per wavefront) owns
// one candidate partial incidence matrix.  Every piece of per-candidate state
// lives in the group's lanes:
//   lane l :  rows[l]   original row l as a column bit mask   (l < depth)
//             cells[l]  column-orbit start mask of row l     (l < depth)
//             perm[l]   image of row l under the current row permutation
//             orb[l]    row orbit label (smallest element of the orbit)
//             pcol[l]   column permutation, position l -> source column (l < ncols)
// Control flow is uniform inside a group; cross-lane work uses ballots,
// readlane/bpermute and ds_permute.
//
// The search replays the reference's scalar algorithm step for step so that
// the verdict, the first witness, the automorphisms found, |Aut| and the
// final permutation state are identical to it:
//   - the row-permutation walk of CGroupOrder::next_permutation
//     (k-Factorization/Utils/Include/GroupOrder.h:136-248),
//   - the row-by-row / orbit-by-orbit comparison of
//     CCanonicityChecker::TestCanonicity (Source/CanonicityChecker.h:326-477)
//     and checkColOrbit (Source/CanonicityChecker.cpp:191-223): a row of the
//     permuted matrix is compared with the original one column orbit at a
//     time by the number of ones; the first orbit that differs decides, and
//     the column permutation is stably partitioned (ones first) inside every
//     orbit that compared equal before it,
//   - addAutomorphism / updateGroupOrder (GroupOrder.h:56-102, :30-49).
#pragma once
#include <stdint.h>
#include "ise.h"

#ifndef ISE_DEV
#define ISE_DEV __device__ __forceinline__
#endif

namespace ise {

// ---- cross-lane primitives (CDNA wave64) ---------------------------------
ISE_DEV int lane_id() { return (int)__lane_id(); }
ISE_DEV uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }
ISE_DEV int pull(int x, int src_lane) { return __builtin_amdgcn_ds_bpermute(src_lane << 2, x); }
ISE_DEV int push(int x, int dst_lane) { return __builtin_amdgcn_ds_permute(dst_lane << 2, x); }
ISE_DEV int popc64(uint64_t x) { return __popcll(x); }
ISE_DEV int ctz64(uint64_t x) { return __builtin_ctzll(x); }
ISE_DEV int hib64(uint64_t x) { return 63 - __builtin_clzll(x); }
ISE_DEV uint64_t lowmask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

// Group-local views: G lanes starting at `base`.
template <int G> struct Grp {
    int ln;     // lane index inside the group
    int base;   // first wave lane of the group
    ISE_DEV Grp() { int l = lane_id(); ln = l & (G - 1); base = l - ln; }
    ISE_DEV uint64_t ballot(bool p) const {
        uint64_t m = ballot64(p);
        if constexpr (G == 64) return m;
        else return (m >> base) & ((1ull << G) - 1);
    }
    // value of x held by group lane i (i uniform in the group)
    ISE_DEV int read(int x, int i) const {
        if constexpr (G == 64) return __builtin_amdgcn_readlane(x, i);
        else return pull(x, base + i);
    }
    ISE_DEV uint64_t read64(uint64_t x, int i) const {
        uint32_t lo = (uint32_t)read((int)(uint32_t)x, i);
        uint32_t hi = (uint32_t)read((int)(uint32_t)(x >> 32), i);
        return ((uint64_t)hi << 32) | lo;
    }
    // A value the whole group holds (in a full wave: pinned to a scalar
    // register, so that loops driven by it compile to scalar branches).
    ISE_DEV int uni(int x) const {
        if constexpr (G == 64) return __builtin_amdgcn_readfirstlane(x);
        else return x;
    }
    ISE_DEV uint64_t uni64(uint64_t x) const {
        if constexpr (G == 64)
            return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32)) << 32) |
                   (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
        else return x;
    }
    ISE_DEV int gather(int x, int i) const { return pull(x, base + i); }
    ISE_DEV int scatter(int x, int i) const { return push(x, base + i); }
};

constexpr int EMAX = 255;   // ELEMENT_MAX for 8-bit indices
constexpr int IMAX = 254;   // IDX_MAX

// Row-permutation state of one candidate.
struct PermState {
    int stab;        // stabiliser length (m_nStabLength)
    int stab_aut;    // m_nStabLengthAut
    uint64_t order;  // |Aut| accumulated along the stabiliser chain, bits 0..63
    uint64_t order_hi;   // bits 64..127 (the reference's group order is a 128-bit UInt)
};

// perm[piv+1 .. n-1] reversed
template <int G> ISE_DEV int reverse_tail(const Grp<G> &g, int pr, int n, int piv) {
    int l = g.ln;
    int src = (l > piv && l < n) ? (n + piv - l) : l;
    return g.gather(pr, src);
}

// Next row permutation fixing the first `fixed` rows (GroupOrder.h:136-248).
// `at`: EMAX plain successor, IMAX after an automorphism, else refuted prefix
// position.  Returns the first changed position or EMAX.
template <int G>
ISE_DEV int next_row_perm(const Grp<G> &g, PermState &st, int &pr, int ob, int n, int at, int fixed) {
    const int l = g.ln;
    for (;;) {
        int piv, succ, old;
        if (at == IMAX && g.read(pr, st.stab) == n - 1)
            at = EMAX;

        if (at == IMAX) {
            piv = st.stab;
            old = g.read(pr, piv);
            at = piv;
            // positions piv+1..old take piv..old-1, positions old+1..n-1 keep their index
            if (l > piv && l <= old) pr = l - 1;
            else if (l > old && l < n) pr = l;
            succ = old + 1;
        } else if (at == EMAX) {
            int prev = g.gather(pr, l > 0 ? l - 1 : 0);
            uint64_t asc = g.ballot(l >= 1 && l < n && prev < pr);
            piv = asc ? hib64(asc) : 0;
            if (piv == fixed)
                return EMAX;
            piv -= 1;
            old = g.read(pr, piv);
            uint64_t big = g.ballot(l > piv && l < n && pr > old);
            succ = hib64(big);
        } else {
            piv = at;
            old = g.read(pr, piv);
            uint64_t big = g.ballot(l > piv && l < n && pr > old);
            if (!big) {
                pr = reverse_tail(g, pr, n, piv);
                at = EMAX;
                fixed = 0;
                continue;   // the reference recurses with default arguments
            }
            succ = ctz64(big);
        }

        if (st.stab == piv) {
            // the image of the stabiliser point must be an orbit leader
            bool step_back = false;
            int img = g.read(pr, succ);
            bool leader = g.gather(ob, pr) == pr;
            if (at >= IMAX) {
                uint64_t lead = g.ballot(leader && l > piv && l <= succ);
                int q = lead ? hib64(lead) : piv;
                if (q != succ) {
                    if (q == 0)
                        return EMAX;
                    step_back = (q == piv);
                    img = g.read(pr, q);
                    int src = (l >= q && l < succ) ? l + 1 : l;
                    pr = g.gather(pr, src);
                }
            } else {
                uint64_t lead = g.ballot(leader && l >= succ && l < n);
                int q = lead ? ctz64(lead) : n;
                if (q != succ) {
                    step_back = (q == n);
                    if (step_back) {
                        if (piv == 0)
                            return EMAX;
                        if (l >= at && l < succ) pr = l;
                    } else {
                        img = g.read(pr, q);
                        int src = (l > succ && l <= q) ? l - 1 : l;
                        pr = g.gather(pr, src);
                    }
                }
            }
            if (l == succ) pr = img;
            if (step_back) {
                succ = at >= IMAX ? n - 1 : piv;
                piv -= 1;
                old = g.read(pr, piv);
                st.stab = piv;
            }
        }

        int at_succ = g.read(pr, succ);
        if (l == piv) pr = at_succ;
        if (l == succ) pr = old;
        if (at >= IMAX) {
            if (st.stab > piv)
                st.stab = piv;
            pr = reverse_tail(g, pr, n, piv);
        }
        return piv;
    }
}

// |Aut| *= length of the orbit of the current stabiliser point (GroupOrder.h:63-73)
template <int G> ISE_DEV void grow_order(const Grp<G> &g, PermState &st, int ob, int n) {
    if (st.stab_aut >= n) return;
    uint64_t m = g.ballot(g.ln > st.stab_aut && g.ln < n && ob == st.stab_aut);
    const uint64_t len = (uint64_t)(1 + popc64(m));
    st.order_hi = st.order_hi * len + __umul64hi(st.order, len);
    st.order *= len;
}

// New automorphism `pr` (GroupOrder.h:56-62, :76-102, Update_Orbits :30-49)
template <int G>
ISE_DEV void add_automorphism(const Grp<G> &g, PermState &st, int pr, int &ob, int n, bool row_permut) {
    const int l = g.ln;
    uint64_t moved = g.ballot(l < n && pr != l);
    int first = ctz64(moved);
    if (row_permut && st.stab_aut > first)
        grow_order(g, st, ob, n);
    st.stab_aut = first;
    st.stab = first;
    for (int p = first; p < n; p++) {
        int a = g.read(ob, p);
        int c = g.read(ob, g.read(pr, p));
        if (a == c) continue;
        if (c < a) { int t = a; a = c; c = t; }
        if (ob == c) ob = a;
    }
}

// Outputs of one batch (device pointers, structure of arrays).
struct OutPtrs {
    uint8_t *verdict, *fail_row, *fail_col, *row_to_change;
    uint64_t *group_order;
    uint64_t *group_order_hi;   // bits 64..127 of |Aut|, may be null
    uint8_t *perm_row;   // [ISE_MAX_ROWS] per candidate
    uint8_t *perm_col;   // [ISE_MAX_COLS] per candidate
    uint8_t *orbits;     // [ISE_MAX_ROWS] per candidate
    uint16_t *n_gens;
    uint8_t *gens;       // [gen_cap][ISE_MAX_COLS] per candidate, may be null
    uint32_t *counters;  // [4] per candidate, may be null
    uint32_t gen_cap;
};

// position of the k-th (0-based) set bit of m (k < popcount(m))
ISE_DEV int select_bit(uint64_t m, int k) {
    int lo = 0, hi = 63;            // answer p: popc(m & lowmask(p+1)) == k+1
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (popc64(m & lowmask(mid + 1)) >= k + 1) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// One candidate staged in a group's lanes.
struct Staged {
    int d, b, fixed;
    bool row_permut;
    uint64_t bmask, rowv, cellv, cell_next, below_l;
};

// Load candidate `ci` into the lanes (rows, cell masks of every row and of
// the row after the last).
template <int G>
ISE_DEV Staged load_candidate(const Grp<G> &g, uint32_t ci, uint32_t meta, const uint64_t *rows,
                              const uint64_t *cells, uint32_t stride) {
    const int l = g.ln;
    Staged c;
    c.d = meta & 0xff;
    c.b = (meta >> 8) & 0xff;
    c.fixed = (meta >> 16) & 0xff;
    c.row_permut = ((meta >> 24) & ISE_SAVE_ROW_PERMUTATIONS) != 0;
    c.bmask = lowmask(c.b);
    const int d = c.d, b = c.b;
    // lane l reads word l (coalesced)
    const uint64_t *rp = rows + (uint64_t)ci * stride;
    c.rowv = (l < d) ? rp[l] : 0ull;
    if (cells) {
        // column 0 always starts an orbit; columns past ncols do not exist
        c.cellv = (l < d) ? ((cells[(uint64_t)ci * stride + l] | 1ull) & c.bmask) : 0ull;
    } else {
        // column c starts an orbit of row l when an earlier row changes value at c
        uint64_t m = 1;
        for (int j = 0; j < d; j++) {
            uint64_t r = g.read64(c.rowv, j);
            if (j < l) m |= r ^ (r << 1);
        }
        c.cellv = m & c.bmask;
    }
    // orbits induced by all d rows (indexing the column-orbit generators)
    const uint64_t last = g.read64(c.rowv, d - 1);
    c.cell_next = (g.read64(c.cellv, d - 1) | (last ^ (last << 1))) & c.bmask;
    c.below_l = lowmask(l);                                        // columns 0..l-1
    (void)b;
    return c;
}

// Fill `geo` (the group's LDS scratch of d*G words): word row*G+l holds, for
// the column orbit of row `row` that contains column l, its first column, its
// length and the number of ones of the original row inside it, so that row
// comparisons do not recompute orbit boundaries.
template <int G>
ISE_DEV void write_geo(const Grp<G> &g, const Staged &c, uint32_t *geo) {
    const int l = g.ln, d = c.d, b = c.b;
    const uint64_t le = (l == 63) ? ~0ull : ((2ull << l) - 1);   // columns 0..l
    for (int r = 0; r < d; r++) {
        const uint64_t C = g.read64(c.cellv, r), R = g.read64(c.rowv, r);
        int s0 = hib64(C & le);
        const uint64_t above = C & ~le & c.bmask;
        const int e0 = above ? ctz64(above) : b;
        int len = e0 - s0;
        int cR0 = popc64(R & (lowmask(e0) & ~lowmask(s0)));
        if (l >= b) { s0 = l; len = 1; cR0 = 0; }
        geo[r * G + l] = (uint32_t)s0 | ((uint32_t)len << 8) | ((uint32_t)cR0 << 16);
    }
}

template <int G>
ISE_DEV Staged stage(const Grp<G> &g, uint32_t ci, uint32_t meta, const uint64_t *rows,
                     const uint64_t *cells, uint32_t stride, uint32_t *geo) {
    const Staged c = load_candidate(g, ci, meta, rows, cells, stride);
    write_geo(g, c, geo);
    return c;
}

// Compare original row `row` with row `prow` placed at position `row`
// (checkColOrbit over every orbit at once): +1 the permuted row is smaller,
// -1 larger, 0 equal.  The column permutation is stably partitioned inside
// every orbit before the deciding one; *dec receives the deciding orbit's
// first column.
template <int G, bool COUNT>
ISE_DEV int compare_row(const Grp<G> &g, const Staged &c, const uint32_t *geo, int row, int prow, int &pc,
                        int *dec, uint32_t &orbs_cmp, uint32_t &cells_cmp) {
    const int l = g.ln;
    const uint32_t gi = geo[row * G + l];
    const int s0 = gi & 0xff, len = (gi >> 8) & 0xff, cR = (gi >> 16) & 0xff;
    const uint64_t om = (~0ull >> (64 - len)) << s0;
    const uint64_t Pw = g.read64(c.rowv, prow);
    const uint64_t C = g.read64(c.cellv, row);
    const bool act = l < c.b;
    const int bit = act ? (int)((Pw >> pc) & 1ull) : 0;
    const uint64_t P = g.ballot(bit != 0);
    const int cP = popc64(P & om);
    // orbit heads whose counts differ (C marks the heads)
    const uint64_t bad = g.ballot(cP != cR) & C;
    int lim = c.b, outcome = 0;
    if (bad) {
        lim = ctz64(bad);
        const uint64_t more = g.ballot(cR > cP);
        outcome = ((more >> lim) & 1ull) ? 1 : -1;
        *dec = lim;
    }
    if constexpr (COUNT) {
        if (bad) {
            orbs_cmp += popc64(C & lowmask(lim + 1));
            const uint64_t after = C & ~lowmask(lim + 1);
            cells_cmp += (uint32_t)(after ? ctz64(after) : c.b);
        } else {
            orbs_cmp += popc64(C);
            cells_cmp += (uint32_t)c.b;
        }
    }
    // stable partition (ones first) inside every orbit before `lim`
    const int ones_before = popc64(P & om & c.below_l);
    const int np = bit ? s0 + ones_before : cP + l - ones_before;
    const int dst = (act && s0 < lim) ? np : l;
    if (g.ballot(dst != l))
        pc = g.scatter(pc, dst);
    return outcome;
}

// Result of one candidate before it is written.
struct Verdict {
    int verdict, frow, fcol, rtc;
    uint64_t order, order_hi;
    uint32_t perms, rows_cmp, orbs_cmp, cells_cmp, ngen;
};

// rowToChange (CanonicityChecker.cpp:320-330): max(row, perm[0..row])
template <int G> ISE_DEV int row_to_change(const Grp<G> &g, int pr, int d, int row) {
    const int l = g.ln;
    const int inv = g.scatter(l, (l < d) ? pr : l);
    const uint64_t seen = g.ballot(l < d && inv <= row);
    const int mx = hib64(seen);
    return mx > row ? mx : row;
}

// The reference's search (plain order) on a staged candidate.
template <int G, bool COUNT>
ISE_DEV void plain_search(const Grp<G> &g, const Staged &c, const uint32_t *geo, uint32_t ci,
                          const OutPtrs &o, uint32_t max_perms, int &pr, int &ob, int &pc, Verdict &v) {
    const int l = g.ln, d = c.d, b = c.b;
    PermState st;
    st.stab = d - 1;
    st.stab_aut = EMAX;
    st.order = 1;
    st.order_hi = 0;
    int row = EMAX;
    for (;;) {
        row = next_row_perm(g, st, pr, ob, d, row, c.fixed);
        if (row == EMAX)
            break;
        if (++v.perms > max_perms) { v.verdict = ISE_VERDICT_LIMIT; break; }
        int outcome = 0, dec = 0;
        for (; row < d; row++) {
            if constexpr (COUNT) v.rows_cmp++;
            outcome = compare_row<G, COUNT>(g, c, geo, row, g.read(pr, row), pc, &dec, v.orbs_cmp, v.cells_cmp);
            if (outcome) break;
        }
        if (outcome > 0)
            continue;                 // refuted prefix: advance position `row`
        if (outcome < 0) {            // the permuted matrix is larger: witness
            v.verdict = 0;
            v.frow = row;
            v.fcol = dec;
            v.rtc = row_to_change(g, pr, d, row);
            break;
        }
        // automorphism: record a generator
        if (o.gens && v.ngen < o.gen_cap) {
            uint8_t *gp = o.gens + ((uint64_t)ci * o.gen_cap + v.ngen) * ISE_MAX_COLS;
            if (!c.row_permut) {
                // permutation induced on the column orbits of the next row
                // (CanonicityChecker.h:427-450)
                const int norb = popc64(c.cell_next);
                for (int x0 = 0; x0 < ISE_MAX_COLS; x0 += G) {
                    const int x = x0 + l;
                    const int src = x < norb ? select_bit(c.cell_next, x) : 0;
                    const int pcs = g.gather(pc, src < G ? src : 0);
                    gp[x] = (uint8_t)(x < norb ? popc64(c.cell_next & lowmask(pcs)) : 0xff);
                }
            } else {
                for (int x0 = 0; x0 < ISE_MAX_COLS; x0 += G) {
                    const int x = x0 + l;
                    const int vv = g.gather(pr, x < G ? x : 0);
                    gp[x] = (uint8_t)(x < d ? vv : 0xff);
                }
            }
        }
        v.ngen++;
        add_automorphism(g, st, pr, ob, d, c.row_permut);
        row = IMAX;
    }
    if (v.verdict == 1 && c.row_permut)
        grow_order(g, st, ob, d);
    v.order = st.order;
    v.order_hi = st.order_hi;
    (void)b;
}

// A word of the set stack holds the set in its low b bits and, in its top
// byte, the number of ones of original row i in the level-i orbit holding
// position l (what every permuted row's count there is compared with):
// written once by stage_levels and carried by every later store to the word,
// so the stack is the group's only LDS.  Matrices wider than PACKED_COLS
// columns run the replayed search.
constexpr int PACKED_COLS = 56;
constexpr uint64_t COUNT_BYTE = 0xff00000000000000ull;
ISE_DEV int level_count(uint64_t w) { return (int)(uint32_t)(w >> 56); }

// Column-set state of the search (used when the candidate is in orbit form,
// see orbit_form).  The column permutation of the reference only matters
// through the set of source columns it places in each column orbit: a row
// compares orbit by orbit by the number of ones, which does not depend on the
// order inside an orbit, and the ones-first partition of an orbit that
// compared equal splits its set into the sets of the two orbits of the next
// row.  So lane l keeps S = the set of source columns placed in the orbit of
// level `row` that contains position l, one 64-bit word, and
//   count of the permuted row in that orbit = popc(R[perm[row]] & S),
//   descending to row+1:  S &= (bit l of original row `row`) ? R : ~R.
// Level-i sets are kept on a per-candidate LDS stack (word i*G + l) so that
// the walk can return to any earlier position.  What this drops is the order
// inside an orbit, i.e. the column permutation itself: kernels that have to
// report it run the replayed search.
//
// Orbit form, which makes the above exact: inside every column orbit of row i
// the ones of row i come first, and the orbits of row i+1 are those of row i
// split after these ones.  Rows written as the reference's MakeRow writes them
// (Source/MatrixCanonChecker.cpp:14-93) have it.
// Orbit form, lane-parallel: lane r checks row r -- no zero followed by a
// one inside one of its column orbits (the ones come first) and, for given
// cell masks, the orbits of row r+1 are those of row r split after its ones.
template <int G>
ISE_DEV bool orbit_form(const Grp<G> &g, const Staged &c, const uint64_t *cells) {
    const int l = g.ln;
    const uint64_t R = c.rowv, C = c.cellv;
    bool ok = ((R >> 1) & ~R & ~(C >> 1) & lowmask(c.b - 1)) == 0;
    if (cells) {
        const int nx = l + 1 < c.d ? l + 1 : l;
        const uint64_t Cn = ((uint64_t)(uint32_t)g.gather((int)(uint32_t)(C >> 32), nx) << 32) |
                            (uint32_t)g.gather((int)(uint32_t)C, nx);
        if (l + 1 < c.d) ok = ok && Cn == ((C | (R ^ (R << 1))) & c.bmask);
    }
    return g.ballot(l < c.d && !ok) == 0;
}

// The identity sets of levels lo..hi-1 onto the stack, each word with its
// count byte.  The search stages a level before it first reads it: a walk
// that ends near the last rows never stages the others.
template <int G>
ISE_DEV void stage_levels(const Grp<G> &g, const Staged &c, uint64_t *stk, int lo, int hi) {
    const int l = g.ln, b = c.b;
    const uint64_t le = (l == 63) ? ~0ull : ((2ull << l) - 1);
    for (int r = lo; r < hi; r++) {
        const uint64_t C = g.read64(c.cellv, r), R = g.read64(c.rowv, r);
        const int s0 = hib64(C & le);
        const uint64_t above = C & ~le & c.bmask;
        const int e0 = above ? ctz64(above) : b;
        uint64_t om = lowmask(e0) & ~lowmask(s0);
        int cR0 = popc64(R & om);
        if (l >= b) { om = 0; cR0 = 0; }
        stk[r * G + l] = om | ((uint64_t)cR0 << 56);
    }
}

// Generator of an automorphism found by the column-set search: the
// permutation it induces on the column orbits of the next row
// (CanonicityChecker.h:427-450) -- the orbit whose first column is the
// smallest source column placed in orbit x (columns of one orbit keep their
// relative order under the ones-first partitions) -- or, for a full matrix
// with t_saveRowPermutations, the row permutation.  S holds the level-d sets.
template <int G>
ISE_DEV void set_generator(const Grp<G> &g, const Staged &c, uint64_t S, int pr, uint32_t ci, const OutPtrs &o,
                           uint32_t ngen) {
    const int l = g.ln, d = c.d;
    if (!o.gens || ngen >= o.gen_cap)
        return;
    uint8_t *gp = o.gens + ((uint64_t)ci * o.gen_cap + ngen) * ISE_MAX_COLS;
    if (!c.row_permut) {
        const int norb = popc64(c.cell_next);
        for (int x0 = 0; x0 < ISE_MAX_COLS; x0 += G) {
            const int x = x0 + l;
            const int src = x < norb ? select_bit(c.cell_next, x) : 0;
            const int sl = src < G ? src : 0;
            const uint32_t lo = (uint32_t)g.gather((int)(uint32_t)S, sl);
            const uint32_t hi = (uint32_t)g.gather((int)(uint32_t)(S >> 32), sl);
            const uint64_t Sx = ((uint64_t)hi << 32) | lo;
            const int pcs = Sx ? ctz64(Sx) : 0;
            gp[x] = (uint8_t)(x < norb ? popc64(c.cell_next & lowmask(pcs)) : 0xff);
        }
    } else {
        for (int x0 = 0; x0 < ISE_MAX_COLS; x0 += G) {
            const int x = x0 + l;
            const int vv = g.gather(pr, x < G ? x : 0);
            gp[x] = (uint8_t)(x < d ? vv : 0xff);
        }
    }
}

// Positions i+1.. of the permutation: the rows of `rest` in increasing order
// (the walk keeps the tail after the last changed position sorted).
template <int G> ISE_DEV int fill_tail(const Grp<G> &g, int pr, int i, int d, uint32_t rest) {
    const int l = g.ln;
    return (l > i && l < d) ? select_bit((uint64_t)rest, l - i - 1) : pr;
}

// Rows of U that one of level i's first Q column orbits refutes, for all of
// them at once: lane r*Q + q compares row r with orbit q (Q = G / RG orbits,
// RG = 1 << rg_log = 16 or 32 rows), counting as the walk counts there.  A
// row whose first difference among those orbits has the larger count in the
// original row is refuted (sibling_run's `more` test) whatever the later
// orbits hold, so the walk moves past it without comparing it again.
// rrow: row (lane / Q) of the candidate; ost (lane i): the first columns of
// level i's first four orbits, a byte each (0xff: no such orbit).  The
// level's stack words are current.
template <int G>
ISE_DEV uint32_t level_refuted(const Grp<G> &g, const uint64_t *stk, int i, uint32_t U, uint64_t rrow,
                               uint32_t ost, int rg_log) {
    const int l = g.ln, qlog = __builtin_ctz(G) - rg_log, q = l & ((1 << qlog) - 1), r = l >> qlog;
    const uint32_t s = ((uint32_t)g.read((int)ost, i) >> (8 * q)) & 0xffu;
    const bool live = s != 0xffu && ((U >> r) & 1u);
    const uint64_t w = stk[i * G + (live ? (int)s : 0)];
    const int cP = popc64(rrow & w);             // rows hold no bit in the count byte
    const int cR = level_count(w);
    const uint64_t ne = g.ballot(live && cP != cR);
    const uint64_t gt = g.ballot(live && cR > cP);
    // lane j < RG: row j's slots, refuted when the first slot that differs is one of gt's
    const bool row_lane = l < (1 << rg_log);
    const int sh = (row_lane ? l : 0) << qlog;
    const uint32_t qm = (1u << (1 << qlog)) - 1u;
    const uint32_t n = row_lane ? (uint32_t)(ne >> sh) & qm : 0u;
    const uint32_t t = (uint32_t)(gt >> sh) & qm;
    return (uint32_t)g.ballot((n & (0u - n) & t) != 0u) & U;
}

// Byte k of lane i: the first column of the k-th orbit of level i (k < 4),
// 0xff past the last orbit.
template <int G> ISE_DEV uint32_t orbit_starts(const Staged &c) {
    uint64_t C = c.cellv;
    uint32_t o = 0;
    for (int k = 0; k < 4; k++) {
        o |= (C ? (uint32_t)ctz64(C) : 0xffu) << (8 * k);
        C &= C - 1;
    }
    return o;
}

// Rows tried at one position of the depth-first phase: compares row `cur`
// and, while it is refuted, moves on to the next larger unused row (the bits
// of U above cur).  Stops on an equal row (outcome SIB_EQUAL), a larger one
// (SIB_LARGER, fcol = the first column of the deciding orbit) or a refuted
// one with no larger unused row left (SIB_REFUTED); with CHECKED also when
// `left` permutations have been used (SIB_REFUTED).  Returns the number of
// rows moved on to.  Rows of `ref` (known refuted at this position,
// level_refuted) are moved past uncompared.
constexpr int SIB_EQUAL = 0, SIB_LARGER = 1, SIB_REFUTED = 2;

template <int G, bool COUNT, bool CHECKED>
ISE_DEV uint32_t sibling_run(const Grp<G> &g, const Staged &c, uint64_t S, int cR, uint64_t C, uint32_t U,
                             uint32_t ref, uint32_t left, int &cur, int &outcome, int &fcol, Verdict &v) {
    uint32_t k = 0;
    for (;;) {
        if constexpr (!CHECKED && !COUNT) {
            if ((ref >> cur) & 1u) {
                const uint32_t abv = U & (0xfffffffeu << cur);
                const uint32_t open = abv & ~ref;
                if (!open) {
                    // every larger unused row is refuted too: the walk ends on the last
                    if (abv) {
                        k += (uint32_t)__popc(abv);
                        cur = 31 - __builtin_clz(abv);
                    }
                    outcome = SIB_REFUTED;
                    return k;
                }
                const int nx = __builtin_ctz(open);
                k += (uint32_t)__popc(abv & ((2u << nx) - 1u));
                cur = nx;
            }
        }
        const uint64_t Rp = g.read64(c.rowv, cur);
        const int cP = popc64(Rp & S);
        const uint64_t bad = g.ballot(cP != cR) & C;
        const uint64_t more = g.ballot(cR > cP);
        if constexpr (COUNT) {
            v.rows_cmp++;
            if (bad) {
                const int lim = ctz64(bad);
                v.orbs_cmp += popc64(C & lowmask(lim + 1));
                const uint64_t after = C & ~lowmask(lim + 1);
                v.cells_cmp += (uint32_t)(after ? ctz64(after) : c.b);
            } else {
                v.orbs_cmp += popc64(C);
                v.cells_cmp += (uint32_t)c.b;
            }
        }
        if (!bad) {
            outcome = SIB_EQUAL;
            break;
        }
        const int lim = ctz64(bad);
        if (!((more >> lim) & 1ull)) {
            outcome = SIB_LARGER;              // larger: a witness
            fcol = lim;
            break;
        }
        outcome = SIB_REFUTED;
        const uint32_t above = U & (0xfffffffeu << cur);
        if (!above)
            break;                             // no larger unused row
        if constexpr (CHECKED) {
            if (k == left)
                break;                         // out of permutations
        }
        k++;
        cur = ctz64(above);
    }
    return k;
}

// The reference's search (plain order) on the column-set state: the same
// walk, verdict, witness row permutation, fail row/column, rowToChange,
// automorphisms, generators, |Aut|, orbits and counters as plain_search.
//
// Until the first automorphism the row orbits are trivial, so the walk of
// next_row_perm (GroupOrder.h:136-248) is a plain depth-first search over
// row prefixes in increasing row order: a row refuted at position i is
// followed by the next larger row not used before i; when there is none the
// search backs up to the deepest position that has one (the walk's
// reversal and lexicographic successor), which is also where the walk puts
// its stabiliser point.  That phase runs on a scalar mask U of the rows not
// used before the current position, with positions after it implied (the
// unused rows in increasing order), and writes out the full permutation only
// when it stops.  From the first automorphism on, the replayed walk takes
// over with the permutation, stabiliser point and orbits the walk would have.
template <int G, bool COUNT>
ISE_DEV void set_search(const Grp<G> &g, const Staged &c, uint64_t *stk, int floor, uint32_t ci,
                        const OutPtrs &o, uint32_t max_perms, int &pr, int &ob, Verdict &v) {
    const int l = g.ln, d = c.d;
    PermState st;
    st.stab = d - 1;
    st.stab_aut = EMAX;
    st.order = 1;
    st.order_hi = 0;
    uint64_t S = 0;
    int row = EMAX;
    bool aut = false;

    // ---- depth-first phase -------------------------------------------------
    {
        int i = d - 1;                         // start: the identity, backed up from its last row
        uint32_t U = 1u << (d - 1);            // rows not used before position i (perm[i] included)
        uint32_t left = max_perms > v.perms ? max_perms - v.perms : 0u;   // permutations still allowed
        bool first = true;
        // rows refuted by a level's first orbits (level_refuted), found when the
        // walk enters the level: lane i keeps level i's, refd marks the levels
        // found for their current prefix.  Not with counters, which count every
        // row the walk compares.
        const int rg_log = (d <= 16 || G == 16) ? 4 : 5;   // a row past RG is never marked
        uint64_t rrow = 0;
        uint32_t refd = 0, ost = 0;
        int refv = 0;
        if constexpr (!COUNT) {
            const int src = l >> (__builtin_ctz(G) - rg_log);
            rrow = ((uint64_t)(uint32_t)g.gather((int)(uint32_t)(c.rowv >> 32), src) << 32) |
                   (uint32_t)g.gather((int)(uint32_t)c.rowv, src);
            ost = orbit_starts<G>(c);
        }
        for (;;) {
            // back up to the deepest position with a larger unused row
            int p = g.uni(i), nxt = -1;
            U = (uint32_t)g.uni((int)U);
            int e = g.read(pr, p);
            for (;;) {
                const uint32_t above = U & (0xfffffffeu << e);
                if (above) {
                    bool through = false;
                    if constexpr (!COUNT) {
                        // every larger row refuted by the level's first orbits: the walk
                        // moves through them there and backs up again
                        if (((refd >> p) & 1u) && left >= 32u && p >= (first ? c.fixed : 0))
                            through = (above & ~(uint32_t)g.read(refv, p)) == 0u;
                    }
                    if (!through) { nxt = ctz64(above); break; }
                    left -= (uint32_t)__popc(above);
                    if (st.stab > p) st.stab = p;
                    first = false;
                }
                if (p == 0) break;
                p--;
                e = g.read(pr, p);
                U |= 1u << e;
            }
            if (nxt < 0 || p < (first ? c.fixed : 0)) {
                // exhausted: the walk ends on the descending permutation
                // (on the identity when it could not move at all)
                if (!first && l < d) pr = d - 1 - l;
                break;
            }
            first = false;
            if (st.stab > p) st.stab = p;
            i = g.uni(p);
            if (i < floor) {
                // levels i and above can be reached from here on
                stage_levels(g, c, stk, i, floor);
                floor = i;
            }
            int cur = g.uni(nxt);
            if (l == i) pr = cur;
            const uint64_t w = stk[i * G + l];
            S = w & ~COUNT_BYTE;
            uint64_t C = g.uni64(g.read64(c.cellv, i));
            int cR = level_count(w);
            uint32_t ref = 0;
            if constexpr (!COUNT) {
                if ((refd >> i) & 1u) {
                    ref = (uint32_t)g.read(refv, i);
                } else {
                    ref = level_refuted(g, stk, i, U, rrow, ost, rg_log);
                    if (l == i) refv = (int)ref;
                    refd |= 1u << i;
                }
            }
            bool fresh = true;                 // cur makes a new permutation (not a descent)
            int why = 0;                       // 1 witness, 2 automorphism, 3 limit
            for (;;) {
                if (fresh) {
                    if (left == 0) { why = 3; break; }
                    left--;
                }
                // rows tried at position i until one is not refuted: only cur changes
                int outcome = SIB_REFUTED, fcol = 0;
                if (left >= 32u)               // at most 31 more rows at one position
                    left -= sibling_run<G, COUNT, false>(g, c, S, cR, C, U, ref, left, cur, outcome, fcol, v);
                else
                    left -= sibling_run<G, COUNT, true>(g, c, S, cR, C, U, 0u, left, cur, outcome, fcol, v);
                if (l == i) pr = cur;
                if (outcome == SIB_EQUAL) {
                    // equal: descend to the next position (Rp: the row just compared)
                    const uint64_t Rp = g.read64(c.rowv, cur);
                    const uint64_t Ro = g.read64(c.rowv, i);
                    S &= ((Ro >> l) & 1ull) ? Rp : ~Rp;
                    U &= ~(1u << cur);
                    if (++i == d) { why = 2; break; }
                    cur = ctz64(U);
                    if (l == i) pr = cur;
                    const uint64_t wn = stk[i * G + l];
                    stk[i * G + l] = S | (wn & COUNT_BYTE);
                    C = g.uni64(g.read64(c.cellv, i));
                    cR = level_count(wn);
                    i = g.uni(i);
                    cur = g.uni(cur);
                    U = (uint32_t)g.uni((int)U);
                    if constexpr (!COUNT) {
                        ref = level_refuted(g, stk, i, U, rrow, ost, rg_log);
                        if (l == i) refv = (int)ref;
                        refd |= 1u << i;
                    }
                    fresh = false;
                    continue;
                }
                if (outcome == SIB_LARGER) {
                    // the permuted matrix is larger: witness
                    pr = fill_tail(g, pr, i, d, U & ~(1u << cur));
                    v.verdict = 0;
                    v.frow = i;
                    v.fcol = fcol;
                    v.rtc = row_to_change(g, pr, d, i);
                    why = 1;
                    break;
                }
                const uint32_t above = U & (0xfffffffeu << cur);
                if (above) {
                    // out of permutations on the next larger row
                    cur = ctz64(above);
                    if (l == i) pr = cur;
                    why = 3;
                }
                break;                         // no larger unused row: back up
            }
            if (why == 3) {
                pr = fill_tail(g, pr, i, d, U & ~(1u << cur));
                v.verdict = ISE_VERDICT_LIMIT;
            }
            if (why == 2) aut = true;
            if (why) break;
        }
        v.perms = max_perms - left;            // the hunt's and this phase's
    }
    if (!aut) {
        if (v.verdict == 1 && c.row_permut)
            grow_order(g, st, ob, d);
        v.order = st.order;
    v.order_hi = st.order_hi;
        return;
    }

    // ---- replayed walk from the first automorphism on ------------------------
    // (levels never reached hold the identity sets: the rows there never moved)
    if (floor > 0) stage_levels(g, c, stk, 0, floor);
    for (;;) {
        if (aut) {
            set_generator(g, c, S, pr, ci, o, v.ngen);
            v.ngen++;
            add_automorphism(g, st, pr, ob, d, c.row_permut);
            row = IMAX;
            aut = false;
        }
        row = g.uni(next_row_perm(g, st, pr, ob, d, row, c.fixed));
        if (row == EMAX)
            break;
        if (++v.perms > max_perms) { v.verdict = ISE_VERDICT_LIMIT; break; }
        const uint64_t w = stk[row * G + l];
        S = w & ~COUNT_BYTE;
        uint64_t C = g.read64(c.cellv, row);
        int cR = level_count(w);
        int outcome = 0, dec = 0;
        bool limit = false;
        for (;;) {
            row = g.uni(row);
            C = g.uni64(C);
            st.stab = g.uni(st.stab);
            v.perms = (uint32_t)g.uni((int)v.perms);
            if constexpr (COUNT) v.rows_cmp++;
            const uint64_t Rp = g.read64(c.rowv, g.read(pr, row));
            const int cP = popc64(Rp & S);
            const uint64_t bad = g.ballot(cP != cR) & C;
            outcome = 0;
            if (!bad) {
                if constexpr (COUNT) { v.orbs_cmp += popc64(C); v.cells_cmp += (uint32_t)c.b; }
                const uint64_t Ro = g.read64(c.rowv, row);
                S &= ((Ro >> l) & 1ull) ? Rp : ~Rp;
                if (++row == d)
                    break;                    // every row matched: automorphism
                const uint64_t wn = stk[row * G + l];
                stk[row * G + l] = S | (wn & COUNT_BYTE);
                C = g.read64(c.cellv, row);
                cR = level_count(wn);
                continue;
            }
            const int lim = ctz64(bad);
            const uint64_t more = g.ballot(cR > cP);
            outcome = ((more >> lim) & 1ull) ? 1 : -1;
            dec = lim;
            if constexpr (COUNT) {
                v.orbs_cmp += popc64(C & lowmask(lim + 1));
                const uint64_t after = C & ~lowmask(lim + 1);
                v.cells_cmp += (uint32_t)(after ? ctz64(after) : c.b);
            }
            if (outcome < 0 || row == st.stab)
                break;
            // refuted: next larger unused row at this position, if any
            const int old = g.read(pr, row);
            const uint64_t big = g.ballot(l > row && l < d && pr > old);
            if (!big)
                break;
            const int succ = ctz64(big);
            const int nw = g.read(pr, succ);
            if (l == row) pr = nw;
            if (l == succ) pr = old;
            if (++v.perms > max_perms) { limit = true; break; }
        }
        if (limit) { v.verdict = ISE_VERDICT_LIMIT; break; }
        if (outcome > 0)
            continue;                 // refuted prefix: the walk advances position `row`
        if (outcome < 0) {            // the permuted matrix is larger: witness
            v.verdict = 0;
            v.frow = row;
            v.fcol = dec;
            v.rtc = row_to_change(g, pr, d, row);
            break;
        }
        aut = true;                   // automorphism: S holds the level-d sets
    }
    if (v.verdict == 1 && c.row_permut)
        grow_order(g, st, ob, d);
    v.order = st.order;
    v.order_hi = st.order_hi;
}

// Witness hunt in a proposed order: depth-first over row prefixes, position i
// trying rows in the order order[i*od + 0..d-1] (ascending when `order` is
// null), with the same comparison and column state as the plain search but
// no automorphism pruning; at most `budget` rows compared.  Extending the
// identity prefix by its own row before the first comparison is equal by
// construction and not compared.
// Returns true with a witness in pr (prefix, then the unused rows in
// increasing order) and v.frow / v.fcol / v.rtc set.
template <int G, bool COUNT>
ISE_DEV bool hunt(const Grp<G> &g, const Staged &c, const uint32_t *geo, const uint8_t *order, int od,
                  uint32_t budget, int &pr, int &pc, Verdict &v) {
    const int l = g.ln, d = c.d;
    int cursor = 0;            // lane i: next index into order row i
    uint32_t used = 0;         // rows placed at positions 0..i-1
    int i = 0;
    int idlen = 0;             // positions 0..idlen-1 hold the identity
    uint32_t nodes = 0;
    for (;;) {
        if (i == d) {          // the whole matrix matched: backtrack
            i--;
            used &= ~(1u << g.read(pr, i));
            if (idlen > i) idlen = i;
            continue;
        }
        int k = g.read(cursor, i);
        int j = -1;
        while (k < d) {
            const int cand = order ? (int)order[i * od + k] : k;
            k++;
            // entries that name no row of this candidate are skipped
            if (cand < d && !((used >> cand) & 1)) { j = cand; break; }
        }
        if (l == i) cursor = k;
        if (j < 0) {           // depth exhausted
            if (i == 0) return false;
            i--;
            used &= ~(1u << g.read(pr, i));
            if (idlen > i) idlen = i;
            continue;
        }
        if (l == i) pr = j;
        if (j == i && idlen == i && nodes == 0) {
            // identity prefix extended by its own row before any comparison:
            // the column permutation is still the identity, the rows are
            // equal and comparing would move no column
            used |= 1u << j;
            idlen = ++i;
            if (l == i) cursor = 0;
            continue;
        }
        if (nodes == budget) return false;
        nodes++;
        v.perms++;
        if constexpr (COUNT) v.rows_cmp++;
        int dec = 0;
        const int outcome = compare_row<G, COUNT>(g, c, geo, i, j, pc, &dec, v.orbs_cmp, v.cells_cmp);
        if (outcome < 0) {
            // complete the permutation with the unused rows in increasing order
            const uint32_t u = ~(used | (1u << j)) & (uint32_t)lowmask(d);
            if (l > i && l < d) pr = select_bit((uint64_t)u, l - i - 1);
            v.verdict = 0;
            v.frow = i;
            v.fcol = dec;
            v.rtc = row_to_change(g, pr, d, i);
            return true;
        }
        if (outcome > 0) continue;
        used |= 1u << j;
        i++;
        if (l == i) cursor = 0;
    }
}

// The witness hunt on the column-set state: the same depth-first order,
// budget, counters and witness as `hunt`, comparisons as in set_search.  The
// set stack holds the identity sets on entry (as stage_levels leaves it) and
// the sets of the hunt's prefixes on return.
template <int G, bool COUNT>
ISE_DEV bool hunt_sets(const Grp<G> &g, const Staged &c, uint64_t *stk, const uint8_t *order,
                       int od, uint32_t budget, int &pr, Verdict &v) {
    const int l = g.ln, d = c.d;
    int cursor = 0;            // lane i: next index into order row i
    uint32_t used = 0;         // rows placed at positions 0..i-1
    int i = 0;
    int idlen = 0;             // positions 0..idlen-1 hold the identity
    uint32_t nodes = 0;
    for (;;) {
        i = g.uni(i);
        if (i == d) {          // the whole matrix matched: backtrack
            i--;
            used &= ~(1u << g.read(pr, i));
            if (idlen > i) idlen = i;
            continue;
        }
        int k = g.read(cursor, i);
        int j = -1;
        while (k < d) {
            const int cand = order ? (int)order[i * od + k] : k;
            k++;
            // entries that name no row of this candidate are skipped
            if (cand < d && !((used >> cand) & 1)) { j = cand; break; }
        }
        if (l == i) cursor = k;
        if (j < 0) {           // depth exhausted
            if (i == 0) return false;
            i--;
            used &= ~(1u << g.read(pr, i));
            if (idlen > i) idlen = i;
            continue;
        }
        j = g.uni(j);
        if (l == i) pr = j;
        if (j == i && idlen == i && nodes == 0) {
            // identity prefix extended by its own row before any comparison
            // (the set stack already holds the identity sets)
            used |= 1u << j;
            idlen = ++i;
            if (l == i) cursor = 0;
            continue;
        }
        if (nodes == budget) return false;
        nodes++;
        v.perms++;
        const uint64_t w = stk[i * G + l];
        const uint64_t S = w & ~COUNT_BYTE;
        const uint64_t C = g.uni64(g.read64(c.cellv, i));
        const int cR = level_count(w);
        const uint64_t Rp = g.read64(c.rowv, j);
        const int cP = popc64(Rp & S);
        const uint64_t bad = g.ballot(cP != cR) & C;
        const uint64_t more = g.ballot(cR > cP);
        if constexpr (COUNT) v.rows_cmp++;
        if (bad) {
            const int lim = ctz64(bad);
            if constexpr (COUNT) {
                v.orbs_cmp += popc64(C & lowmask(lim + 1));
                const uint64_t after = C & ~lowmask(lim + 1);
                v.cells_cmp += (uint32_t)(after ? ctz64(after) : c.b);
            }
            if ((more >> lim) & 1ull)
                continue;      // refuted: next row at this position
            // larger: witness; the unused rows follow in increasing order
            const uint32_t u = ~(used | (1u << j)) & (uint32_t)lowmask(d);
            if (l > i && l < d) pr = select_bit((uint64_t)u, l - i - 1);
            v.verdict = 0;
            v.frow = i;
            v.fcol = lim;
            v.rtc = row_to_change(g, pr, d, i);
            return true;
        }
        if constexpr (COUNT) { v.orbs_cmp += popc64(C); v.cells_cmp += (uint32_t)c.b; }
        if (i + 1 < d) {
            const uint64_t Ro = g.read64(c.rowv, i);
            const uint64_t wn = stk[(i + 1) * G + l];
            stk[(i + 1) * G + l] = (S & (((Ro >> l) & 1ull) ? Rp : ~Rp)) | (wn & COUNT_BYTE);
        }
        used |= 1u << j;
        i++;
        if (l == i) cursor = 0;
    }
}

// Write the record of candidate `ci`: the scalar fields from lane 0, the
// row permutation and orbits (and the column permutation when asked for)
// from every lane of the group.
template <int G, bool COUNT>
ISE_DEV void write_record(const Grp<G> &g, uint32_t ci, const OutPtrs &o, const Verdict &v, int pr, int ob, int pc,
                          int d, int b) {
    const int l = g.ln;
    if (l == 0) {
        o.verdict[ci] = (uint8_t)v.verdict;
        if (o.fail_row) o.fail_row[ci] = (uint8_t)v.frow;
        if (o.fail_col) o.fail_col[ci] = (uint8_t)v.fcol;
        if (o.row_to_change) o.row_to_change[ci] = (uint8_t)v.rtc;
        if (o.group_order) o.group_order[ci] = v.verdict == 1 ? v.order : 0ull;
        if (o.group_order_hi) o.group_order_hi[ci] = v.verdict == 1 ? v.order_hi : 0ull;
        if (o.n_gens) o.n_gens[ci] = (uint16_t)(v.ngen > 0xffffu ? 0xffffu : v.ngen);
        if (COUNT && o.counters) {
            uint32_t *cc = o.counters + (uint64_t)ci * 4;
            cc[0] = v.perms; cc[1] = v.rows_cmp; cc[2] = v.orbs_cmp; cc[3] = v.cells_cmp;
        }
    }
    // every lane takes part in the gathers; lanes past the record skip the store
    for (int x0 = 0; x0 < ISE_MAX_ROWS; x0 += G) {
        const int x = x0 + l;
        const int vp = g.gather(pr, x < G ? x : 0), vo = g.gather(ob, x < G ? x : 0);
        if (x < ISE_MAX_ROWS) {
            if (o.perm_row) o.perm_row[(uint64_t)ci * ISE_MAX_ROWS + x] = (uint8_t)(x < d ? vp : 0xff);
            if (o.orbits) o.orbits[(uint64_t)ci * ISE_MAX_ROWS + x] = (uint8_t)(x < d ? vo : 0xff);
        }
    }
    if (o.perm_col) {
        for (int x0 = 0; x0 < ISE_MAX_COLS; x0 += G) {
            const int x = x0 + l;
            const int vc = g.gather(pc, x < G ? x : 0);
            o.perm_col[(uint64_t)ci * ISE_MAX_COLS + x] = (uint8_t)(x < b ? vc : 0xff);
        }
    }
}

// Full canonicity test of candidate `ci` by the group g.  With HUNT, a
// budgeted witness hunt in the proposed order runs first; the plain search
// decides whatever the hunt leaves open (canonical candidates, |Aut|).  With
// SETS (no column permutation reported), a candidate in orbit form with at
// most PACKED_COLS columns runs the column-set search, any other the replayed
// one.  `lds` is the group's
// scratch: d*G orbit-geometry words, or in SETS mode d*G 64-bit set words.
//
// With `defer` set, a candidate whose search runs past max_perms is appended
// to that list (count in *defer_count) and no record is written: it is
// checked again from the start by a later launch.
template <int G, bool COUNT, bool HUNT, bool SETS>
ISE_DEV void check_candidate(const Grp<G> &g, uint32_t ci, uint32_t meta, const uint64_t *rows,
                             const uint64_t *cells, uint32_t stride, const OutPtrs &o, uint32_t max_perms,
                             uint32_t *lds, const uint8_t *order, int od, uint32_t hunt_budget,
                             uint32_t *defer, uint32_t *defer_count) {
    const int l = g.ln;
    const Staged c = load_candidate(g, ci, meta, rows, cells, stride);
    const int d = c.d, b = c.b;
    int pr = l, ob = l, pc = l;
    Verdict v{1, EMAX, EMAX, EMAX, 1ull, 0ull, 0u, 0u, 0u, 0u, 0u};
    bool found = false;
    uint32_t *geo = lds;
    if constexpr (SETS) {
        uint64_t *stk = reinterpret_cast<uint64_t *>(lds);
        if (c.b <= PACKED_COLS && orbit_form(g, c, cells)) {
            bool hfound = false;
            int floor = d;   // levels floor..d-1 staged
            if constexpr (HUNT) {
                // an order table narrower than the candidate cannot guide it: plain search only
                if (!order || d <= od) {
                    stage_levels(g, c, stk, 0, d);   // the hunt starts at the first row
                    hfound = hunt_sets<G, COUNT>(g, c, stk, order ? order + (uint64_t)ci * od * od : nullptr,
                                                 od, hunt_budget, pr, v);
                    if (!hfound) {
                        pr = l;
                        stage_levels(g, c, stk, 0, d);   // identity sets again
                        floor = 0;
                    }
                }
            }
            if (!hfound)
                set_search<G, COUNT>(g, c, stk, floor, ci, o, max_perms, pr, ob, v);
            found = true;
        }
    }
    if (!found) {
        // replayed search (column permutation reported, or not in orbit form)
        write_geo(g, c, geo);
        bool hfound = false;
        if constexpr (HUNT) {
            // an order table narrower than the candidate cannot guide it: plain search only
            if (!order || d <= od)
                hfound = hunt<G, COUNT>(g, c, geo, order ? order + (uint64_t)ci * od * od : nullptr, od,
                                        hunt_budget, pr, pc, v);
            if (!hfound) { pr = l; pc = l; }
        }
        if (!hfound)
            plain_search<G, COUNT>(g, c, geo, ci, o, max_perms, pr, ob, pc, v);
    }

    if (defer && v.verdict == ISE_VERDICT_LIMIT) {
        if (l == 0) defer[atomicAdd(defer_count, 1u)] = ci;
        return;
    }
    write_record<G, COUNT>(g, ci, o, v, pr, ob, pc, d, b);
}

} // namespace ise
