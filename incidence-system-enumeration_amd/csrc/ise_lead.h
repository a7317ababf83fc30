// ise_lead.h -- the learned order of the first row, on the column-set form of
// the search (one candidate per wave, depth <= 16).
//
// Until the walk of CGroupOrder::next_permutation (GroupOrder.h:136-248) meets
// its first automorphism it is a depth-first search over row prefixes in
// increasing row order, pruned where a prefix compares smaller
// (CanonicityChecker.h:326-416).  The prefixes that compare equal do not
// depend on the order in which the rows of one position are tried, so a
// search that tries the rows of position 0 in a proposed order, and the rows
// of every later position in increasing order:
//   - meets an automorphism iff the group is not trivial: the candidate then
//     takes the plain search, whose |Aut|, orbits and generators are exact;
//   - ends with no witness and no automorphism iff the matrix is canonical
//     with a trivial group: the plain search's results, written directly
//     (verdict 1, |Aut| 1, no generators, the walk's last permutation: the
//     descending one);
//   - otherwise stops at a witness of its own order (a larger matrix).
// The subtree of each first row is the plain search's depth-first phase from
// position 1 on (`lead_subtree`); the first row 0 continues the identity the
// way the plain walk starts.  The rows of position 0 come from the row-order
// network scoring every row as the first one (`lead_order`), or from a given
// table.
#pragma once
#include "ise_wave.h"
#include "ise_mlp_dev.h"

namespace ise {

constexpr int LEAD_MAX_ROWS = 16;   // one 16-row tile of the network

// Waves per SIMD the lead kernel's registers must leave room for: 5 is what
// its 8 KiB of LDS per wave (16 rows) allows; the compiler then keeps it to
// 96 VGPRs.
#ifndef ISE_LEAD_WAVES_PER_EU
#define ISE_LEAD_WAVES_PER_EU 5
#endif
constexpr int LEAD_EXHAUSTED = 0, LEAD_WITNESS = 1, LEAD_AUT = 2, LEAD_LIMIT = 3;

// Feature `id` of row j as the row of position 0 (the oracle's
// orc_lead_features): A[s] = ones of row j in the columns of weight m - s
// (s < 3) or <= m - 3 (s = 3), weight = rows holding a one there, m = the
// largest weight; A0 the same of row 0.
struct LeadStats {
    int A[4], A0[4];
    int sg;     // sign of A against A0, lexicographic
    int m0;     // largest weight of a column row j shares with row 0 (0 for row 0)
    int mn;     // the same with row d-1 (0 for it)
    int ties;   // rows with the same A
};

ISE_DEV float lead_feature(int id, const LeadStats &s, int j, int d) {
    if (j >= d) return 0.f;
    if (id < 4) return (float)s.A[id] * 0.0625f;
    if (id < 8) return (float)(s.A[id - 4] - s.A0[id - 4]) * 0.0625f;
    switch (id) {
    case 8: return (float)s.sg;
    case 9: return j == 0 ? 1.f : 0.f;
    case 10: return j == d - 1 ? 1.f : 0.f;
    case 11: return (float)j * 0.0625f;
    case 12: return (float)d * 0.0625f;
    case 13: return (float)s.m0 * 0.0625f;
    case 14: return (float)s.mn * 0.0625f;
    default: return (float)s.ties * 0.0625f;
    }
}

// Statistics of row j = lane % 16 (every lane; lanes 16.. repeat rows 0..15).
ISE_DEV LeadStats lead_stats(const Grp<64> &g, const Staged &c) {
    const int l = g.ln, d = c.d, b = c.b, j = l & 15;
    LeadStats s;
    int wc = 0;                                   // lane l: weight of column l
    for (int i = 0; i < d; i++) {
        const uint64_t r = g.read64(c.rowv, i);
        wc += (int)((r >> l) & 1ull);
    }
    int m = 0;
    for (int t = d; t >= 1; t--)
        if (g.ballot(wc >= t)) { m = t; break; }
    const uint64_t M[4] = {g.ballot(wc == m), g.ballot(wc == m - 1), g.ballot(wc == m - 2),
                           g.ballot(l < b && wc <= m - 3)};
    const uint64_t Rj = ((uint64_t)(uint32_t)g.gather((int)(uint32_t)(c.rowv >> 32), j) << 32) |
                        (uint32_t)g.gather((int)(uint32_t)c.rowv, j);
    const uint64_t R0 = g.read64(c.rowv, 0), Rn = g.read64(c.rowv, d - 1);
    int key = 0;
    s.sg = 0;
    for (int q = 0; q < 4; q++) {
        s.A[q] = j < d ? popc64(Rj & M[q]) : 0;
        s.A0[q] = g.read(s.A[q], 0);
        if (!s.sg) s.sg = s.A[q] > s.A0[q] ? 1 : (s.A[q] < s.A0[q] ? -1 : 0);
        key |= s.A[q] << (6 * q);
    }
    const uint64_t X0 = (j != 0 && j < d) ? (Rj & R0) : 0ull, Xn = (j != d - 1 && j < d) ? (Rj & Rn) : 0ull;
    s.m0 = 0;
    s.mn = 0;
    for (int t = 1; t <= m; t++) {                // weights form nested masks: count the levels met
        const uint64_t T = g.ballot(l < b && wc >= t);
        s.m0 += (X0 & T) != 0ull;
        s.mn += (Xn & T) != 0ull;
    }
    s.ties = 0;
    for (int q = 0; q < d; q++) s.ties += g.read(key, q) == key;
    return s;
}

// The rows of position 0 by decreasing network score (ties by increasing
// row): lane r < d returns the row of rank r; *score (lane j < 16) the score
// of row j.  frag: the weights as A fragments (global memory), w: the packed
// fp32 parameters (biases).
ISE_DEV int lead_order(const Grp<64> &g, const Staged &c, const bf16x4 *frag, const float *w, float *score) {
    const int l = g.ln, d = c.d, q = l >> 4, j = l & 15;
    const LeadStats s = lead_stats(g, c);
    bf16x4 xb;
    for (int t = 0; t < 4; t++) {
        // feature 4q + t: the four candidates are compile-time ids, selected by q
        const float f0 = lead_feature(t, s, j, d), f1 = lead_feature(4 + t, s, j, d);
        const float f2 = lead_feature(8 + t, s, j, d), f3 = lead_feature(12 + t, s, j, d);
        xb[t] = (short)f2bf(q == 0 ? f0 : q == 1 ? f1 : q == 2 ? f2 : f3);
    }
    float sc = mlp_tile_streamed(xb, frag, w, l);   // lane l < 16: row l
    if (!(sc == sc)) sc = -__builtin_inff();      // a NaN ranks last
    *score = sc;
    int rank = 0;
    for (int k = 0; k < d; k++) {
        const float sk = __int_as_float(g.read(__float_as_int(sc), k));
        rank += (sk > sc) || (sk == sc && k < l);
    }
    return g.scatter(l, l < d ? rank : l);
}

// Depth-first search below a fixed row at position 0, over positions 1..d-1
// in increasing row order: the plain search's depth-first phase (set_search)
// with position 1 as its lowest.  enter: start by trying `cur` at position i
// (a new first row, whose compare at position 0 began the permutation); else
// start by backing up from position i (the identity, as the plain walk
// starts).  Permutations are counted as the walk counts them: a new row at a
// position after a back-up or a refutation is one, a descent is not.  Levels floor..d-1 hold the identity sets and
// are restaged when the back-up first reaches them.
template <bool COUNT>
ISE_DEV int lead_subtree(const Grp<64> &g, const Staged &c, uint64_t *stk, int i, uint32_t U, int cur, bool enter,
                         int floor, uint32_t &left, int &pr, Verdict &v, uint64_t rrow, uint32_t ost) {
    constexpr int G = 64;
    const int l = g.ln, d = c.d;
    constexpr int rg_log = 4;                     // depth <= 16
    uint32_t refd = 0;
    int refv = 0;
    for (;;) {
        if (!enter) {
            // back up to the deepest position >= 1 with a larger unused row
            int p = g.uni(i), nxt = -1;
            U = (uint32_t)g.uni((int)U);
            int e = g.read(pr, p);
            for (;;) {
                const uint32_t above = U & (0xfffffffeu << e);
                if (above) {
                    bool through = false;
                    if constexpr (!COUNT) {
                        if (((refd >> p) & 1u) && left >= 32u)
                            through = (above & ~(uint32_t)g.read(refv, p)) == 0u;
                    }
                    if (!through) { nxt = ctz64(above); break; }
                    left -= (uint32_t)__popc(above);
                }
                if (p == 1) break;
                p--;
                e = g.read(pr, p);
                U |= 1u << e;
            }
            if (nxt < 0)
                return LEAD_EXHAUSTED;
            i = g.uni(p);
            if (i < floor) {
                stage_levels(g, c, stk, i, floor);
                floor = i;
            }
            cur = g.uni(nxt);
        }
        if (l == i) pr = cur;
        const uint64_t w = stk[i * G + l];
        uint64_t S = w & ~COUNT_BYTE;
        uint64_t C = g.uni64(g.read64(c.cellv, i));
        int cR = level_count(w);
        uint32_t ref = 0;
        if constexpr (!COUNT) {
            if (!enter && ((refd >> i) & 1u)) {
                ref = (uint32_t)g.read(refv, i);
            } else {
                ref = level_refuted(g, stk, i, U, rrow, ost, rg_log);
                if (l == i) refv = (int)ref;
                refd |= 1u << i;
            }
        }
        bool fresh = !enter;                     // a new first row's first compare continues its permutation
        enter = false;
        int why = 0;                              // 1 witness, 2 automorphism, 3 limit
        for (;;) {
            if (fresh) {
                if (left == 0) { why = 3; break; }
                left--;
            }
            int outcome = SIB_REFUTED, fcol = 0;
            if (left >= 32u)
                left -= sibling_run<G, COUNT, false>(g, c, S, cR, C, U, ref, left, cur, outcome, fcol, v);
            else
                left -= sibling_run<G, COUNT, true>(g, c, S, cR, C, U, 0u, left, cur, outcome, fcol, v);
            if (l == i) pr = cur;
            if (outcome == SIB_EQUAL) {
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
                pr = fill_tail(g, pr, i, d, U & ~(1u << cur));
                v.verdict = 0;
                v.frow = i;
                v.fcol = fcol;
                v.rtc = row_to_change(g, pr, d, i);
                why = 1;
                break;
            }
            if (U & (0xfffffffeu << cur)) why = 3;   // out of permutations before the next row
            break;                                    // no larger unused row: back up
        }
        if (why == 1) return LEAD_WITNESS;
        if (why == 2) return LEAD_AUT;
        if (why == 3) return LEAD_LIMIT;
    }
}

// The search with the rows of position 0 in the order lane k of `lead` holds
// (k < 16; rows >= d, repeated rows and lanes past d skipped), then any row
// not listed, in increasing order.  The stack is restaged here.
template <bool COUNT>
ISE_DEV int lead_search(const Grp<64> &g, const Staged &c, uint64_t *stk, int lead, uint32_t max_perms, int &pr,
                        Verdict &v) {
    constexpr int G = 64;
    const int l = g.ln, d = c.d;
    uint32_t left = max_perms;
    uint64_t rrow = 0;
    uint32_t ost = 0;
    if constexpr (!COUNT) {
        const int src = l >> 2;                   // level_refuted: lane r*4 + q holds row r
        rrow = ((uint64_t)(uint32_t)g.gather((int)(uint32_t)(c.rowv >> 32), src) << 32) |
               (uint32_t)g.gather((int)(uint32_t)c.rowv, src);
        ost = orbit_starts<G>(c);
    }
    stage_levels(g, c, stk, 0, d);                // count bytes of every level, identity sets
    const uint64_t w0 = stk[l];
    const uint64_t S0 = w0 & ~COUNT_BYTE;
    const int cR0 = level_count(w0);
    const uint64_t C0 = g.uni64(g.read64(c.cellv, 0));
    const uint64_t R0 = g.read64(c.rowv, 0);
    const uint32_t all = (uint32_t)lowmask(d);
    uint32_t tried = 0;
    bool staged = true;                           // the stack holds the identity sets
    for (int k = 0; k < LEAD_MAX_ROWS + d; k++) {
        int j0 = k < LEAD_MAX_ROWS ? g.uni(g.read(lead, k)) : k - LEAD_MAX_ROWS;
        if (j0 < 0 || j0 >= d || ((tried >> j0) & 1u)) continue;
        tried |= 1u << j0;
        int r;
        if (j0 == 0) {
            // the identity: back up from its last row, as the plain walk starts
            pr = l;
            r = lead_subtree<COUNT>(g, c, stk, d - 1, 1u << (d - 1), 0, false, staged ? 0 : d, left, pr, v,
                                    rrow, ost);
        } else {
            if (left == 0) {
                v.perms += max_perms;
                return LEAD_LIMIT;
            }
            left--;
            const uint64_t Rp = g.read64(c.rowv, j0);
            const int cP = popc64(Rp & S0);
            const uint64_t bad = g.ballot(cP != cR0) & C0;
            if constexpr (COUNT) {
                v.rows_cmp++;
                if (bad) {
                    const int lim = ctz64(bad);
                    v.orbs_cmp += popc64(C0 & lowmask(lim + 1));
                    const uint64_t after = C0 & ~lowmask(lim + 1);
                    v.cells_cmp += (uint32_t)(after ? ctz64(after) : c.b);
                } else {
                    v.orbs_cmp += popc64(C0);
                    v.cells_cmp += (uint32_t)c.b;
                }
            }
            const uint32_t rest = all & ~(1u << j0);
            if (bad) {
                const int lim = ctz64(bad);
                if ((g.ballot(cR0 > cP) >> lim) & 1ull)
                    continue;                     // refuted at position 0
                pr = l == 0 ? j0 : fill_tail(g, pr, 0, d, rest);
                v.verdict = 0;
                v.frow = 0;
                v.fcol = lim;
                v.rtc = row_to_change(g, pr, d, 0);
                v.perms += max_perms - left;
                return LEAD_WITNESS;
            }
            const uint64_t S1 = S0 & (((R0 >> l) & 1ull) ? Rp : ~Rp);
            stk[G + l] = S1 | (stk[G + l] & COUNT_BYTE);
            staged = false;
            pr = l == 0 ? j0 : fill_tail(g, pr, 0, d, rest);
            r = lead_subtree<COUNT>(g, c, stk, 1, rest, ctz64(rest), true, 0, left, pr, v, rrow, ost);
        }
        if (r != LEAD_EXHAUSTED) {
            v.perms += max_perms - left;
            return r;
        }
        staged = false;                           // descents overwrote the sets
    }
    v.perms += max_perms - left;
    return LEAD_EXHAUSTED;
}

// Learned-order test of candidate `ci` (G = 64, orbit form, depth 2..16,
// lenStab 0, at most PACKED_COLS columns; others take check_candidate): the
// plain search with a budget of `gate` permutations; when that leaves it open
// and found no automorphism, the lead search; when that meets an automorphism
// or its permutation limit, the plain search from the start.  lead_in (a
// table [n][16]) replaces the network's order; lead_out / score_out receive
// the order used and the scores.
template <bool COUNT>
ISE_DEV void lead_candidate(const Grp<64> &g, uint32_t ci, const Staged &c, const OutPtrs &o, uint32_t max_perms,
                            uint32_t gate, uint64_t *stk, const bf16x4 *frag, const float *w,
                            const uint8_t *lead_in, uint8_t *lead_out, float *score_out) {
    const int l = g.ln, d = c.d;
    int pr = l, ob = l;
    Verdict v{1, EMAX, EMAX, EMAX, 1ull, 0ull, 0u, 0u, 0u, 0u, 0u};
    set_search<64, COUNT>(g, c, stk, d, ci, o, gate < max_perms ? gate : max_perms, pr, ob, v);
    if (v.verdict == ISE_VERDICT_LIMIT && gate < max_perms && v.ngen == 0) {
        int lead;
        float sc = 0.f;
        if (lead_in) {
            lead = l < LEAD_MAX_ROWS ? (int)lead_in[(uint64_t)ci * LEAD_MAX_ROWS + l] : 0xff;
        } else {
            lead = lead_order(g, c, frag, w, &sc);
        }
        if (lead_out && l < LEAD_MAX_ROWS)
            lead_out[(uint64_t)ci * LEAD_MAX_ROWS + l] = (uint8_t)(l < d ? lead : 0xff);
        if (score_out && l < LEAD_MAX_ROWS)
            score_out[(uint64_t)ci * LEAD_MAX_ROWS + l] = l < d ? sc : 0.f;
        // counters: the gate's work, then the lead search's
        Verdict lv{1, EMAX, EMAX, EMAX, 1ull, 0ull, v.perms, v.rows_cmp, v.orbs_cmp, v.cells_cmp, 0u};
        pr = l;
        const int r = lead_search<COUNT>(g, c, stk, lead, max_perms, pr, lv);
        if (r == LEAD_WITNESS) {
            v = lv;
            ob = l;
        } else if (r == LEAD_EXHAUSTED) {
            // canonical with a trivial group: the plain search's results
            v = lv;
            v.verdict = 1;
            v.order = 1ull;
            v.order_hi = 0ull;
            pr = l < d ? d - 1 - l : l;
            ob = l;
        } else {
            // counters: the gate and lead search's work, then the plain search's
            Verdict pv{1, EMAX, EMAX, EMAX, 1ull, 0ull, 0u, lv.rows_cmp, lv.orbs_cmp, lv.cells_cmp, 0u};
            pr = l;
            ob = l;
            set_search<64, COUNT>(g, c, stk, d, ci, o, max_perms, pr, ob, pv);
            pv.perms += lv.perms;
            v = pv;
        }
    } else if (v.verdict == ISE_VERDICT_LIMIT && gate < max_perms) {
        // the gate met an automorphism: the plain search from the start
        Verdict pv{1, EMAX, EMAX, EMAX, 1ull, 0ull, 0u, 0u, 0u, 0u, 0u};
        pr = l;
        ob = l;
        set_search<64, COUNT>(g, c, stk, d, ci, o, max_perms, pr, ob, pv);
        v = pv;
    }
    write_record<64, COUNT>(g, ci, o, v, pr, ob, l, d, c.b);
}

} // namespace ise
