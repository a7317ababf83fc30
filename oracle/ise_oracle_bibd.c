/*
 * ise_oracle_bibd.c -- CPU oracle: canonicity test of (partial) incidence
 * matrices and a depth-first BIBD enumerator.
 *
 * TEST INFRASTRUCTURE ONLY (see ise_oracle.h).  Works on an unpacked uint8
 * matrix and walks column orbits one by one, the way the reference's scalar
 * checker does, so that it is an independent formulation of the bit-mask /
 * lane-parallel one used by the HIP kernel.
 *
 * The permutation generator keeps the reference's 8-bit index arithmetic
 * (T = unsigned char, ELEMENT_MAX = 255, IDX_MAX = 254; Source/DataTypes.h:186-188),
 * because the order in which row permutations are visited decides which
 * witness is reported first and which automorphisms are found.
 */
#include "ise_oracle.h"
#include <string.h>
#include <stdlib.h>

typedef unsigned char u8;
#define EMAX ((u8)255)  /* ELEMENT_MAX */
#define IMAX ((u8)254)  /* IDX_MAX     */

/* Group-order bookkeeping of CGroupOrder (GroupOrder.h:51-107). */
typedef struct {
    unsigned short stab;      /* m_nStabLength    */
    unsigned short stab_aut;  /* m_nStabLengthAut */
    uint64_t order;           /* m_nGroupOrder, bits 0..63 (a 128-bit UInt in the reference) */
    uint64_t order_hi;        /* bits 64..127 */
} grp_t;

/* Reverse perm[lo+1 .. hi-1]  (GroupOrder.h:25-28, `revert`). */
static void reverse_tail(u8 *perm, u8 hi, u8 lo)
{
    while (++lo < --hi) { u8 t = perm[lo]; perm[lo] = perm[hi]; perm[hi] = t; }
}

/* Merge row orbits under an automorphism, starting at position `from`
 * (GroupOrder.h:30-49).  Orbits are labelled by their smallest element. */
static void merge_orbits(const u8 *perm, unsigned short len, u8 *orb, u8 from)
{
    u8 p = from;
    do {
        u8 a = orb[p], c = orb[perm[p]];
        if (a == c) continue;
        if (c < a) { u8 t = a; a = c; c = t; }
        for (unsigned q = len; q--;)
            if (orb[q] == c) orb[q] = a;
    } while (++p < len);
}

/* Multiply |Aut| by the length of the orbit of the current stabiliser point
 * (GroupOrder.h:63-73). */
static void grow_group_order(grp_t *g, unsigned short degree, const u8 *orb)
{
    uint64_t len = 1;
    unsigned short p = g->stab_aut;
    while (++p < degree)
        if (orb[p] == g->stab_aut) len++;
    const unsigned __int128 o = (((unsigned __int128)g->order_hi << 64) | g->order) * len;
    g->order = (uint64_t)o;
    g->order_hi = (uint64_t)(o >> 64);
}

/* Record a new automorphism (GroupOrder.h:56-62, :76-79, :88-102). */
static void record_automorphism(grp_t *g, unsigned short degree, const u8 *perm,
                                u8 *orb, int row_permut)
{
    u8 first_moved = 0;
    while (first_moved == perm[first_moved]) first_moved++;
    if (row_permut && g->stab_aut > first_moved)
        grow_group_order(g, degree, orb);
    g->stab_aut = first_moved;
    g->stab = first_moved;
    merge_orbits(perm, degree, orb, first_moved);
}

/*
 * Next row permutation fixing the first `fixed` rows, skipping images at the
 * stabiliser position that are not orbit leaders (GroupOrder.h:136-248).
 *   at == EMAX : plain lexicographic successor
 *   at == IMAX : first call after an automorphism was recorded
 *   at <  IMAX : the prefix perm[0..at] was refuted; change position `at`
 * Returns the first position whose image changed, or EMAX when exhausted.
 */
static u8 next_row_perm(grp_t *g, u8 *perm, const u8 *orb, u8 n, u8 at, u8 fixed)
{
    u8 piv, succ, old;

    if (at == IMAX && perm[g->stab] == n - 1)
        at = EMAX;

    if (at == IMAX) {
        /* restart the search just after the stabiliser position */
        piv = (u8)g->stab;
        old = perm[piv];
        at = piv;
        for (succ = n; --succ > old;)
            perm[succ] = succ;
        for (u8 q = succ++; q-- > piv;)
            perm[q + 1] = q;
    } else if (at >= IMAX) {
        succ = piv = n;
        while (--piv > 0 && perm[piv - 1] >= perm[piv])
            ;
        if (piv == fixed)
            return EMAX;
        old = perm[--piv];
        while (perm[--succ] <= old)
            ;
    } else {
        piv = succ = at;
        old = perm[piv];
        while (++succ < n && perm[succ] <= old)
            ;
        if (succ >= n) {
            reverse_tail(perm, n, piv);
            return next_row_perm(g, perm, orb, n, EMAX, 0);
        }
    }

    if (g->stab == piv) {
        /* the image of the stabiliser point must be an orbit leader */
        int step_back = 0;
        u8 q = succ, img = perm[succ];
        if (at >= IMAX) {
            while (q > piv && orb[perm[q]] != perm[q])
                q--;
            if (q != succ) {
                if (!q)
                    return EMAX;
                step_back = (q == piv);
                img = perm[q--];
                while (++q < succ)
                    perm[q] = perm[q + 1];
            }
        } else {
            while (q < n && orb[perm[q]] != perm[q])
                q++;
            if (q != succ) {
                step_back = (q == n);
                if (step_back) {
                    if (!piv)
                        return EMAX;
                    q = at - 1;
                    while (++q < succ)
                        perm[q] = q;
                } else {
                    img = perm[q++];
                    while (--q > succ)
                        perm[q] = perm[q - 1];
                }
            }
        }
        perm[succ] = img;
        if (step_back) {
            succ = at >= IMAX ? n - 1 : piv;
            old = perm[--piv];
            g->stab = piv;
        }
    }

    perm[piv] = perm[succ];
    perm[succ] = old;
    if (at >= IMAX) {
        if (g->stab > piv)
            g->stab = piv;
        reverse_tail(perm, n, piv);
    }
    return piv;
}

int orc_walk_step(uint8_t *perm, const uint8_t *orbits, int n, int at, int fixed, uint16_t *state)
{
    grp_t g;
    g.stab = state[0];
    g.stab_aut = state[1];
    g.order = 1;
    g.order_hi = 0;
    u8 r = next_row_perm(&g, perm, orbits, (u8)n, (u8)at, (u8)fixed);
    state[0] = g.stab;
    state[1] = g.stab_aut;
    return r;
}

/* Column-orbit start list of row `row` from its start mask. */
static int orbit_starts(uint64_t mask, int ncols, u8 *starts, u8 *lens)
{
    int n = 0;
    for (int c = 0; c < ncols; c++)
        if ((mask >> c) & 1) starts[n++] = (u8)c;
    for (int o = 0; o < n; o++)
        lens[o] = (u8)((o + 1 < n ? starts[o + 1] : ncols) - starts[o]);
    return n;
}

void orc_cell_masks(const uint64_t *rows, int nrows, int ncols, uint64_t *out)
{
    /* column c starts an orbit of row i when c == 0 or columns c-1 and c
     * differ somewhere in rows 0..i-1 (ColOrbitManager.h:115-212 keeps the
     * same partition as linked lists) */
    uint64_t full = ncols >= 64 ? ~0ull : ((1ull << ncols) - 1);
    uint64_t m = 1;
    for (int i = 0; i < nrows; i++) {
        out[i] = m & full;
        uint64_t r = rows[i];
        m |= (r ^ (r << 1)) & full;
    }
}

/* One column orbit of the current row against the permuted row
 * (CanonicityChecker.cpp:191-223, rank 2).  Returns +1 when the original has
 * more ones (the permuted matrix is smaller), -1 when fewer (it is larger),
 * 0 when equal; on 0 the orbit's columns of pcol are stably reordered so that
 * the columns holding a one in the permuted row come first. */
static int compare_orbit(int start, int len, const u8 *row, const u8 *prow, u8 *pcol)
{
    int ones_orig = 0, ones_perm = 0;
    u8 ones[ORC_MAX_COLS], zeros[ORC_MAX_COLS];
    int n1 = 0, n0 = 0;
    for (int q = 0; q < len; q++) {
        u8 col = pcol[start + q];
        ones_orig += row[start + q];
        if (prow[col]) ones[n1++] = col; else zeros[n0++] = col;
    }
    ones_perm = n1;
    if (ones_orig != ones_perm)
        return ones_orig > ones_perm ? 1 : -1;
    for (int q = 0; q < n1; q++) pcol[start + q] = ones[q];
    for (int q = 0; q < n0; q++) pcol[start + n1 + q] = zeros[q];
    return 0;
}

static u8 row_to_change(const u8 *perm, u8 row)
{
    u8 best = row;
    for (int q = row; q >= 0; q--)
        if (perm[q] > best) best = perm[q];
    return best;
}

typedef struct {
    int nrows, ncols, len_stab;
    unsigned out_info;
    const u8 *mat;                 /* nrows x ncols */
    u8 ostart[ORC_MAX_ROWS + 1][ORC_MAX_COLS];
    u8 olen[ORC_MAX_ROWS + 1][ORC_MAX_COLS];
    int norb[ORC_MAX_ROWS + 1];
} cand_t;

static int run_check_from(const cand_t *cd, orc_results *res, size_t idx, uint32_t gen_cap,
                          const uint32_t *init);

/* The improved solution of the last test that failed on its last row
 * (CCanonicityChecker::reconstructSolution, CanonicityChecker.cpp:227-277),
 * captured at the failure: the current solution's entries for the orbits that
 * compared equal, the deciding orbit's count of the permuted row, then fresh
 * counts of the permuted row through the column permutation for the orbits
 * after it. */
static __thread u8 improved_sol[ORC_MAX_COLS];
static __thread int improved_len = -1;

static void reconstruct_solution(const cand_t *cd, int row, int deciding, const u8 *prow, const u8 *pcol)
{
    const u8 *orig = cd->mat + (size_t)row * cd->ncols;
    int o = 0;
    for (; o < deciding; o++) {          /* copied from the current solution */
        int ones = 0;
        for (int q = 0; q < cd->olen[row][o]; q++) ones += orig[cd->ostart[row][o] + q];
        improved_sol[o] = (u8)ones;
    }
    for (; o < cd->norb[row]; o++) {     /* counters of the permuted row */
        int ones = 0;
        for (int q = 0; q < cd->olen[row][o]; q++) ones += prow[pcol[cd->ostart[row][o] + q]];
        improved_sol[o] = (u8)ones;
    }
    improved_len = cd->norb[row];
}

/* Permutations the walk may try before it stops unfinished (the learned
 * order's gate); UINT32_MAX: no limit. */
static __thread uint32_t walk_limit = 0xffffffffu;

static int run_check(const cand_t *cd, orc_results *res, size_t idx, uint32_t gen_cap)
{
    return run_check_from(cd, res, idx, gen_cap, NULL);
}

static int run_check_from(const cand_t *cd, orc_results *res, size_t idx, uint32_t gen_cap,
                          const uint32_t *init)
{
    const int d = cd->nrows, b = cd->ncols;
    const int row_permut = (cd->out_info & ORC_SAVE_ROW_PERMUTATIONS) != 0;
    u8 perm[ORC_MAX_ROWS], orb[ORC_MAX_ROWS], pcol[ORC_MAX_COLS];
    grp_t g;
    uint32_t perms = init ? init[0] : 0, rows_cmp = init ? init[1] : 0;
    uint32_t orbs_cmp = init ? init[2] : 0, cells = init ? init[3] : 0;
    uint16_t ngen = 0;
    int verdict = 1;
    u8 frow = EMAX, fcol = EMAX, rtc = EMAX;

    for (int q = 0; q < d; q++) perm[q] = orb[q] = (u8)q;
    for (int q = 0; q < b; q++) pcol[q] = (u8)q;
    g.stab = (unsigned short)(d - 1);
    g.stab_aut = EMAX;
    g.order = 1;
    g.order_hi = 0;

    u8 row = EMAX;
    for (;;) {
        row = next_row_perm(&g, perm, orb, (u8)d, row, (u8)cd->len_stab);
        if (row == EMAX)
            break;
        if (perms == walk_limit) {         /* unfinished: the counters so far */
            verdict = 2;
            goto done;
        }
        perms++;
        int refuted = 0;
        for (; row < d; row++) {
            const u8 *orig = cd->mat + (size_t)row * b;
            const u8 *prow = cd->mat + (size_t)perm[row] * b;
            rows_cmp++;
            int diff = 0, o;
            for (o = 0; o < cd->norb[row]; o++) {
                int s = cd->ostart[row][o], len = cd->olen[row][o];
                orbs_cmp++;
                cells += (uint32_t)len;
                if (len > 1)
                    diff = compare_orbit(s, len, orig, prow, pcol);
                else
                    diff = (int)orig[s] - (int)prow[pcol[s]];
                if (diff) break;
            }
            if (diff > 0) { refuted = 1; break; }
            if (diff < 0) {
                verdict = 0;
                frow = row;
                fcol = cd->ostart[row][o];
                rtc = row_to_change(perm, row);
                if (row == d - 1)
                    reconstruct_solution(cd, row, o, prow, pcol);
                goto done;
            }
        }
        if (refuted)
            continue;
        /* every row matched: perm is an automorphism */
        if (!row_permut) {
            /* permutation it induces on the column orbits of the next row
             * (CanonicityChecker.h:427-450) */
            if (res->gens && ngen < gen_cap) {
                u8 *gp = res->gens + ((size_t)idx * gen_cap + ngen) * ORC_MAX_COLS;
                memset(gp, 0xff, ORC_MAX_COLS);
                u8 colidx[ORC_MAX_COLS];
                memset(colidx, 0xff, sizeof colidx);
                for (int o = 0; o < cd->norb[d]; o++) colidx[cd->ostart[d][o]] = (u8)o;
                for (int o = 0; o < cd->norb[d]; o++) gp[o] = colidx[pcol[cd->ostart[d][o]]];
            }
        } else if (res->gens && ngen < gen_cap) {
            u8 *gp = res->gens + ((size_t)idx * gen_cap + ngen) * ORC_MAX_COLS;
            memset(gp, 0xff, ORC_MAX_COLS);
            memcpy(gp, perm, (size_t)d);
        }
        ngen++;
        record_automorphism(&g, (unsigned short)d, perm, orb, row_permut);
        row = IMAX;
    }
    if (row_permut)
        grow_group_order(&g, (unsigned short)d, orb);

done:
    if (res->verdict) res->verdict[idx] = (uint8_t)verdict;
    if (res->fail_row) res->fail_row[idx] = frow;
    if (res->fail_col) res->fail_col[idx] = fcol;
    if (res->row_to_change) res->row_to_change[idx] = rtc;
    if (res->group_order) res->group_order[idx] = verdict ? g.order : 0;
    if (res->group_order_hi) res->group_order_hi[idx] = verdict ? g.order_hi : 0;
    if (res->perm_row) {
        memset(res->perm_row + idx * ORC_MAX_ROWS, 0xff, ORC_MAX_ROWS);
        memcpy(res->perm_row + idx * ORC_MAX_ROWS, perm, (size_t)d);
    }
    if (res->perm_col) {
        memset(res->perm_col + idx * ORC_MAX_COLS, 0xff, ORC_MAX_COLS);
        memcpy(res->perm_col + idx * ORC_MAX_COLS, pcol, (size_t)b);
    }
    if (res->orbits) {
        memset(res->orbits + idx * ORC_MAX_ROWS, 0xff, ORC_MAX_ROWS);
        memcpy(res->orbits + idx * ORC_MAX_ROWS, orb, (size_t)d);
    }
    if (res->n_gens) res->n_gens[idx] = ngen;
    if (res->counters) {
        uint32_t *c = res->counters + idx * 4;
        c[0] = perms; c[1] = rows_cmp; c[2] = orbs_cmp; c[3] = cells;
    }
    return verdict;
}

static void build_orbits(cand_t *cd, const uint64_t *cells_in, const uint64_t *rows)
{
    uint64_t masks[ORC_MAX_ROWS + 1];
    if (cells_in) {
        uint64_t full = cd->ncols >= 64 ? ~0ull : ((1ull << cd->ncols) - 1);
        /* column 0 always starts an orbit; columns past ncols do not exist */
        for (int i = 0; i < cd->nrows; i++) masks[i] = (cells_in[i] | 1ull) & full;
        uint64_t r = rows[cd->nrows - 1];
        masks[cd->nrows] = (masks[cd->nrows - 1] | (r ^ (r << 1))) & full;
    } else {
        uint64_t tmp[ORC_MAX_ROWS + 1];
        uint64_t ext[ORC_MAX_ROWS + 1];
        memcpy(ext, rows, sizeof(uint64_t) * cd->nrows);
        ext[cd->nrows] = 0;
        orc_cell_masks(ext, cd->nrows + 1, cd->ncols, tmp);
        memcpy(masks, tmp, sizeof(uint64_t) * (cd->nrows + 1));
    }
    for (int i = 0; i <= cd->nrows; i++)
        cd->norb[i] = orbit_starts(masks[i], cd->ncols, cd->ostart[i], cd->olen[i]);
}

int orc_check_batch(uint32_t n, uint32_t stride, const uint64_t *rows,
                    const uint64_t *cells, const uint32_t *meta,
                    uint32_t gen_cap, orc_results *res)
{
    static __thread cand_t cd;
    static __thread u8 mat[ORC_MAX_ROWS * ORC_MAX_COLS];
    for (uint32_t c = 0; c < n; c++) {
        uint32_t m = meta[c];
        int d = m & 0xff, b = (m >> 8) & 0xff;
        if (d < 1 || d > ORC_MAX_ROWS || d > (int)stride || b < 1 || b > ORC_MAX_COLS)
            return -1;
        const uint64_t *rp = rows + (size_t)c * stride;
        cd.nrows = d; cd.ncols = b;
        cd.len_stab = (m >> 16) & 0xff;
        cd.out_info = (m >> 24) & 0xff;
        for (int i = 0; i < d; i++)
            for (int q = 0; q < b; q++)
                mat[i * b + q] = (u8)((rp[i] >> q) & 1);
        cd.mat = mat;
        build_orbits(&cd, cells ? cells + (size_t)c * stride : NULL, rp);
        run_check(&cd, res, c, gen_cap);
    }
    return 0;
}

int orc_improved_solution(const uint8_t *mat, int nrows, int ncols, int len_stab, uint8_t *out)
{
    uint8_t verdict = 1;
    orc_results res;
    memset(&res, 0, sizeof res);
    res.verdict = &verdict;
    improved_len = -1;
    if (orc_test_canonicity(mat, nrows, ncols, len_stab, 0, &res, 0) < 0) return -1;
    if (verdict != 0 || improved_len < 0) return -1;
    memcpy(out, improved_sol, (size_t)improved_len);
    return improved_len;
}

int orc_test_canonicity(const uint8_t *mat, int nrows, int ncols, int len_stab,
                        unsigned out_info, orc_results *res, uint32_t gen_cap)
{
    if (nrows < 1 || nrows > ORC_MAX_ROWS || ncols < 1 || ncols > ORC_MAX_COLS)
        return -1;
    cand_t *cd = (cand_t *)malloc(sizeof(cand_t));
    uint64_t rows[ORC_MAX_ROWS];
    for (int i = 0; i < nrows; i++) {
        rows[i] = 0;
        for (int q = 0; q < ncols; q++)
            if (mat[i * ncols + q]) rows[i] |= 1ull << q;
    }
    cd->nrows = nrows; cd->ncols = ncols; cd->len_stab = len_stab;
    cd->out_info = out_info; cd->mat = mat;
    build_orbits(cd, NULL, rows);
    int v = run_check(cd, res, 0, gen_cap);
    free(cd);
    return v;
}

/* Column order of GenerateBinaryColumnOrbits (MatrixCanonChecker.cpp:186-263)
 * applied to every row in turn: inside each orbit of row i the ones are moved
 * first (the first zero swapped with the last one, the swap carried through
 * the rows below), which leaves the columns in decreasing order. */
static void orbit_form(u8 *mat, int nrows, int ncols)
{
    u8 start[ORC_MAX_COLS + 1];
    int nst = 1;
    start[0] = 0;
    for (int i = 0; i < nrows; i++) {
        u8 nstart[ORC_MAX_COLS + 1];
        int nn = 0;
        for (int o = 0; o < nst; o++) {
            const int lo = start[o], hi = o + 1 < nst ? start[o + 1] : ncols;
            int ones = 0;
            for (int q = lo; q < hi; q++) ones += mat[i * ncols + q];
            if (ones > 0 && ones < hi - lo) {
                const int jlast = lo + ones;
                int j1 = lo, j2 = hi;
                while (j1 < jlast) {
                    while (mat[i * ncols + j1]) j1++;       /* first zero */
                    if (j1 == jlast) break;
                    while (!mat[i * ncols + --j2]) {}       /* last one   */
                    for (int r = i; r < nrows; r++) {       /* swap the two columns */
                        u8 t = mat[r * ncols + j1];
                        mat[r * ncols + j1] = mat[r * ncols + j2];
                        mat[r * ncols + j2] = t;
                    }
                    j1++;
                }
            }
            nstart[nn++] = (u8)lo;
            if (ones > 0 && ones < hi - lo) nstart[nn++] = (u8)(lo + ones);
        }
        memcpy(start, nstart, (size_t)nn);
        nst = nn;
    }
}

int orc_canonize(const uint8_t *mat, int nrows, int ncols, int len_stab, int max_iters, uint8_t *out,
                 uint64_t *aut, uint32_t *iters)
{
    if (nrows < 1 || nrows > ORC_MAX_ROWS || ncols < 1 || ncols > ORC_MAX_COLS || len_stab < 0 ||
        len_stab >= nrows)
        return -1;
    u8 cur[ORC_MAX_ROWS * ORC_MAX_COLS], tmp[ORC_MAX_ROWS * ORC_MAX_COLS];
    memcpy(cur, mat, (size_t)(nrows * ncols));
    orbit_form(cur, nrows, ncols);
    uint32_t it = 0;
    int done = 0;
    while (it < (uint32_t)max_iters) {
        uint8_t verdict = 0, prow[ORC_MAX_ROWS];
        uint64_t order = 0;
        orc_results res;
        memset(&res, 0, sizeof res);
        res.verdict = &verdict;
        res.group_order = &order;
        res.perm_row = prow;
        orc_test_canonicity(cur, nrows, ncols, len_stab, ORC_SAVE_ROW_PERMUTATIONS, &res, 0);
        it++;
        if (verdict == 1) {
            if (aut) *aut = order;
            done = 1;
            break;
        }
        /* rows rearranged by the witness: row i of the new matrix is row prow[i] */
        for (int i = 0; i < nrows; i++)
            memcpy(tmp + i * ncols, cur + prow[i] * ncols, (size_t)ncols);
        memcpy(cur, tmp, (size_t)(nrows * ncols));
        orbit_form(cur, nrows, ncols);
    }
    memcpy(out, cur, (size_t)(nrows * ncols));
    if (iters) *iters = it;
    if (!done && aut) *aut = 0;
    return done;
}

/* ------------------------------------------------------------------------ */
/* Brute force for tiny matrices.                                            */

static int cmp_col_desc(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

/* Columns as integers with row 0 as the most significant bit, sorted in
 * decreasing order, give the largest matrix for a fixed row order. */
static void canon_for_rows(const uint8_t *mat, int nrows, int ncols, const int *rp, uint64_t *colv)
{
    for (int q = 0; q < ncols; q++) {
        uint64_t x = 0;
        for (int i = 0; i < nrows; i++)
            x = (x << 1) | mat[rp[i] * ncols + q];
        colv[q] = x;
    }
    qsort(colv, (size_t)ncols, sizeof(uint64_t), cmp_col_desc);
}

int orc_bruteforce_canon(const uint8_t *mat, int nrows, int ncols,
                         uint8_t *canon_out, uint64_t *n_row_auts)
{
    if (nrows > 10 || ncols > 64) return -1;
    int rp[16], c[16];
    uint64_t best[64], cur[64], self[64];
    int have = 0;
    uint64_t auts = 0;
    for (int i = 0; i < nrows; i++) { rp[i] = i; c[i] = 0; }
    canon_for_rows(mat, nrows, ncols, rp, self);
    /* Heap's algorithm over all row orders */
    int i = 0;
    for (;;) {
        canon_for_rows(mat, nrows, ncols, rp, cur);
        if (!have || memcmp(cur, best, sizeof(uint64_t) * ncols) != 0) {
            /* compare the two column-sorted matrices row by row, column 0 first */
            int better = !have;
            for (int r = 0; r < nrows && have; r++) {
                int decided = 0;
                for (int q = 0; q < ncols; q++) {
                    int x = (int)((cur[q] >> (nrows - 1 - r)) & 1), y = (int)((best[q] >> (nrows - 1 - r)) & 1);
                    if (x != y) { better = x > y; decided = 1; break; }
                }
                if (decided) break;
            }
            if (better) { memcpy(best, cur, sizeof(uint64_t) * ncols); have = 1; }
        }
        if (memcmp(cur, self, sizeof(uint64_t) * ncols) == 0) auts++;
        while (i < nrows && c[i] >= i) { c[i] = 0; i++; }
        if (i >= nrows) break;
        int t;
        if (i % 2 == 0) { t = rp[0]; rp[0] = rp[i]; rp[i] = t; }
        else { t = rp[c[i]]; rp[c[i]] = rp[i]; rp[i] = t; }
        c[i]++;
        i = 0;
    }
    if (canon_out)
        for (int r = 0; r < nrows; r++)
            for (int q = 0; q < ncols; q++)
                canon_out[r * ncols + q] = (uint8_t)((best[q] >> (nrows - 1 - r)) & 1);
    if (n_row_auts) *n_row_auts = auts;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Depth-first enumerator (Enumerator.cpp:540-1024 in shape; own row solver) */

typedef struct {
    int v, b, r, k, lambda;
    uint64_t rows[ORC_MAX_ROWS];
    uint64_t *designs; uint64_t *auts; long max_designs; long ndesigns;
    uint64_t *stats;
    uint64_t *cand_rows; uint32_t *cand_meta; long cand_cap; int stride, dmin, dmax;
    u8 mat[ORC_MAX_ROWS * ORC_MAX_COLS];
    /* when set, solutions of the row system are appended here instead of
     * being recursed into (level-by-level enumeration) */
    uint64_t *collect; long ncollect, collect_cap;
} enum_t;

/* orbit data of the row about to be built */
typedef struct {
    int n;
    u8 start[ORC_MAX_COLS], len[ORC_MAX_COLS], weight[ORC_MAX_COLS];
    uint32_t inrow[ORC_MAX_COLS];  /* bit i: orbit has a one in row i */
    int lo[ORC_MAX_COLS], hi[ORC_MAX_COLS];
} orbs_t;

static int check_partial(enum_t *e, int depth, uint64_t *aut)
{
    const int b = e->b;
    for (int i = 0; i < depth; i++)
        for (int q = 0; q < b; q++)
            e->mat[i * b + q] = (u8)((e->rows[i] >> q) & 1);
    if (e->cand_rows && depth >= e->dmin && depth <= e->dmax && (long)e->stats[4] < e->cand_cap) {
        long c = (long)e->stats[4]++;
        memset(e->cand_rows + (size_t)c * e->stride, 0, sizeof(uint64_t) * e->stride);
        memcpy(e->cand_rows + (size_t)c * e->stride, e->rows, sizeof(uint64_t) * depth);
        unsigned oi = depth == e->v ? (ORC_SAVE_ROW_PERMUTATIONS | ORC_SAVE_ROW_TO_CHANGE) : 0;
        e->cand_meta[c] = (uint32_t)depth | ((uint32_t)b << 8) | (oi << 24);
    }
    uint8_t verdict = 0; uint64_t order = 0;
    orc_results res;
    memset(&res, 0, sizeof res);
    res.verdict = &verdict; res.group_order = &order;
    unsigned oi = depth == e->v ? (ORC_SAVE_ROW_PERMUTATIONS | ORC_SAVE_ROW_TO_CHANGE) : 0;
    orc_test_canonicity(e->mat, depth, b, 0, oi, &res, 0);
    e->stats[0]++;
    if (aut) *aut = order;
    return verdict;
}

static void dfs_row(enum_t *e, int depth);

/* Enumerate orbit counts x[o] for the row `depth` (the reference solves the
 * same system with CInSysSolver, InSysSolver.h:92-207). */
static void solve_orbit(enum_t *e, int depth, const orbs_t *ob, int o, int *x,
                        int sum, int *lam, const int *cap_r, const int *cap_l)
{
    if (o == ob->n) {
        if (sum != e->r) return;
        for (int i = 0; i < depth; i++) if (lam[i] != e->lambda) return;
        uint64_t row = 0;
        for (int p = 0; p < ob->n; p++)
            if (x[p]) row |= (((x[p] >= 64) ? ~0ull : ((1ull << x[p]) - 1)) << ob->start[p]);
        e->rows[depth] = row;
        if (e->collect) {
            if (e->ncollect < e->collect_cap) e->collect[e->ncollect] = row;
            e->ncollect++;
            return;
        }
        dfs_row(e, depth + 1);
        return;
    }
    /* remaining capacity after this orbit */
    for (int val = ob->hi[o]; val >= ob->lo[o]; val--) {
        int ns = sum + val;
        if (ns > e->r || ns + cap_r[o + 1] < e->r) continue;
        int ok = 1;
        for (int i = 0; i < depth; i++) {
            int nl = lam[i] + (((ob->inrow[o] >> i) & 1) ? val : 0);
            if (nl > e->lambda || nl + cap_l[(o + 1) * ORC_MAX_ROWS + i] < e->lambda) { ok = 0; break; }
        }
        if (!ok) continue;
        x[o] = val;
        for (int i = 0; i < depth; i++) if ((ob->inrow[o] >> i) & 1) lam[i] += val;
        solve_orbit(e, depth, ob, o + 1, x, ns, lam, cap_r, cap_l);
        for (int i = 0; i < depth; i++) if ((ob->inrow[o] >> i) & 1) lam[i] -= val;
    }
}

static void expand(enum_t *e, int depth);

static void dfs_row(enum_t *e, int depth)
{
    if (depth > 0) {
        uint64_t aut = 0;
        int canon = check_partial(e, depth, &aut);
        if (!canon) return;
        if (depth == e->v) {
            e->stats[2]++;
            if (e->ndesigns < e->max_designs && e->designs) {
                memcpy(e->designs + (size_t)e->ndesigns * e->v, e->rows, sizeof(uint64_t) * e->v);
                if (e->auts) e->auts[e->ndesigns] = aut;
            }
            e->ndesigns++;
            e->stats[3]++;
            return;
        }
        e->stats[1]++;
    }
    expand(e, depth);
}

/* Solve the system of row `depth` over the column orbits of rows 0..depth-1. */
static void expand(enum_t *e, int depth)
{
    /* orbits of row `depth` */
    orbs_t ob;
    uint64_t masks[ORC_MAX_ROWS + 1];
    uint64_t ext[ORC_MAX_ROWS + 1];
    memcpy(ext, e->rows, sizeof(uint64_t) * depth);
    ext[depth] = 0;
    orc_cell_masks(ext, depth + 1, e->b, masks);
    ob.n = orbit_starts(masks[depth], e->b, ob.start, ob.len);
    const int remaining_after = e->v - depth - 1;
    for (int o = 0; o < ob.n; o++) {
        int s = ob.start[o], w = 0;
        uint32_t in = 0;
        for (int i = 0; i < depth; i++)
            if ((e->rows[i] >> s) & 1) { w++; in |= 1u << i; }
        ob.weight[o] = (u8)w;
        ob.inrow[o] = in;
        int need = e->k - w;               /* ones still missing in each column */
        if (need < 0 || need > remaining_after + 1) return;  /* dead */
        ob.hi[o] = need > 0 ? ob.len[o] : 0;
        ob.lo[o] = need > remaining_after ? ob.len[o] : 0;
    }
    int cap_r[ORC_MAX_COLS + 1];
    static __thread int cap_l[(ORC_MAX_COLS + 1) * ORC_MAX_ROWS];
    cap_r[ob.n] = 0;
    for (int i = 0; i < depth; i++) cap_l[ob.n * ORC_MAX_ROWS + i] = 0;
    for (int o = ob.n - 1; o >= 0; o--) {
        cap_r[o] = cap_r[o + 1] + ob.hi[o];
        for (int i = 0; i < depth; i++)
            cap_l[o * ORC_MAX_ROWS + i] = cap_l[(o + 1) * ORC_MAX_ROWS + i] +
                (((ob.inrow[o] >> i) & 1) ? ob.hi[o] : 0);
    }
    int x[ORC_MAX_COLS], lam[ORC_MAX_ROWS];
    memset(lam, 0, sizeof lam);
    /* cap arrays are indexed by depth-specific content: keep a private copy */
    int *capl = (int *)malloc(sizeof(int) * (ob.n + 1) * ORC_MAX_ROWS);
    memcpy(capl, cap_l, sizeof(int) * (ob.n + 1) * ORC_MAX_ROWS);
    solve_orbit(e, depth, &ob, 0, x, 0, lam, cap_r, capl);
    free(capl);
}

long orc_enumerate_bibd(int v, int k, int lambda, uint64_t *designs,
                        uint64_t *auts, long max_designs, uint64_t *stats,
                        uint64_t *cand_rows, uint32_t *cand_meta, long cand_cap,
                        int stride, int dmin, int dmax)
{
    if (v < 2 || k < 2 || k > v || lambda < 1) return -1;
    if ((lambda * (v - 1)) % (k - 1)) return -2;
    int r = lambda * (v - 1) / (k - 1);
    if ((v * r) % k) return -2;
    int b = v * r / k;
    if (v > ORC_MAX_ROWS || b > ORC_MAX_COLS) return -3;
    enum_t *e = (enum_t *)calloc(1, sizeof(enum_t));
    e->v = v; e->b = b; e->r = r; e->k = k; e->lambda = lambda;
    e->designs = designs; e->auts = auts; e->max_designs = max_designs;
    uint64_t st[5] = {0, 0, 0, 0, 0};
    e->stats = st;
    e->cand_rows = cand_rows; e->cand_meta = cand_meta; e->cand_cap = cand_cap;
    e->stride = stride; e->dmin = dmin; e->dmax = dmax;
    dfs_row(e, 0);
    long nd = e->ndesigns;
    if (stats) memcpy(stats, st, sizeof st);
    free(e);
    return nd;
}

/* Solutions of the next-row system of one partial matrix (rows[0..depth-1]),
 * in solver order (up to cap written); returns the number of solutions. */
long orc_row_solutions(int v, int k, int lambda, const uint64_t *rows, int depth, uint64_t *out, long cap)
{
    if (v < 2 || k < 2 || k > v || lambda < 1 || depth < 0 || depth >= v) return -1;
    if ((lambda * (v - 1)) % (k - 1)) return -2;
    int r = lambda * (v - 1) / (k - 1);
    if ((v * r) % k) return -2;
    int b = v * r / k;
    if (v > ORC_MAX_ROWS || b > ORC_MAX_COLS) return -3;
    enum_t *e = (enum_t *)calloc(1, sizeof(enum_t));
    e->v = v; e->b = b; e->r = r; e->k = k; e->lambda = lambda;
    uint64_t st[5] = {0, 0, 0, 0, 0};
    e->stats = st;
    memcpy(e->rows, rows, sizeof(uint64_t) * depth);
    e->collect = out; e->ncollect = 0; e->collect_cap = cap;
    expand(e, depth);
    long n = e->ncollect;
    free(e);
    return n;
}

/* ------------------------------------------------------------------------ */
/* Beam-sampled level-by-level enumeration.                                  */
/*
 * Level d holds the children (row-system solutions, in solver order) of the
 * canonical candidates kept at level d-1; every child is tested and those in
 * [dmin, dmax] are written out.  When more than `beam` candidates of a level
 * are canonical, `beam` of them are kept: a partial Fisher-Yates shuffle of
 * their indices driven by splitmix64 from `seed` (t = 0..beam-1: swap idx[t]
 * with idx[t + next() % (count - t)]), the chosen indices then sorted.  The
 * generator state carries over from level to level.  beam == 0 keeps all.
 */
static uint64_t splitmix_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

long orc_enumerate_sampled(int v, int k, int lambda, uint32_t beam, uint64_t seed, int dmin, int dmax,
                           int stride, uint64_t *cand_rows, uint32_t *cand_meta, long cand_cap,
                           uint64_t *level_counts)
{
    if (v < 2 || k < 2 || k > v || lambda < 1 || dmin < 1 || dmax < dmin || stride < dmax) return -1;
    if ((lambda * (v - 1)) % (k - 1)) return -2;
    int r = lambda * (v - 1) / (k - 1);
    if ((v * r) % k) return -2;
    int b = v * r / k;
    if (v > ORC_MAX_ROWS || b > ORC_MAX_COLS) return -3;
    enum_t *e = (enum_t *)calloc(1, sizeof(enum_t));
    e->v = v; e->b = b; e->r = r; e->k = k; e->lambda = lambda;
    uint64_t st[5] = {0, 0, 0, 0, 0};
    e->stats = st;
    size_t cap_par = 1024, npar = 1;
    uint64_t *par = (uint64_t *)calloc(cap_par * ORC_MAX_ROWS, sizeof(uint64_t));
    size_t cap_kid = 1 << 16;
    uint64_t *kid = (uint64_t *)malloc(cap_kid * ORC_MAX_ROWS * sizeof(uint64_t));
    uint64_t rng = seed, sols[4096];
    long nout = 0;
    for (int depth = 0; depth < v && npar; depth++) {
        /* children of every kept parent, in parent order */
        size_t nkid = 0;
        for (size_t p = 0; p < npar; p++) {
            memcpy(e->rows, par + p * ORC_MAX_ROWS, sizeof(uint64_t) * ORC_MAX_ROWS);
            e->collect = sols; e->ncollect = 0; e->collect_cap = 4096;
            expand(e, depth);
            e->collect = NULL;
            if (e->ncollect > 4096) { free(par); free(kid); free(e); return -4; }
            for (long q = 0; q < e->ncollect; q++) {
                if (nkid == cap_kid) {
                    cap_kid *= 2;
                    kid = (uint64_t *)realloc(kid, cap_kid * ORC_MAX_ROWS * sizeof(uint64_t));
                }
                uint64_t *dst = kid + nkid * ORC_MAX_ROWS;
                memcpy(dst, par + p * ORC_MAX_ROWS, sizeof(uint64_t) * ORC_MAX_ROWS);
                dst[depth] = sols[q];
                nkid++;
            }
        }
        const int nd = depth + 1;
        if (level_counts) level_counts[nd] = nkid;
        /* test every child; keep the canonical ones */
        uint32_t *canon = (uint32_t *)malloc((nkid + 1) * sizeof(uint32_t));
        size_t ncan = 0;
        for (size_t c = 0; c < nkid; c++) {
            const uint64_t *src = kid + c * ORC_MAX_ROWS;
            if (cand_rows && nd >= dmin && nd <= dmax && nout < cand_cap) {
                memset(cand_rows + (size_t)nout * stride, 0, sizeof(uint64_t) * stride);
                memcpy(cand_rows + (size_t)nout * stride, src, sizeof(uint64_t) * nd);
                unsigned oi = nd == v ? (ORC_SAVE_ROW_PERMUTATIONS | ORC_SAVE_ROW_TO_CHANGE) : 0;
                cand_meta[nout] = (uint32_t)nd | ((uint32_t)b << 8) | (oi << 24);
                nout++;
            }
            memcpy(e->rows, src, sizeof(uint64_t) * ORC_MAX_ROWS);
            if (check_partial(e, nd, NULL)) canon[ncan++] = (uint32_t)c;
        }
        if (beam && ncan > beam) {
            for (uint32_t t = 0; t < beam; t++) {
                size_t j = t + (size_t)(splitmix_next(&rng) % (uint64_t)(ncan - t));
                uint32_t x = canon[t]; canon[t] = canon[j]; canon[j] = x;
            }
            ncan = beam;
            qsort(canon, ncan, sizeof(uint32_t), cmp_u32);
        }
        if (ncan > cap_par) {
            cap_par = ncan;
            par = (uint64_t *)realloc(par, cap_par * ORC_MAX_ROWS * sizeof(uint64_t));
        }
        for (size_t q = 0; q < ncan; q++)
            memcpy(par + q * ORC_MAX_ROWS, kid + (size_t)canon[q] * ORC_MAX_ROWS, sizeof(uint64_t) * ORC_MAX_ROWS);
        npar = nd < v ? ncan : 0;
        free(canon);
    }
    free(par); free(kid); free(e);
    return nout;
}

/* ------------------------------------------------------------------------ */
/* Witness hunt in a proposed order (the learned-order search).              */
/*
 * Depth-first search over row prefixes in which position i tries the rows in
 * the order order[i][0..d-1] (rows already placed are skipped).  A row is
 * compared with the same orbit-by-orbit rule and column-permutation state as
 * the plain search (compare_orbit); "smaller" prunes the child, "larger" is a
 * witness (stop), "equal" descends.  At most `budget` children are compared.
 * No automorphism pruning: the hunt only looks for a witness; canonicity is
 * proved by the plain search.  Returns 1 with the witness in perm_out (prefix
 * perm[0..fail_row], then the unused rows in increasing order) or 0 when the
 * budget or the tree is exhausted; *nodes receives the children compared.
 */
static int hunt_core(const cand_t *cd, const uint8_t *order, uint32_t budget, u8 *perm_out,
                     u8 *pcol_out, u8 *fail_row, u8 *fail_col, uint32_t *cnt)
{
    const int d = cd->nrows, b = cd->ncols;
    const u8 *mat = cd->mat;
    u8 perm[ORC_MAX_ROWS], pcol[ORC_MAX_COLS];
    int cursor[ORC_MAX_ROWS + 1];
    uint32_t used = 0, n = 0;
    for (int q = 0; q < b; q++) pcol[q] = (u8)q;
    int i = 0, found = 0, idlen = 0;      /* perm[0..idlen-1] is the identity */
    u8 dec = 0;
    cursor[0] = 0;
    for (;;) {
        if (i == d) {                      /* the whole matrix matched: backtrack */
            i--;
            used &= ~(1u << perm[i]);
            if (idlen > i) idlen = i;
            continue;
        }
        /* entries naming no row of this candidate (>= d) are skipped */
        while (cursor[i] < d && (order ? (order[i * d + cursor[i]] >= d ||
                                          ((used >> order[i * d + cursor[i]]) & 1))
                                       : ((used >> cursor[i]) & 1)))
            cursor[i]++;
        if (cursor[i] == d) {
            if (i == 0) break;
            i--;
            used &= ~(1u << perm[i]);
            if (idlen > i) idlen = i;
            continue;
        }
        const u8 j = order ? order[i * d + cursor[i]] : (u8)cursor[i];
        cursor[i]++;
        perm[i] = j;
        if (j == i && idlen == i && n == 0) {
            /* identity prefix extended by its own row before any comparison
             * (the column permutation is still the identity): equal, and the
             * comparison would leave the columns in place */
            used |= 1u << j;
            idlen = ++i;
            cursor[i] = 0;
            continue;
        }
        if (n == budget) break;
        n++;
        cnt[0]++; cnt[1]++;
        const u8 *orig = mat + (size_t)i * b, *prow = mat + (size_t)j * b;
        int diff = 0, o;
        for (o = 0; o < cd->norb[i]; o++) {
            int s = cd->ostart[i][o], len = cd->olen[i][o];
            cnt[2]++;
            cnt[3] += (uint32_t)len;
            diff = len > 1 ? compare_orbit(s, len, orig, prow, pcol) : (int)orig[s] - (int)prow[pcol[s]];
            if (diff) break;
        }
        if (diff < 0) { found = 1; dec = cd->ostart[i][o]; break; }
        if (diff > 0) continue;
        used |= 1u << j;
        i++;
        cursor[i] = 0;
    }
    if (found) {
        uint32_t u = used | (1u << perm[i]);
        int k = i + 1;
        for (int r = 0; r < d; r++)
            if (!((u >> r) & 1)) perm[k++] = (u8)r;
        if (perm_out) memcpy(perm_out, perm, (size_t)d);
        if (pcol_out) memcpy(pcol_out, pcol, (size_t)b);
        if (fail_row) *fail_row = (u8)i;
        if (fail_col) *fail_col = dec;
    }
    return found;
}

int orc_hunt(const uint8_t *mat, int d, int b, const uint8_t *order, uint32_t budget,
             uint8_t *perm_out, uint8_t *fail_row, uint32_t *nodes)
{
    if (d < 1 || d > ORC_MAX_ROWS || b < 1 || b > ORC_MAX_COLS) return -1;
    cand_t *cd = (cand_t *)malloc(sizeof(cand_t));
    uint64_t rows[ORC_MAX_ROWS];
    for (int i = 0; i < d; i++) {
        rows[i] = 0;
        for (int q = 0; q < b; q++)
            if (mat[i * b + q]) rows[i] |= 1ull << q;
    }
    cd->nrows = d; cd->ncols = b; cd->len_stab = 0; cd->out_info = 0; cd->mat = mat;
    build_orbits(cd, NULL, rows);
    uint32_t cnt[4] = {0, 0, 0, 0};
    int found = hunt_core(cd, order, budget, perm_out, NULL, fail_row, NULL, cnt);
    if (nodes) *nodes = cnt[0];
    free(cd);
    return found;
}

int orc_check_batch_hunt(uint32_t n, uint32_t stride, const uint64_t *rows, const uint64_t *cells,
                         const uint32_t *meta, const uint8_t *order, uint32_t order_dim,
                         uint32_t budget, uint32_t gen_cap, orc_results *res)
{
    static __thread cand_t cd;
    static __thread u8 mat[ORC_MAX_ROWS * ORC_MAX_COLS];
    for (uint32_t c = 0; c < n; c++) {
        uint32_t m = meta[c];
        int d = m & 0xff, b = (m >> 8) & 0xff;
        if (d < 1 || d > ORC_MAX_ROWS || d > (int)stride || b < 1 || b > ORC_MAX_COLS)
            return -1;
        /* a table narrower than the candidate cannot guide it: plain search only */
        const int use_hunt = !order || d <= (int)order_dim;
        const uint64_t *rp = rows + (size_t)c * stride;
        cd.nrows = d; cd.ncols = b;
        cd.len_stab = (m >> 16) & 0xff;
        cd.out_info = (m >> 24) & 0xff;
        for (int i = 0; i < d; i++)
            for (int q = 0; q < b; q++)
                mat[i * b + q] = (u8)((rp[i] >> q) & 1);
        cd.mat = mat;
        build_orbits(&cd, cells ? cells + (size_t)c * stride : NULL, rp);
        /* the order table of a candidate is order_dim x order_dim; compact it to d x d */
        u8 ord[ORC_MAX_ROWS * ORC_MAX_ROWS];
        if (order && use_hunt)
            for (int i = 0; i < d; i++)
                for (int k = 0; k < d; k++)
                    ord[i * d + k] = order[(size_t)c * order_dim * order_dim + (size_t)i * order_dim + k];
        uint32_t cnt[4] = {0, 0, 0, 0};
        u8 perm[ORC_MAX_ROWS], pcol[ORC_MAX_COLS], frow = 255, fcol = 255;
        if (use_hunt && hunt_core(&cd, order ? ord : NULL, budget, perm, pcol, &frow, &fcol, cnt)) {
            if (res->verdict) res->verdict[c] = 0;
            if (res->fail_row) res->fail_row[c] = frow;
            if (res->fail_col) res->fail_col[c] = fcol;
            if (res->row_to_change) res->row_to_change[c] = row_to_change(perm, frow);
            if (res->group_order) res->group_order[c] = 0;
            if (res->group_order_hi) res->group_order_hi[c] = 0;
            if (res->perm_row) {
                memset(res->perm_row + c * ORC_MAX_ROWS, 0xff, ORC_MAX_ROWS);
                memcpy(res->perm_row + c * ORC_MAX_ROWS, perm, (size_t)d);
            }
            if (res->perm_col) {
                memset(res->perm_col + c * ORC_MAX_COLS, 0xff, ORC_MAX_COLS);
                memcpy(res->perm_col + c * ORC_MAX_COLS, pcol, (size_t)b);
            }
            if (res->orbits) {
                memset(res->orbits + c * ORC_MAX_ROWS, 0xff, ORC_MAX_ROWS);
                for (int i = 0; i < d; i++) res->orbits[c * ORC_MAX_ROWS + i] = (u8)i;
            }
            if (res->n_gens) res->n_gens[c] = 0;
            if (res->counters) memcpy(res->counters + c * 4, cnt, sizeof cnt);
        } else {
            run_check_from(&cd, res, c, gen_cap, cnt);
        }
    }
    return 0;
}

/* ---- lead-order search (the learned order of the first row) --------------
 *
 * The project's learned order, not a reference algorithm.  Until the walk of
 * next_row_perm meets its first automorphism it is a depth-first search over
 * row prefixes in increasing row order (lexicographic order of the row
 * permutations, GroupOrder.h:136-248 with trivial orbits), pruned where a
 * prefix compares smaller.  The set of prefixes that compare equal does not
 * depend on the order in which the rows of a position are tried, so a
 * depth-first search that tries the rows of position 0 in another order and
 * the rows of every later position in increasing order
 *   - meets an automorphism iff the group is not trivial (the search then
 *     stops and the candidate takes the plain search),
 *   - ends without a witness and without an automorphism iff the matrix is
 *     canonical with a trivial group (the plain search's results), and
 *   - otherwise stops at a witness of its own order (a larger matrix, first
 *     at *frow).
 * The node "prefix 0..i-1 is the identity, row i at position i" is equal by
 * construction and not counted (the plain walk starts on the identity and
 * never compares it); it still partitions the columns as a comparison would.
 * lead: the rows of position 0 in order (entries >= d skipped).  Returns 0
 * (exhausted), 1 (witness: perm_out, *frow, *fcol) or 2 (automorphism). */
static int lead_dfs(const cand_t *cd, const u8 *lead, int nlead, u8 *perm_out, u8 *frow, u8 *fcol,
                    uint32_t *cnt)
{
    const int d = cd->nrows, b = cd->ncols;
    const u8 *mat = cd->mat;
    u8 perm[ORC_MAX_ROWS], pcol[ORC_MAX_COLS];
    int cursor[ORC_MAX_ROWS + 1];
    uint32_t used = 0, tried = 0;
    for (int q = 0; q < b; q++) pcol[q] = (u8)q;
    int i = 0, idlen = 0, desc = 0;               /* desc: the last step descended */
    cursor[0] = 0;
    for (;;) {
        if (i == d) {
            if (idlen < d) return 2;               /* a non-identity permutation matched */
            i--;                                   /* the identity: back up */
            used &= ~(1u << perm[i]);
            idlen = i;
            desc = 0;
            continue;
        }
        int j = -1;
        if (i == 0) {
            /* the listed rows (each once, rows >= d skipped), then the others in increasing order */
            while (j < 0 && cursor[0] < nlead + d) {
                const int k = cursor[0]++;
                const int r = k < nlead ? lead[k] : k - nlead;
                if (r < d && !((tried >> r) & 1)) j = r;
            }
            if (j >= 0) tried |= 1u << j;
        } else {
            while (cursor[i] < d && ((used >> cursor[i]) & 1)) cursor[i]++;
            if (cursor[i] < d) j = cursor[i]++;
        }
        if (j < 0) {
            if (i == 0) return 0;
            i--;
            used &= ~(1u << perm[i]);
            if (idlen > i) idlen = i;
            desc = 0;
            continue;
        }
        perm[i] = (u8)j;
        const int ident = idlen == i && j == i;
        /* the walk counts a permutation per new row after a back-up or a
         * refutation; a descent continues the permutation */
        if (!ident) { cnt[0] += !desc; cnt[1]++; }
        desc = 0;
        const u8 *orig = mat + (size_t)i * b, *prow = mat + (size_t)j * b;
        int diff = 0, o;
        for (o = 0; o < cd->norb[i]; o++) {
            int s = cd->ostart[i][o], len = cd->olen[i][o];
            if (!ident) { cnt[2]++; cnt[3] += (uint32_t)len; }
            diff = len > 1 ? compare_orbit(s, len, orig, prow, pcol) : (int)orig[s] - (int)prow[pcol[s]];
            if (diff) break;
        }
        if (diff > 0) continue;
        if (diff < 0) {
            uint32_t u = used | (1u << j);
            int k = i + 1;
            for (int r = 0; r < d; r++)
                if (!((u >> r) & 1)) perm[k++] = (u8)r;
            memcpy(perm_out, perm, (size_t)d);
            *frow = (u8)i;
            *fcol = cd->ostart[i][o];
            return 1;
        }
        used |= 1u << j;
        if (ident) idlen = i + 1;
        i++;
        cursor[i] = 0;
        desc = 1;
    }
}

/* Orbit form (what the column-set search needs): inside every column orbit
 * of a row the ones come first.  With given cell masks, the orbits of row
 * i+1 must also be those of row i split after its ones. */
static int cand_orbit_form(const cand_t *cd, const uint64_t *cells_in, const uint64_t *rows)
{
    const int d = cd->nrows, b = cd->ncols;
    for (int i = 0; i < d; i++)
        for (int o = 0; o < cd->norb[i]; o++) {
            const u8 *r = cd->mat + (size_t)i * b + cd->ostart[i][o];
            for (int q = 1; q < cd->olen[i][o]; q++)
                if (r[q] > r[q - 1]) return 0;
        }
    if (cells_in) {
        const uint64_t full = b >= 64 ? ~0ull : ((1ull << b) - 1);
        for (int i = 0; i + 1 < d; i++) {
            uint64_t ci = (cells_in[i] | 1ull) & full, cn = (cells_in[i + 1] | 1ull) & full;
            if (cn != ((ci | (rows[i] ^ (rows[i] << 1))) & full)) return 0;
        }
    }
    return 1;
}

/* Features of the rows of one candidate as the row of position 0 (16 per row,
 * small integers scaled by powers of two, exact in bf16):
 * with w(c) = the number of rows holding a one in column c and m = max w,
 * A_j[s] = ones of row j in columns of weight m - s (s = 0, 1, 2) and of
 * weight <= m - 3 (s = 3):
 *   0-3  A_j / 16            4-7  (A_j - A_0) / 16
 *   8    sign of A_j against A_0 (lexicographic, +1 when larger)
 *   9    j == 0              10   j == d - 1
 *   11   j / 16              12   d / 16
 *   13   largest weight of a column shared with row 0 (0 for row 0) / 16
 *   14   largest weight of a column shared with row d-1 (0 for it) / 16
 *   15   rows with the same A as row j / 16 */
static void lead_features_one(const uint64_t *R, int d, int b, float *F)
{
    int w[ORC_MAX_COLS];
    int m = 0;
    for (int c = 0; c < b; c++) {
        w[c] = 0;
        for (int i = 0; i < d; i++) w[c] += (int)((R[i] >> c) & 1);
        if (w[c] > m) m = w[c];
    }
    int A[ORC_MAX_ROWS][4];
    for (int j = 0; j < d; j++) {
        for (int s = 0; s < 4; s++) A[j][s] = 0;
        for (int c = 0; c < b; c++) {
            if (!((R[j] >> c) & 1)) continue;
            int s = m - w[c];
            A[j][s < 3 ? s : 3]++;
        }
    }
    for (int j = 0; j < d; j++) {
        float *f = F + (size_t)j * 16;
        int sg = 0;
        for (int s = 0; s < 4 && !sg; s++) sg = A[j][s] > A[0][s] ? 1 : (A[j][s] < A[0][s] ? -1 : 0);
        int m0 = 0, mn = 0, ties = 0;
        for (int c = 0; c < b; c++) {
            if (j != 0 && ((R[j] & R[0]) >> c) & 1 && w[c] > m0) m0 = w[c];
            if (j != d - 1 && ((R[j] & R[d - 1]) >> c) & 1 && w[c] > mn) mn = w[c];
        }
        for (int q = 0; q < d; q++)
            ties += A[q][0] == A[j][0] && A[q][1] == A[j][1] && A[q][2] == A[j][2] && A[q][3] == A[j][3];
        for (int s = 0; s < 4; s++) {
            f[s] = (float)A[j][s] / 16.f;
            f[4 + s] = (float)(A[j][s] - A[0][s]) / 16.f;
        }
        f[8] = (float)sg;
        f[9] = j == 0 ? 1.f : 0.f;
        f[10] = j == d - 1 ? 1.f : 0.f;
        f[11] = (float)j / 16.f;
        f[12] = (float)d / 16.f;
        f[13] = (float)m0 / 16.f;
        f[14] = (float)mn / 16.f;
        f[15] = (float)ties / 16.f;
    }
}

int orc_lead_features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta, float *X)
{
    memset(X, 0, sizeof(float) * (size_t)n * 16 * 16);
    for (uint32_t c = 0; c < n; c++) {
        const int d = meta[c] & 0xff, b = (meta[c] >> 8) & 0xff;
        if (d < 1 || d > 16 || d > (int)stride || b < 1 || b > ORC_MAX_COLS) continue;
        lead_features_one(rows + (size_t)c * stride, d, b, X + (size_t)c * 256);
    }
    return 0;
}

int orc_check_batch_lead(uint32_t n, uint32_t stride, const uint64_t *rows, const uint64_t *cells,
                         const uint32_t *meta, const uint8_t *lead, uint32_t gate_perms, uint32_t gen_cap,
                         orc_results *res, uint8_t *mode)
{
    static __thread cand_t cd;
    static __thread u8 mat[ORC_MAX_ROWS * ORC_MAX_COLS];
    for (uint32_t c = 0; c < n; c++) {
        uint32_t m = meta[c];
        int d = m & 0xff, b = (m >> 8) & 0xff;
        if (d < 1 || d > ORC_MAX_ROWS || d > (int)stride || b < 1 || b > ORC_MAX_COLS)
            return -1;
        const uint64_t *rp = rows + (size_t)c * stride;
        const uint64_t *cp = cells ? cells + (size_t)c * stride : NULL;
        cd.nrows = d; cd.ncols = b;
        cd.len_stab = (m >> 16) & 0xff;
        cd.out_info = (m >> 24) & 0xff;
        for (int i = 0; i < d; i++)
            for (int q = 0; q < b; q++)
                mat[i * b + q] = (u8)((rp[i] >> q) & 1);
        cd.mat = mat;
        build_orbits(&cd, cp, rp);
        run_check(&cd, res, c, gen_cap);
        uint8_t md = 0;
        const int eligible = lead && d >= 2 && d <= 16 && cd.len_stab == 0 && b <= 56 &&
                             cand_orbit_form(&cd, cp, rp);
        if (eligible) {
            if (!res->counters)            /* the gate reads the plain search's permutations */
                return -2;
            uint32_t plain_cnt[4];
            memcpy(plain_cnt, res->counters + (size_t)c * 4, sizeof plain_cnt);
            md = 1;
            if (plain_cnt[0] > gate_perms) {
                /* the gate: the walk's first gate_perms permutations */
                uint32_t gate_cnt[4] = {0, 0, 0, 0};
                uint16_t gate_ngen = 0;
                uint8_t gv = 0;
                orc_results gr;
                memset(&gr, 0, sizeof gr);
                gr.verdict = &gv;
                gr.counters = gate_cnt;
                gr.n_gens = &gate_ngen;
                walk_limit = gate_perms;
                run_check(&cd, &gr, 0, 0);
                walk_limit = 0xffffffffu;
                if (gate_ngen > 0) {
                    md = 4;                        /* an automorphism within the gate: plain search */
                } else {
                    uint32_t cnt[4];
                    memcpy(cnt, gate_cnt, sizeof cnt);
                    u8 perm[ORC_MAX_ROWS], frow = 255, fcol = 255;
                    int r = lead_dfs(&cd, lead + (size_t)c * 16, 16, perm, &frow, &fcol, cnt);
                    if (r == 1) {
                        md = 2;
                        if (res->verdict) res->verdict[c] = 0;
                        if (res->fail_row) res->fail_row[c] = frow;
                        if (res->fail_col) res->fail_col[c] = fcol;
                        if (res->row_to_change) res->row_to_change[c] = row_to_change(perm, frow);
                        if (res->group_order) res->group_order[c] = 0;
                        if (res->group_order_hi) res->group_order_hi[c] = 0;
                        if (res->perm_row) {
                            memset(res->perm_row + c * ORC_MAX_ROWS, 0xff, ORC_MAX_ROWS);
                            memcpy(res->perm_row + c * ORC_MAX_ROWS, perm, (size_t)d);
                        }
                        if (res->orbits) {
                            memset(res->orbits + c * ORC_MAX_ROWS, 0xff, ORC_MAX_ROWS);
                            for (int i = 0; i < d; i++) res->orbits[c * ORC_MAX_ROWS + i] = (u8)i;
                        }
                        if (res->n_gens) res->n_gens[c] = 0;
                    } else {
                        md = r == 0 ? 3 : 4;
                    }
                    /* a candidate that met an automorphism also runs the whole plain search */
                    if (r == 2)
                        for (int q = 0; q < 4; q++) cnt[q] += plain_cnt[q];
                    memcpy(res->counters + (size_t)c * 4, cnt, sizeof cnt);
                }
            }
        }
        if (mode) mode[c] = md;
    }
    return 0;
}
