/*
 * ise_oracle_cdtools.c -- CPU oracle: the k-Sys binary canonizer
 * (createCanonizer / runCanonizer, Source/CDTools.cpp:5-51).
 *
 * TEST INFRASTRUCTURE ONLY (see ise_oracle.h).
 *
 * runCanonizer turns a k-system given as days of v players (each day lists its
 * v/k groups of k players) into the (v + 1) x b matrix of
 * C_InSys::prepareFirstMatrixRow / convertToBinaryMatrix (Source/matrix.h:
 * 192-210): row 0 holds the day labels (the columns of day j hold
 * days - 1 - j), row 1 + p the incidence of player p, column j * v/k + g group
 * g of day j.  CMatrixCanonChecker::CanonizeMatrix(k, NULL, numClasses)
 * (Source/MatrixCanonChecker.cpp:420-564) then
 *   - repeats the canonicity test with row permutations saved on the matrix
 *     rearranged by each witness (rows 1..v move, row 0 stays: a player row at
 *     position 0 holds none of the largest label and loses at once), which
 *     ends on the largest matrix over the permutations of the player rows and
 *     of the columns inside each day (the columns of one label form the first
 *     row's column orbits);
 *   - when every day is present (numClasses = days), walks the permutations
 *     of the days (CGroupOrder::next_permutation with the automorphisms of the
 *     best matrix so far), canonizes each, keeps the larger matrix and starts
 *     the walk again from the identity whenever the best matrix grows
 *     (initOrbits after cmp == 1; the 0/1 rows make memcmp's result +-1).  A
 *     walk that ends has found no day permutation that gives a larger matrix,
 *     so the result is the largest matrix over the day permutations too.
 * The function returns rows 1..v, the v x b player rows.
 *
 * Restatement: the label row is the initial column partition of the
 * canonicity test (its orbits: the days, day 0 first), given to the test as
 * the orbit-start masks of every level; the v player rows are canonized under
 * it (orc_canonize_part), and for full days every day order is canonized and
 * the largest result kept.
 */
#include "ise_oracle.h"
#include <string.h>

typedef uint8_t u8;

/* Columns in decreasing order inside each block of `part` (bit c set: column
 * c starts a block), row 0 most significant: an orbit form of the matrix under
 * that initial partition (stable, so equal columns keep their order). */
static void block_sort(u8 *mat, int nrows, int ncols, uint64_t part)
{
    uint32_t key[ORC_MAX_COLS];
    int idx[ORC_MAX_COLS];
    for (int q = 0; q < ncols; q++) {
        uint32_t x = 0;
        for (int i = 0; i < nrows; i++) x = (x << 1) | mat[i * ncols + q];
        key[q] = x;
        idx[q] = q;
    }
    for (int lo = 0; lo < ncols;) {
        int hi = lo + 1;
        while (hi < ncols && !((part >> hi) & 1ull)) hi++;
        for (int a = lo + 1; a < hi; a++)          /* insertion sort, decreasing */
            for (int c = a; c > lo && key[idx[c - 1]] < key[idx[c]]; c--) {
                int t = idx[c]; idx[c] = idx[c - 1]; idx[c - 1] = t;
            }
        lo = hi;
    }
    u8 tmp[ORC_MAX_ROWS * ORC_MAX_COLS];
    for (int i = 0; i < nrows; i++)
        for (int q = 0; q < ncols; q++) tmp[i * ncols + q] = mat[i * ncols + idx[q]];
    memcpy(mat, tmp, (size_t)(nrows * ncols));
}

int orc_canonize_part(const uint8_t *mat, int nrows, int ncols, uint64_t part, int max_iters, uint8_t *out,
                      uint64_t *aut, uint32_t *iters)
{
    if (nrows < 1 || nrows > ORC_MAX_ROWS || ncols < 1 || ncols > ORC_MAX_COLS) return -1;
    const uint64_t full = ncols >= 64 ? ~0ull : ((1ull << ncols) - 1);
    part = (part | 1ull) & full;
    u8 cur[ORC_MAX_ROWS * ORC_MAX_COLS], tmp[ORC_MAX_ROWS * ORC_MAX_COLS];
    memcpy(cur, mat, (size_t)(nrows * ncols));
    block_sort(cur, nrows, ncols, part);
    uint32_t it = 0;
    int done = 0;
    while (it < (uint32_t)max_iters) {
        uint64_t rows[ORC_MAX_ROWS], cells[ORC_MAX_ROWS];
        uint64_t m = part;
        for (int i = 0; i < nrows; i++) {
            uint64_t r = 0;
            for (int q = 0; q < ncols; q++) r |= (uint64_t)cur[i * ncols + q] << q;
            rows[i] = r;
            cells[i] = m;
            m = (m | (r ^ (r << 1))) & full;
        }
        const uint32_t meta = (uint32_t)nrows | ((uint32_t)ncols << 8) | ((uint32_t)ORC_SAVE_ROW_PERMUTATIONS << 24);
        uint8_t verdict = 0, prow[ORC_MAX_ROWS];
        uint64_t order = 0;
        orc_results res;
        memset(&res, 0, sizeof res);
        res.verdict = &verdict;
        res.group_order = &order;
        res.perm_row = prow;
        if (orc_check_batch(1, (uint32_t)nrows, rows, cells, &meta, 0, &res) < 0) return -1;
        it++;
        if (verdict == 1) {
            if (aut) *aut = order;
            done = 1;
            break;
        }
        for (int i = 0; i < nrows; i++)
            memcpy(tmp + i * ncols, cur + prow[i] * ncols, (size_t)ncols);
        memcpy(cur, tmp, (size_t)(nrows * ncols));
        block_sort(cur, nrows, ncols, part);
    }
    memcpy(out, cur, (size_t)(nrows * ncols));
    if (iters) *iters = it;
    if (!done && aut) *aut = 0;
    return done;
}

/* The v x b player rows of `days` days taken in the order `order` (position j
 * holds day order[j]); column j * v/k + g is group g of that day. */
static void cd_binary(const uint8_t *mat, int v, int k, int days, const int *order, uint8_t *M)
{
    const int G = v / k, b = days * G;
    memset(M, 0, (size_t)(v * b));
    for (int j = 0; j < days; j++) {
        const uint8_t *day = mat + order[j] * v;
        for (int g = 0; g < G; g++)
            for (int n = 0; n < k; n++) M[day[g * k + n] * b + j * G + g] = 1;
    }
}

/* Lexicographic successor of order[0..n-1]; 0 after the last. */
static int next_order(int *a, int n)
{
    int i = n - 2;
    while (i >= 0 && a[i] >= a[i + 1]) i--;
    if (i < 0) return 0;
    int j = n - 1;
    while (a[j] <= a[i]) j--;
    int t = a[i]; a[i] = a[j]; a[j] = t;
    for (int lo = i + 1, hi = n - 1; lo < hi; lo++, hi--) { t = a[lo]; a[lo] = a[hi]; a[hi] = t; }
    return 1;
}

uint64_t orc_cd_day_starts(int days, int groups)
{
    uint64_t p = 0;
    for (int j = 0; j < days; j++) p |= 1ull << (j * groups);
    return p;
}

int orc_cd_canonize_orders(const uint8_t *mat, int v, int k, int days, const uint8_t *orders, long n_orders,
                           uint8_t *best)
{
    if (k < 2 || v < k || v % k || days < 1) return -1;
    const int G = v / k, b = days * G;
    if (v > ORC_MAX_ROWS || b > ORC_MAX_COLS) return -1;
    const uint64_t part = orc_cd_day_starts(days, G);
    u8 M[ORC_MAX_ROWS * ORC_MAX_COLS], C[ORC_MAX_ROWS * ORC_MAX_COLS];
    int order[ORC_MAX_COLS];
    for (long t = 0; t < n_orders; t++) {
        for (int j = 0; j < days; j++) order[j] = orders[t * days + j];
        cd_binary(mat, v, k, days, order, M);
        if (orc_canonize_part(M, v, b, part, 1 << 16, C, NULL, NULL) != 1) return -2;
        if (!t || memcmp(C, best, (size_t)(v * b)) > 0) memcpy(best, C, (size_t)(v * b));
    }
    return 0;
}

int orc_cd_run_canonizer_ex(const uint8_t *mat, int v, int k, int day, uint8_t *out, int *ncols,
                            uint64_t *day_order, uint64_t *fixed_order)
{
    if (k < 2 || v < k || v % k || (v - 1) % (k - 1)) return -1;
    const int days_max = (v - 1) / (k - 1);
    const int days = day <= 0 ? days_max : day;
    if (days > days_max) return -1;
    const int G = v / k, b = days * G;
    if (v > ORC_MAX_ROWS || b > ORC_MAX_COLS) return -1;
    for (int i = 0; i < days * v; i++)
        if (mat[i] >= v) return -1;
    const uint64_t part = orc_cd_day_starts(days, G);
    int order[ORC_MAX_COLS];
    for (int j = 0; j < days; j++) order[j] = j;
    u8 M[ORC_MAX_ROWS * ORC_MAX_COLS], C[ORC_MAX_ROWS * ORC_MAX_COLS], best[ORC_MAX_ROWS * ORC_MAX_COLS];
    int have = 0;
    uint64_t reach = 0, fixed = 0;
    const int walk = days == days_max;   /* numClasses = days only when every day is present */
    do {
        cd_binary(mat, v, k, days, order, M);
        uint64_t aut = 0;
        if (orc_canonize_part(M, v, b, part, 1 << 16, C, &aut, NULL) != 1) return -2;
        const int c = have ? memcmp(C, best, (size_t)(v * b)) : 1;
        if (c > 0) {
            memcpy(best, C, (size_t)(v * b));
            have = 1;
            reach = 1;
            fixed = aut;
        } else if (c == 0) {
            reach++;
        }
    } while (walk && next_order(order, days));
    memcpy(out, best, (size_t)(v * b));
    if (ncols) *ncols = b;
    if (day_order) *day_order = reach;
    if (fixed_order) *fixed_order = fixed;
    return 0;
}

int orc_cd_run_canonizer(const uint8_t *mat, int v, int k, int day, uint8_t *out, int *ncols)
{
    return orc_cd_run_canonizer_ex(mat, v, k, day, out, ncols, NULL, NULL);
}
