/*
 * ise_oracle_kfact.c -- CPU oracle: canonicity of partial 1-factorizations of
 * K_n (groups of size 2, complete graph).
 *
 * TEST INFRASTRUCTURE ONLY (see ise_oracle.h).
 *
 * Restates alldata::canonizator / cnvCheckTgNew / cnvCheckKm / cnvCheckKm1
 * (k-Factorization/TripleSys/Source/cnv.cpp:65-351) with the generic normal
 * form kmProcessMatrix / kmSortMatrixForReorderedPlayers / kmSortGroups /
 * kmSortGroupsByFirstValue / kmSortRowsBy2ndValue
 * (k-Factorization/TripleSys/Source/kmSupport.cpp:43-166, 220-313), for the
 * default switches USE_EQUAL == 0 (TripleSys.h:79) and early exit on the first
 * smaller matrix (m_doNotExitEarlyIfNotCanonical == false, TripleSys.h:387).
 *
 * A matrix is nrounds rows of n players; each row lists the n/2 pairs of one
 * round.  The relabellings tried, in order:
 *   for every arrangement of the n/2 pair slots (Fuchs' countdown QuickPerm
 *   order, cnv.cpp:283-315), for every choice of orientation of each pair
 *   (odometer over {0,1}^(n/2), last pair fastest, cnv.cpp:193-208), for every
 *   day d (0..nrounds-1, cnv.cpp:91-163): the relabelling that sends the pair
 *   in slot i of round d to the labels (a[i], a[i]+1) in that orientation.
 * Each relabelled matrix is normalised (pairs sorted, pairs of a row sorted by
 * their first element, rows sorted by the partner of player 0) and compared
 * row by row with the original: smaller -> not canonical (stop), equal ->
 * automorphism.
 */
#include "ise_oracle.h"
#include <string.h>
#include <stdlib.h>

typedef unsigned char u8;
#define KF_MAX_N 16

/* Fuchs' countdown QuickPerm: writes all k! arrangements of 0..k-1 (scaled by
 * 2: the base label of each pair slot) in visiting order; returns k!. */
static long quickperm_table(int k, u8 *out)
{
    u8 a[KF_MAX_N / 2 + 1], p[KF_MAX_N / 2 + 1];
    long cnt = 0;
    for (int i = 0; i < k; i++) { a[i] = (u8)(2 * i); p[i] = (u8)i; }
    p[k] = (u8)k;
    memcpy(out + cnt++ * k, a, (size_t)k);
    int i = 1;
    while (i < k) {
        p[i]--;
        int j = (i % 2) ? p[i] : 0;
        u8 t = a[i]; a[i] = a[j]; a[j] = t;
        memcpy(out + cnt++ * k, a, (size_t)k);
        i = 0;
        while (!p[++i])
            p[i] = (u8)i;
    }
    return cnt;
}

/* partner arrays part[r][x] of an nr x n matrix; 0 unless every round is a
 * perfect matching of players 0..n-1 and no pair appears in two rounds (the
 * checks the device path's host side makes before a launch) */
static int kf_partners(const u8 *mat, int n, int nr, u8 part[][KF_MAX_N])
{
    u8 met[KF_MAX_N][KF_MAX_N];
    memset(met, 0, sizeof met);
    for (int r = 0; r < nr; r++) {
        int seen = 0;
        for (int q = 0; q < n; q += 2) {
            const int x = mat[r * n + q], y = mat[r * n + q + 1];
            if (x >= n || y >= n || x == y || (seen >> x & 1) || (seen >> y & 1) || met[x][y]) return 0;
            seen |= (1 << x) | (1 << y);
            met[x][y] = met[y][x] = 1;
            part[r][x] = (u8)y;
            part[r][y] = (u8)x;
        }
    }
    return 1;
}

/* normal form of matrix mi (nr x n) relabelled by tr, into mo */
static void normal_form(const u8 *mi, int nr, int n, const u8 *tr, u8 *mo)
{
    u8 tmp[KF_MAX_N * KF_MAX_N];
    const int npairs = n / 2;
    for (int r = 0; r < nr; r++) {
        u8 byfirst[KF_MAX_N][2];
        u8 present[KF_MAX_N];
        memset(present, 0, sizeof present);
        for (int q = 0; q < npairs; q++) {
            u8 x = tr[mi[r * n + 2 * q]], y = tr[mi[r * n + 2 * q + 1]];
            if (x > y) { u8 t = x; x = y; y = t; }
            byfirst[x][0] = x; byfirst[x][1] = y; present[x] = 1;
        }
        int k = 0;
        for (int v = 0; v < n; v++)
            if (present[v]) { tmp[r * n + k++] = byfirst[v][0]; tmp[r * n + k++] = byfirst[v][1]; }
    }
    /* rows by their second entry (partner of player 0) */
    int k = 0;
    for (int key = 0; key < n; key++)
        for (int r = 0; r < nr; r++)
            if (tmp[r * n + 1] == key) { memcpy(mo + k * n, tmp + r * n, (size_t)n); k++; }
}

void orc_kf_normalize(uint8_t *mat, int n, int nrounds)
{
    u8 id[KF_MAX_N], out[KF_MAX_N * KF_MAX_N];
    for (int i = 0; i < n; i++) id[i] = (u8)i;
    normal_form(mat, nrounds, n, id, out);
    memcpy(mat, out, (size_t)(n * nrounds));
}

/* The relabelling (player -> label) of the last witness found, for the
 * player index below. */
static __thread u8 witness_tr[KF_MAX_N];
static __thread int witness_valid;

static void keep_witness(const u8 *tr, int n)
{
    memcpy(witness_tr, tr, (size_t)n);
    witness_valid = 1;
}

int orc_kf_check(const uint8_t *mat, int n, int nrounds, uint64_t *witness,
                 uint64_t *n_auts, uint64_t *n_tried)
{
    witness_valid = 0;
    if (n < 4 || n > KF_MAX_N || (n & 1) || nrounds < 1 || nrounds > n - 1) return -1;
    u8 part[KF_MAX_N][KF_MAX_N];
    if (!kf_partners(mat, n, nrounds, part)) return -1;
    const int npairs = n / 2;
    long nperm = 1;
    for (int i = 2; i <= npairs; i++) nperm *= i;
    u8 *qp = (u8 *)malloc((size_t)nperm * npairs);
    quickperm_table(npairs, qp);
    u8 trmk[KF_MAX_N], ttr[KF_MAX_N], mo[KF_MAX_N * KF_MAX_N];
    uint64_t auts = 0, tried = 0, idx = 0;
    int verdict = 1;
    for (long h = 0; h < nperm && verdict; h++) {
        const u8 *a = qp + h * npairs;
        for (unsigned t = 0; t < (1u << npairs) && verdict; t++) {
            for (int i = 0; i < npairs; i++) {
                /* orientation bit of slot i: the odometer's last digit is the fastest */
                int flip = (t >> (npairs - 1 - i)) & 1;
                trmk[2 * i] = (u8)(a[i] + flip);
                trmk[2 * i + 1] = (u8)(a[i] + 1 - flip);
            }
            for (int day = 0; day < nrounds; day++, idx++) {
                const u8 *rd = mat + day * n;
                for (int pos = 0; pos < n; pos++) ttr[rd[pos]] = trmk[pos];
                normal_form(mat, nrounds, n, ttr, mo);
                tried++;
                int cmp = memcmp(mo, mat, (size_t)(n * nrounds));
                /* row-wise early exit gives the same sign as a full compare */
                if (cmp == 0) auts++;
                else if (cmp < 0) {
                    verdict = 0;
                    if (witness) *witness = idx;
                    keep_witness(ttr, n);
                    break;
                }
            }
        }
    }
    free(qp);
    if (n_auts) *n_auts = auts;
    if (n_tried) *n_tried = tried;
    return verdict;
}

/* Brute force for tiny n: canonical <=> no relabelling by any permutation of
 * the players gives a smaller normal form; also counts the permutations that
 * reproduce the matrix. */
int orc_kf_bruteforce(const uint8_t *mat, int n, int nrounds, uint64_t *n_auts)
{
    u8 part[KF_MAX_N][KF_MAX_N];
    if (n < 4 || n > 8 || (n & 1) || nrounds < 1 || nrounds > n - 1 || !kf_partners(mat, n, nrounds, part))
        return -1;
    u8 tr[KF_MAX_N], mo[KF_MAX_N * KF_MAX_N];
    int c[KF_MAX_N], i = 0, canon = 1;
    uint64_t auts = 0;
    for (int q = 0; q < n; q++) { tr[q] = (u8)q; c[q] = 0; }
    for (;;) {
        normal_form(mat, nrounds, n, tr, mo);
        int cmp = memcmp(mo, mat, (size_t)(n * nrounds));
        if (cmp < 0) canon = 0;
        if (cmp == 0) auts++;
        while (i < n && c[i] >= i) { c[i] = 0; i++; }
        if (i >= n) break;
        u8 t;
        if (i % 2 == 0) { t = tr[0]; tr[0] = tr[i]; tr[i] = t; }
        else { t = tr[c[i]]; tr[c[i]] = tr[i]; tr[i] = t; }
        c[i]++;
        i = 0;
    }
    if (n_auts) *n_auts = auts;
    return canon;
}

/*
 * Two-row canonizer for perfect 1-factorizations: alldata::cnvCheck2P1F
 * (k-Factorization/TripleSys/Source/cnvSupport.cpp:3-95) with create2P1FTr
 * (p1fSupport.cpp:3-38) and kmProcessMatrix2p1f / kmProcessOneNot1stRow2
 * (kmSupport.cpp:357-381, 519-598), for the defaults of the P1F runs
 * (AutSaveTestedTrs 0, no row precalculation, early exit on the first
 * smaller matrix).
 *
 * Relabellings, in order: for iRowLast = 1..nr-1, indRow = 0..iRowLast-1,
 * swap = 0, 1 ((r0, r1) = (indRow, iRowLast), then (iRowLast, indRow)), start
 * player k = 0..n-1: walk the cycle of rounds r0, r1 from k and the cycle of
 * rounds 0, 1 from player 0 side by side, sending the players of the first to
 * the players of the second.  Rounds r0, r1 become rounds 0, 1 and are not
 * compared.  Every other round is relabelled, normalised (pairs sorted, pairs
 * by first element) and placed by its second entry v (the partner of player
 * 0): rounds with v = 3..nr are compared, in order of v, with rounds v-1 of the
 * matrix as given (memcmp); a missing v in 3..nr makes the relabelling not
 * smaller (return 2 in the reference).  Fewer than 3 rounds: every relabelling
 * is equal.  |Aut| = relabellings equal before the decision (each is a distinct
 * player permutation).  Returns 1 canonical, 0 not, -1 bad arguments or a pair
 * of rounds that is not one Hamiltonian cycle (not a P1F).
 */
static int kf_is_p1f_pair(const u8 *pa, const u8 *pb, int n)
{
    int v = 0, len = 0;
    do { v = pb[pa[v]]; len += 2; } while (v != 0 && len <= n);
    return len == n;
}

int orc_kf_check_2p1f(const uint8_t *mat, int n, int nrounds, uint64_t *witness,
                      uint64_t *n_auts, uint64_t *n_tried)
{
    if (n < 4 || n > KF_MAX_N || (n & 1) || nrounds < 1 || nrounds > n - 1) return -1;
    u8 part[KF_MAX_N][KF_MAX_N];
    if (!kf_partners(mat, n, nrounds, part)) return -1;
    for (int a = 0; a < nrounds; a++)
        for (int b = a + 1; b < nrounds; b++)
            if (!kf_is_p1f_pair(part[a], part[b], n)) return -1;
    witness_valid = 0;
    uint64_t auts = 0, tried = 0, idx = 0;
    int verdict = 1;
    if (witness) *witness = ~0ull;
    for (int last = 1; last < nrounds && verdict; last++)
        for (int ind = 0; ind < last && verdict; ind++)
            for (int swap = 0; swap < 2 && verdict; swap++) {
                const int r0 = swap ? last : ind, r1 = swap ? ind : last;
                for (int k = 0; k < n; k++, idx++) {
                    u8 tr[KF_MAX_N];
                    int v1 = k, it = 0;   /* create2P1FTr */
                    for (int s = 0; s < n; s += 2) {
                        tr[v1] = (u8)it;
                        const int v2 = part[r0][v1];
                        it = part[0][it];
                        tr[v2] = (u8)it;
                        it = part[1][it];
                        v1 = part[r1][v2];
                    }
                    tried++;
                    int cmp = 0;
                    if (nrounds >= 3) {
                        u8 byv[KF_MAX_N][KF_MAX_N];
                        int have[KF_MAX_N] = {0};
                        for (int r = 0; r < nrounds; r++) {
                            if (r == r0 || r == r1) continue;
                            u8 row[KF_MAX_N], pres[KF_MAX_N] = {0}, sec[KF_MAX_N];
                            for (int q = 0; q < n; q += 2) {
                                u8 x = tr[mat[r * n + q]], y = tr[mat[r * n + q + 1]];
                                if (x > y) { u8 t = x; x = y; y = t; }
                                pres[x] = 1;
                                sec[x] = y;
                            }
                            int m = 0;
                            for (int v = 0; v < n; v++)
                                if (pres[v]) { row[m++] = (u8)v; row[m++] = sec[v]; }
                            memcpy(byv[row[1]], row, (size_t)n);
                            have[row[1]] = 1;
                        }
                        cmp = 0;
                        for (int v = 3; v <= nrounds && cmp == 0; v++) {
                            if (!have[v]) { cmp = 2; break; }
                            const int c = memcmp(byv[v], mat + (v - 1) * n, (size_t)n);
                            cmp = c < 0 ? -1 : (c > 0 ? 1 : 0);
                        }
                    }
                    if (cmp == 0) auts++;
                    else if (cmp < 0) {
                        verdict = 0;
                        if (witness) *witness = idx;
                        keep_witness(tr, n);
                        break;
                    }
                }
            }
    if (n_auts) *n_auts = auts;
    if (n_tried) *n_tried = tried;
    return verdict;
}

/*
 * Two-row canonizer of uniform 1-factorizations: alldata::cnvCheck2U1F
 * (k-Factorization/TripleSys/Source/cnvSupport2.cpp:3-142) with
 * getCyclesAndPathFromNeighbors and sortCycles (p1fCheck.cpp:65-76, 153-228),
 * CycleMapping::InitCycleMapping / ProceedToNextMapping (cycles.cpp:111-211),
 * createU1FTr (TrCycles.cpp:3-50) and the full comparison kmProcessMatrix
 * (kmSupport.cpp:275-313), for the U1F runs' defaults (complete graph,
 * AutSaveTestedTrs 0, Any2RowsConvertToFirst2 1, early exit).
 *
 * Cycles of a pair of rounds (a, b): from the smallest player not yet met,
 * alternately the a-partner and the b-partner until the start returns; the
 * path lists each step as a pair (x, a(x)), (a(x), b(a(x))), ... (2L entries
 * for a cycle of L players).  Cycles are bubble-sorted by length (stable).
 * Pairs of rounds in order (last = 1.., ind < last, (ind, last) then
 * (last, ind)); a pair whose sorted cycle lengths differ from those of rounds
 * 0, 1 makes the matrix not canonical at once (no relabelling is counted for
 * it).  Otherwise every mapping of its cycles onto the cycles of rounds 0, 1
 * is tried: cycles of one length permuted among themselves (lexicographic
 * next permutation, first group of equal lengths fastest), a start offset per
 * cycle (L/2 choices, every other player; mixed radix, first cycle fastest),
 * a direction per cycle (binary, last cycle fastest); directions change
 * fastest, then offsets, then permutations.  Relabelling index = ordered pair
 * index * T + position in that order, T = relabellings per pair.
 */
#define KF_MAXCYC 4

typedef struct { int n, len[KF_MAXCYC], start[KF_MAXCYC]; u8 path[2 * KF_MAX_N]; } kf_cycles;

static void kf_pair_cycles(const u8 *pa, const u8 *pb, int n, kf_cycles *c)
{
    int used[KF_MAX_N] = {0}, ip = 0;
    c->n = 0;
    for (int k0 = 0; k0 < n; k0++) {
        if (used[k0]) continue;
        if (c->n == KF_MAXCYC) { c->n = KF_MAXCYC + 1; return; }   /* more cycles than rounds 0, 1 can have */
        c->start[c->n] = ip / 2;
        int k = k0, len = 0;
        for (;;) {
            const int k1 = pa[k], k2 = pb[k1];
            c->path[ip++] = (u8)k; c->path[ip++] = (u8)k1;
            c->path[ip++] = (u8)k1; c->path[ip++] = (u8)k2;
            used[k] = used[k1] = 1;
            len += 2;
            if (k2 == k0) break;
            k = k2;
        }
        c->len[c->n++] = len;
    }
    for (int j = 1; j < c->n; j++)          /* sortCycles: bubble sort by length */
        for (int i = 0; i < c->n - j; i++)
            if (c->len[i] > c->len[i + 1]) {
                int t = c->len[i]; c->len[i] = c->len[i + 1]; c->len[i + 1] = t;
                t = c->start[i]; c->start[i] = c->start[i + 1]; c->start[i + 1] = t;
            }
}

/* next lexicographic permutation of p[0..m); 0 when p was the last */
static int kf_next_perm(int *p, int m)
{
    int i = m - 2;
    while (i >= 0 && p[i] > p[i + 1]) i--;
    if (i < 0) return 0;
    int j = m - 1;
    while (p[j] < p[i]) j--;
    int t = p[i]; p[i] = p[j]; p[j] = t;
    for (int a = i + 1, b = m - 1; a < b; a++, b--) { t = p[a]; p[a] = p[b]; p[b] = t; }
    return 1;
}

/* One mapping of a pair's cycles onto those of rounds 0, 1 (CycleMapping,
 * cycles.cpp:111-211): direction and start offset per cycle, and the order of
 * the cycles of one length. */
typedef struct { int dir[KF_MAXCYC], off[KF_MAXCYC], perm[KF_MAXCYC], sout[KF_MAXCYC]; } u1f_map;

static void u1f_map_first(u1f_map *m, const kf_cycles *cy)
{
    for (int c = 0; c < KF_MAXCYC; c++) { m->dir[c] = m->off[c] = 0; m->perm[c] = c; }
    for (int c = 0; c < cy->n; c++) m->sout[c] = cy->start[c];
}

/* ProceedToNextMapping: directions fastest (last cycle first), then offsets
 * (first cycle first), then the permutations of each group of equal lengths */
static void u1f_map_next(u1f_map *m, const kf_cycles *cy)
{
    int i = cy->n;
    while (i-- && m->dir[i]) m->dir[i] = 0;
    if (i >= 0) { m->dir[i] = 1; return; }
    int c = 0;
    while (c < cy->n && m->off[c] + 1 == cy->len[c] / 2) m->off[c++] = 0;
    if (c < cy->n) { m->off[c]++; return; }
    for (int g0 = 0, g1; g0 < cy->n; g0 = g1) {
        for (g1 = g0; g1 < cy->n && cy->len[g1] == cy->len[g0]; g1++) {}
        if (kf_next_perm(m->perm + g0, g1 - g0)) {
            for (int j = g0; j < g1; j++) m->sout[j] = cy->start[g0 + m->perm[j] - g0];
            return;
        }
        for (int j = g0; j < g1; j++) m->perm[j] = j;
    }
}

/* createU1FTr: the relabelling sending the pair's cycles, as mapped, onto
 * those of rounds 0, 1 */
static void u1f_map_tr(const u1f_map *m, const kf_cycles *cy, const kf_cycles *ref, u8 *tr)
{
    for (int c = 0; c < cy->n; c++) {
        const int L2 = 2 * cy->len[c];
        int ip = m->dir[c] ? 4 * m->off[c] + 1 : 4 * m->off[c];
        const int step = m->dir[c] ? -1 : 1;
        const u8 *fp = cy->path + 2 * m->sout[c];
        const u8 *fp01 = ref->path + 2 * ref->start[c];
        for (int j = 0; j < L2; j++, ip += step) {
            if (ip >= L2) ip -= L2;
            else if (ip < 0) ip += L2;
            tr[fp[ip]] = fp01[j];
        }
    }
}

/* relabellings per pair: the cycle lengths' product (offsets x directions)
 * times the orders of the cycles of equal lengths */
static uint64_t u1f_per_pair(const kf_cycles *ref)
{
    uint64_t T = 1;
    for (int c = 0; c < ref->n; c++) T *= (uint64_t)ref->len[c];
    for (int c = 0, g = 0; c < ref->n; c = g) {
        for (g = c; g < ref->n && ref->len[g] == ref->len[c]; g++) {}
        for (int f = 2; f <= g - c; f++) T *= (uint64_t)f;
    }
    return T;
}

int orc_kf_check_2u1f(const uint8_t *mat, int n, int nrounds, uint64_t *witness,
                      uint64_t *n_auts, uint64_t *n_tried)
{
    if (n < 4 || n > KF_MAX_N || (n & 1) || nrounds < 1 || nrounds > n - 1) return -1;
    u8 part[KF_MAX_N][KF_MAX_N];
    if (!kf_partners(mat, n, nrounds, part)) return -1;
    uint64_t auts = 0, tried = 0, idx = 0;
    int verdict = 1;
    witness_valid = 0;
    if (witness) *witness = ~0ull;
    if (nrounds >= 2) {
        kf_cycles ref;
        kf_pair_cycles(part[0], part[1], n, &ref);
        if (ref.n > KF_MAXCYC) return -1;
        const uint64_t T = u1f_per_pair(&ref);
        for (int last = 1; last < nrounds && verdict; last++)
            for (int ind = 0; ind < last && verdict; ind++)
                for (int swap = 0; swap < 2 && verdict; swap++) {
                    const int r0 = swap ? last : ind, r1 = swap ? ind : last;
                    kf_cycles cy;
                    kf_pair_cycles(part[r0], part[r1], n, &cy);
                    int same = cy.n == ref.n;
                    for (int c = 0; c < ref.n && same; c++) same = cy.len[c] == ref.len[c];
                    if (!same) { verdict = 0; break; }
                    u1f_map m;
                    u1f_map_first(&m, &cy);
                    for (uint64_t t = 0; t < T; t++, idx++) {
                        if (t) u1f_map_next(&m, &cy);
                        u8 tr[KF_MAX_N];
                        u1f_map_tr(&m, &cy, &ref, tr);
                        u8 mo[KF_MAX_N * KF_MAX_N];
                        normal_form(mat, nrounds, n, tr, mo);
                        tried++;
                        const int cmp = memcmp(mo, mat, (size_t)(n * nrounds));
                        if (cmp == 0) auts++;
                        else if (cmp < 0) {
                            verdict = 0;
                            if (witness) *witness = idx;
                            keep_witness(tr, n);
                            break;
                        }
                    }
                }
    }
    if (n_auts) *n_auts = auts;
    if (n_tried) *n_tried = tried;
    return verdict;
}

/*
 * cnvCheck3U1F (k-Factorization/TripleSys/Source/cnvSupport.cpp:94-339) for
 * groups of 2 on complete graphs, with the defaults of such runs: no
 * second-row database (p1 = result(1), the current set, one pass of the outer
 * loop), AutSaveTestedTrs 0, undefined cycles not allowed, early exit.
 *  - cyclesFor2Rows (p1fCheck.cpp:769-795) collects the cycle set of rounds
 *    0, 1 checked with eCheckErrors: every cycle length must be one the run
 *    allows (cycleLengthOk, p1fCheck.cpp:540-558) and the set one of its types
 *    (cyclesNotOk, :519-538); `types` ([ntypes][4] lengths, ascending, zero
 *    after the last) is the u1fCycles parameter, none meaning one cycle
 *    through every player.  When no set is collected, the two-row test
 *    cyclesOfTwoRowsOk (cnvSupport4.cpp:2-55) fails at two rounds and no pair
 *    selects a set at more: not canonical before any relabelling.
 *  - Pairs of rounds in the order of cnvCheck2U1F; for each, the loop over
 *    nv1 = MAX_CYCLE_SETS (32) cycle sets (cnvSupport.cpp:208-216), where
 *    collectOneCyclesSet (:341-352) recomputes the pair's cycles for groups
 *    of 2 whatever the index; a set whose lengths are not those of the
 *    collected set is not selected, which makes the matrix not canonical
 *    (:288-303).  Every mapping onto the selected set is tried
 *    (createU1FTr, kmProcessMatrix): smaller -> not canonical; equal ->
 *    updateGroup, which keeps each automorphism once.
 * Relabelling index ((ordered pair * 32) + repetition) * T + position.
 */
#define KF_3U1F_REPS 32

int orc_kf_check_3u1f(const uint8_t *mat, int n, int nrounds, const uint8_t *types, int ntypes,
                      uint64_t *witness, uint64_t *n_auts, uint64_t *n_tried)
{
    if (n < 4 || n > KF_MAX_N || (n & 1) || nrounds < 1 || nrounds > n - 1 || ntypes < 0) return -1;
    u8 part[KF_MAX_N][KF_MAX_N];
    if (!kf_partners(mat, n, nrounds, part)) return -1;
    uint64_t tried = 0;
    size_t ngrp = 0, capgrp = 0;
    u8 *grp = NULL;   /* the automorphisms found (updateGroup) */
    int verdict = 1;
    if (witness) *witness = ~0ull;
    if (nrounds >= 2) {
        kf_cycles ref;
        kf_pair_cycles(part[0], part[1], n, &ref);
        if (ref.n > KF_MAXCYC) return -1;
        int allowed = 0;
        if (ntypes == 0) allowed = ref.n == 1 && ref.len[0] == n;
        for (int k = 0; k < ntypes && !allowed; k++) {
            int same = 1, c = 0;
            for (; c < KF_MAXCYC; c++) {
                const int want = types[k * KF_MAXCYC + c];
                if (c < ref.n ? want != ref.len[c] : want != 0) same = 0;
            }
            allowed = same;
        }
        if (!allowed) verdict = 0;
        const uint64_t T = u1f_per_pair(&ref);
        uint64_t po = 0;
        for (int last = 1; last < nrounds && verdict; last++)
            for (int ind = 0; ind < last && verdict; ind++)
                for (int swap = 0; swap < 2 && verdict; swap++, po++) {
                    const int r0 = swap ? last : ind, r1 = swap ? ind : last;
                    for (int rep = 0; rep < KF_3U1F_REPS && verdict; rep++) {
                        kf_cycles cy;
                        kf_pair_cycles(part[r0], part[r1], n, &cy);
                        int same = cy.n == ref.n;
                        for (int c = 0; c < ref.n && same; c++) same = cy.len[c] == ref.len[c];
                        if (!same) { verdict = 0; break; }
                        u1f_map m;
                        u1f_map_first(&m, &cy);
                        for (uint64_t t = 0; t < T; t++) {
                            if (t) u1f_map_next(&m, &cy);
                            u8 tr[KF_MAX_N];
                            u1f_map_tr(&m, &cy, &ref, tr);
                            u8 mo[KF_MAX_N * KF_MAX_N];
                            normal_form(mat, nrounds, n, tr, mo);
                            tried++;
                            const int cmp = memcmp(mo, mat, (size_t)(n * nrounds));
                            if (cmp < 0) {
                                verdict = 0;
                                if (witness) *witness = (po * KF_3U1F_REPS + (uint64_t)rep) * T + t;
                                break;
                            }
                            if (cmp == 0) {
                                size_t g = 0;
                                while (g < ngrp && memcmp(grp + g * KF_MAX_N, tr, (size_t)n)) g++;
                                if (g == ngrp) {
                                    if (ngrp == capgrp) {
                                        capgrp = capgrp ? 2 * capgrp : 64;
                                        u8 *ng = (u8 *)realloc(grp, capgrp * KF_MAX_N);
                                        if (!ng) { free(grp); return -1; }
                                        grp = ng;
                                    }
                                    memcpy(grp + ngrp++ * KF_MAX_N, tr, (size_t)n);
                                }
                            }
                        }
                    }
                }
    }
    free(grp);
    if (n_auts) *n_auts = (uint64_t)ngrp;
    if (n_tried) *n_tried = tried;
    return verdict;
}

/* the QuickPerm visiting order (for tests) */
long orc_kf_quickperm(int k, uint8_t *out) { return quickperm_table(k, out); }

/*
 * The k-Sys backjump position of a test that found a smaller relabelling:
 * kmProcessMatrix's comparison (kmSupport.cpp:275-313) of the witness's normal
 * form with the matrix, and setPlayerIndex / setPlayerIndexByPos
 * (kmSupport.cpp:314-356) at the first smaller row; the position starts at
 * nrows * n - 3 (cnvCheckNew, cnv.cpp:213: nrows * players - groupSize - 1)
 * and only ever decreases.  mode 0: general canonizer, 1: two-row P1F,
 * 2: two-row U1F (each run first, the witness relabelling taken from it).
 * info (optional, 3 ints): the first smaller row of the normal form, the round
 * it came from, and whether the position was set there (1) or kept (0).
 */
int orc_kf_player_index(const uint8_t *mat, int n, int nrounds, int mode, int *info)
{
    uint64_t w, a, t;
    const int v = mode == 1 ? orc_kf_check_2p1f(mat, n, nrounds, &w, &a, &t)
                : mode == 2 ? orc_kf_check_2u1f(mat, n, nrounds, &w, &a, &t)
                            : orc_kf_check(mat, n, nrounds, &w, &a, &t);
    const int start = nrounds * n - 2 - 1;
    if (info) info[0] = info[1] = -1, info[2] = 0;
    if (v < 0) return -1;
    if (v == 1 || !witness_valid) return start;
    /* the witness's rounds, relabelled and normalised, with their sources */
    u8 rel[KF_MAX_N][KF_MAX_N];
    int src_of_key[KF_MAX_N];
    for (int q = 0; q < KF_MAX_N; q++) src_of_key[q] = -1;
    for (int r = 0; r < nrounds; r++) {
        u8 one[KF_MAX_N];
        normal_form(mat + r * n, 1, n, witness_tr, one);
        memcpy(rel[r], one, (size_t)n);
        src_of_key[one[1]] = r;          /* the partner of label 0 orders the rounds */
    }
    int day_max = -1, row = 0;
    for (int key = 0; key < n; key++) {
        const int r = src_of_key[key];
        if (r < 0) continue;
        const int c = memcmp(rel[r], mat + row * n, (size_t)n);
        if (day_max < 0) day_max = r;
        if (c > 0) return start;         /* not smaller: cannot happen for a witness */
        if (c == 0) {
            if (day_max < r) day_max = r;
            row++;
            continue;
        }
        if (info) info[0] = row, info[1] = r;
        if (day_max > r) return start;   /* an earlier row came from a later round */
        /* first position where the relabelled row is smaller (at most n - 3) */
        const u8 *co = rel[r], *ci = mat + row * n;
        int ip = 0;
        while (ip < n - 3 && !(co[ip] < ci[ip])) ip++;
        u8 player_of[KF_MAX_N], pos_in_src[KF_MAX_N];
        for (int q = 0; q < n; q++) {
            player_of[witness_tr[q]] = (u8)q;
            pos_in_src[mat[r * n + q]] = (u8)q;
        }
        int j = 0;
        for (int q = 0; q <= ip; q++)
            if (j < pos_in_src[player_of[co[q]]]) j = pos_in_src[player_of[co[q]]];
        const int idx = r * n + j;
        if (info) info[2] = 1;
        return idx < start ? idx : start;
    }
    return start;
}
