/*
 * ise_oracle.h -- CPU oracle for the batched canonicity test.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark may use it.
 *
 * It restates, in plain C over uint8 0/1 matrices, the reference's CPU
 * algorithms:
 *   - CCanonicityChecker::TestCanonicity   Source/CanonicityChecker.h:255-586
 *     (non-combined designs: numParts == 1, rank == 2, no group on parts)
 *   - CCanonicityChecker::checkColOrbit    Source/CanonicityChecker.cpp:191-223
 *   - CCanonicityChecker::rowToChange      Source/CanonicityChecker.cpp:320-330
 *   - CCanonicityChecker::constructColIndex Source/CanonicityChecker.cpp:332-342
 *   - CGroupOrder::next_permutation         k-Factorization/Utils/Include/GroupOrder.h:136-248
 *   - CGroupOrder::addAutomorphism / updateGroupOrder / udpdateStabLength
 *                                           GroupOrder.h:56-102, Update_Orbits :30-49
 *   - a depth-first row-by-row BIBD enumerator in the shape of
 *     CEnumerator::Enumerate (Source/Enumerator.cpp:540-1024) with its own
 *     row solver (the reference's InSysSolver is replaced by a direct DFS over
 *     column-orbit counts)
 *   - the k-factorization canonizer alldata::cnvCheckNew / canonizator /
 *     cnvCheckKm1 / kmProcessMatrix (k-Factorization/TripleSys/Source/cnv.cpp,
 *     kmSupport.cpp)
 *
 * Parity pins (see tests/): published design counts and |Aut| histograms,
 * canonical matrices printed by the reference, brute-force canonical forms and
 * automorphism-group orders of small matrices.
 */
#ifndef ISE_ORACLE_H
#define ISE_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_ROWS 32
#define ORC_MAX_COLS 64

/* outInfo bits, same values as t_canonOutInfo (Source/CanonicityChecker.h:27-32) */
#define ORC_SAVE_ROW_PERMUTATIONS 1u
#define ORC_SAVE_COL_PERMUTATIONS 2u
#define ORC_SAVE_ROW_TO_CHANGE    4u

/* Per-candidate result arrays (structure of arrays, n entries each). */
typedef struct {
    uint8_t  *verdict;        /* [n]      1 = canonical */
    uint8_t  *fail_row;       /* [n]      row where the permuted matrix won (255 if canonical) */
    uint8_t  *fail_col;       /* [n]      first column of the deciding column orbit (255) */
    uint8_t  *row_to_change;  /* [n]      rowToChange(fail_row) (255 if canonical) */
    uint64_t *group_order;    /* [n]      |Aut| (row action) when ORC_SAVE_ROW_PERMUTATIONS */
    uint8_t  *perm_row;       /* [n][32]  row permutation at exit (witness when non-canonical) */
    uint8_t  *perm_col;       /* [n][64]  column permutation at exit */
    uint8_t  *orbits;         /* [n][32]  row orbits of the automorphisms found */
    uint16_t *n_gens;         /* [n]      automorphisms found (generators recorded) */
    uint8_t  *gens;           /* [n][gen_cap][64] or NULL */
    uint32_t *counters;       /* [n][4]   perms tried, rows compared, orbits compared, cells examined */
    uint32_t gen_cap;         /* unused here (an argument of orc_check_batch); keeps the layout of the
                                 library's ise_results so that one structure serves both */
    uint64_t *group_order_hi; /* [n]      bits 64..127 of |Aut| (the reference's UInt is 128-bit) */
} orc_results;

/* Canonicity test of a batch of bit-packed candidates.
 * rows/cells: [n][stride] uint64 words, bit c = column c.  cells[i] is the
 * column-orbit start mask of row i (orbits induced by rows 0..i-1); when cells
 * is NULL it is derived from the rows.  meta[c] = depth | ncols<<8 |
 * lenStab<<16 | outInfo<<24.  Returns 0 on success, <0 on invalid input. */
int orc_check_batch(uint32_t n, uint32_t stride, const uint64_t *rows,
                    const uint64_t *cells, const uint32_t *meta,
                    uint32_t gen_cap, orc_results *res);

/* Single candidate on a row-major uint8 matrix (nrows x ncols, entries 0/1). */
int orc_test_canonicity(const uint8_t *mat, int nrows, int ncols, int len_stab,
                        unsigned out_info, orc_results *res, uint32_t gen_cap);

/* Column-orbit start masks of a bit-packed matrix: out[i] = orbits of row i. */
void orc_cell_masks(const uint64_t *rows, int nrows, int ncols, uint64_t *out);

/* CMatrixCanonChecker::CanonizeMatrix for a 0/1 matrix (k = 0,
 * Source/MatrixCanonChecker.cpp:420-564): columns first put in orbit form
 * (decreasing column order), then, while the test with row permutations saved
 * finds a witness, the rows are rearranged by it and the columns put in orbit
 * form again.  out receives the final (canonical) matrix, *aut its |Aut|,
 * *iters the tests run.  Returns 1, or 0 when max_iters tests did not
 * finish, -1 on bad sizes. */
int orc_canonize(const uint8_t *mat, int nrows, int ncols, int len_stab, int max_iters, uint8_t *out,
                 uint64_t *aut, uint32_t *iters);

/* The improved solution CCanonicityChecker::reconstructSolution builds when
 * the test of this matrix (out_info 0) fails on its last row
 * (CanonicityChecker.cpp:227-277, called at CanonicityChecker.h:382): one
 * count per column orbit of the last row.  Returns the number of orbits, or -1
 * when the matrix is canonical or fails on an earlier row. */
int orc_improved_solution(const uint8_t *mat, int nrows, int ncols, int len_stab, uint8_t *out);

/* Test hook: one step of the row-permutation walk the test runs
 * (CGroupOrder::next_permutation, GroupOrder.h:136-248) on caller-held state.
 * perm[n], orbits[n] (orbit leader of each row); at = 255 for the plain
 * successor, 254 for the first call after an automorphism, else the position
 * whose prefix was refuted; fixed = rows held in place (lenStab).
 * state[0] = stabiliser point (m_nStabLength), state[1] = m_nStabLengthAut.
 * Returns the first position whose image changed, or 255 when the walk is
 * exhausted. */
int orc_walk_step(uint8_t *perm, const uint8_t *orbits, int n, int at, int fixed, uint16_t *state);

/* Depth-first enumeration of 2-(v,k,lambda) designs (b, r derived).
 * Canonical full designs are written to designs[] as v packed row words each
 * (up to max_designs) with |Aut| in auts[].  stats[0..3] = candidates tested,
 * canonical partials, fully constructed matrices, designs.  Returns the number
 * of designs found, or <0 on error.  If cand_rows is non-NULL every tested
 * candidate with depth in [dmin,dmax] is appended (stride words each, up to
 * cand_cap; the count goes to stats[4]). */
long orc_enumerate_bibd(int v, int k, int lambda, uint64_t *designs,
                        uint64_t *auts, long max_designs, uint64_t *stats,
                        uint64_t *cand_rows, uint32_t *cand_meta, long cand_cap,
                        int stride, int dmin, int dmax);

/* Beam-sampled level-by-level enumeration: every child of the kept canonical
 * candidates of the previous level is tested (and written out when its depth
 * is in [dmin,dmax], up to cand_cap); at most `beam` canonical candidates per
 * level are kept (seeded partial Fisher-Yates with splitmix64, indices then
 * sorted; beam 0 keeps all).  level_counts[d] = candidates at depth d.
 * Returns the number of candidates written, <0 on error. */
long orc_enumerate_sampled(int v, int k, int lambda, uint32_t beam, uint64_t seed, int dmin, int dmax,
                           int stride, uint64_t *cand_rows, uint32_t *cand_meta, long cand_cap,
                           uint64_t *level_counts);

/* Solutions of the next-row system of one partial matrix (rows[0..depth-1]):
 * the rows the enumerator would try next, in its order (up to cap written).
 * Returns the number of solutions or <0 on error. */
long orc_row_solutions(int v, int k, int lambda, const uint64_t *rows, int depth, uint64_t *out, long cap);

/* Brute-force reference for tiny matrices: the lexicographically largest
 * matrix over all row permutations (columns re-sorted in decreasing order)
 * and the number of row permutations that map the matrix to itself. */
int orc_bruteforce_canon(const uint8_t *mat, int nrows, int ncols,
                         uint8_t *canon_out, uint64_t *n_row_auts);

/* Witness hunt: depth-first search over row prefixes where position i tries
 * rows in the order order[i*d + 0..d-1]; same comparison rule and column
 * state as the plain search; at most `budget` children compared.  Returns 1
 * (witness in perm_out, first larger row in *fail_row) or 0. */
int orc_hunt(const uint8_t *mat, int d, int b, const uint8_t *order, uint32_t budget,
             uint8_t *perm_out, uint8_t *fail_row, uint32_t *nodes);

/* Batch test with a witness hunt first (order NULL = increasing row order;
 * else [n][order_dim][order_dim] bytes), then the plain search for the
 * candidates the hunt leaves open.  Same result layout as orc_check_batch;
 * counters add the hunt's comparisons to the plain search's. */
int orc_check_batch_hunt(uint32_t n, uint32_t stride, const uint64_t *rows, const uint64_t *cells,
                         const uint32_t *meta, const uint8_t *order, uint32_t order_dim,
                         uint32_t budget, uint32_t gen_cap, orc_results *res);

/* ---- lead order: the learned order of the rows of position 0 ----
 * Features of every row j < d <= 16 as the row of position 0: X[c][j][16]
 * (zero past d).  Defined in ise_oracle_bibd.c with the search below. */
int orc_lead_features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta, float *X);
/* Learned-order test of a batch: the plain search decides every candidate
 * whose walk needs at most gate_perms permutations (and every candidate the
 * lead order does not take: depth < 2 or > 16, lenStab > 0, more than 56
 * columns or not in orbit form); the others run the depth-first search with
 * the rows of position 0 in the order lead[c][0..15] (entries >= depth
 * skipped): a witness of that order, the plain results when it ends without
 * one (canonical, trivial group), the plain search when it meets an
 * automorphism (as it does when the gate's permutations found one).  The
 * rows of position 0: lead[c][0..15] (rows >= depth and repeats skipped), then
 * the rows not listed, in increasing order.  res->counters is required (the
 * gate reads it); for the candidates of the lead search it receives the
 * gate's work (the walk's first gate_perms permutations), then the lead
 * search's (permutations counted as the walk counts them: a new row after a
 * back-up or a refutation), plus the whole plain search's when it met an
 * automorphism.  mode[c] (optional): 0 plain only, 1 decided within the gate,
 * 2 lead witness, 3 lead search exhausted, 4 automorphism met. */
int orc_check_batch_lead(uint32_t n, uint32_t stride, const uint64_t *rows, const uint64_t *cells,
                         const uint32_t *meta, const uint8_t *lead, uint32_t gate_perms, uint32_t gen_cap,
                         orc_results *res, uint8_t *mode);

/* ---- learned row-order scorer (ise_oracle_mlp.c) ---- */
/* 16 features of every (position i, row j) pair: X[(c*od+i)*od+j][16],
 * zero outside depth x depth. */
int orc_mlp_features(uint32_t n, uint32_t stride, const uint64_t *rows, const uint32_t *meta,
                     uint32_t od, float *X);
/* Logits of M feature rows (params packed as the library's ISE_MLP_PARAMS);
 * bf16_rounding != 0 rounds at the GPU path's points. */
int orc_mlp_forward(const float *params, const float *X, uint32_t M, int bf16_rounding, double *z);
/* Weighted mean binary cross-entropy of the logits against y and its
 * gradient (packed like the parameters). */
int orc_mlp_gradient(const float *params, const float *X, const float *y, const float *w, uint32_t M,
                     int bf16_rounding, double *grad, double *loss);

/* ---- k-factorizations of K_n (n <= 16): rounds of n/2 pairs ---- */
/* Canonicity of a partial 1-factorization (nrounds x n bytes, each round the
 * n players as n/2 ordered pairs).  Returns 1 canonical, 0 not, <0 error
 * (bad sizes, a round that is not a perfect matching of 0..n-1, or a pair in
 * two rounds; the same holds for every orc_kf_* canonizer below).
 * On non-canonical, *witness receives the index of the first relabelling in
 * plain order that produced a smaller normal form; n_auts counts
 * relabellings that reproduced the matrix. */
int orc_kf_check(const uint8_t *mat, int n, int nrounds, uint64_t *witness,
                 uint64_t *n_auts, uint64_t *n_tried);

/* Normal form of a factorization under the identity labelling (pairs
 * sorted, pairs by first element, rows by partner of player 0). */
void orc_kf_normalize(uint8_t *mat, int n, int nrounds);

/* Brute force over all n! player relabellings (n <= 8): canonical or not,
 * and the number of relabellings that reproduce the matrix. */
int orc_kf_bruteforce(const uint8_t *mat, int n, int nrounds, uint64_t *n_auts);
/* cnvCheck2P1F: the two-row canonizer of perfect 1-factorizations (same
 * results fields as orc_kf_check; witness = index in its own order). */
int orc_kf_check_2p1f(const uint8_t *mat, int n, int nrounds, uint64_t *witness,
                      uint64_t *n_auts, uint64_t *n_tried);
/* cnvCheck2U1F: the two-row canonizer of uniform 1-factorizations (a pair of
 * rounds of another cycle type than rounds 0, 1: not canonical, witness ~0). */
int orc_kf_check_2u1f(const uint8_t *mat, int n, int nrounds, uint64_t *witness,
                      uint64_t *n_auts, uint64_t *n_tried);
/* cnvCheck3U1F for groups of 2 on complete graphs: the run's cycle types
 * types[ntypes][4] (ascending lengths, zero after the last; none: one cycle
 * through every player); witness = ((ordered pair * 32) + repetition) * T +
 * position; n_auts = distinct automorphisms found. */
int orc_kf_check_3u1f(const uint8_t *mat, int n, int nrounds, const uint8_t *types, int ntypes,
                      uint64_t *witness, uint64_t *n_auts, uint64_t *n_tried);

/* Fuchs' countdown QuickPerm order of the pair slots (k! rows of k bytes,
 * each byte the base label 2*slot); returns k!. */
long orc_kf_quickperm(int k, uint8_t *out);

/* k-Sys backjump position (setPlayerIndex, kmSupport.cpp:314-356) of the first
 * smaller relabelling found by canonizer `mode` (0 general, 1 two-row P1F,
 * 2 two-row U1F); nrows * n - 3 when there is none.  info (optional, 3 ints):
 * first smaller row of the normal form, the round it came from, 1 if the
 * position was set there. */
int orc_kf_player_index(const uint8_t *mat, int n, int nrounds, int mode, int *info);

/* runCanonizer (Source/CDTools.cpp:25-51): mat = `day` days (day <= 0: all
 * (v - 1) / (k - 1) of them) of v players, groups of k in order; out receives
 * the v x b player rows of the canonical binary matrix, *ncols = b.  Every day
 * order is tried when all days are present.  0 ok, -1 invalid arguments, -2 a
 * canonization did not end. */
int orc_cd_run_canonizer(const uint8_t *mat, int v, int k, int day, uint8_t *out, int *ncols);
/* The same with the group orders of the result: *day_order = day orders whose
 * canonical matrix is the result (1 without a walk), *fixed_order = its |Aut|
 * with every day fixed. */
int orc_cd_run_canonizer_ex(const uint8_t *mat, int v, int k, int day, uint8_t *out, int *ncols,
                            uint64_t *day_order, uint64_t *fixed_order);
/* The largest canonical form (v x b bytes into best) over the given day
 * orders (orders [n_orders][days]: position j takes day orders[t][j]): the
 * inner step of the full-day walk, for timing samples of it. */
int orc_cd_canonize_orders(const uint8_t *mat, int v, int k, int days, const uint8_t *orders, long n_orders,
                           uint8_t *best);
/* Column-partition mask of `days` days of `groups` columns (bit j*groups). */
uint64_t orc_cd_day_starts(int days, int groups);
/* CanonizeMatrix with an initial column partition: part bit c set = column c
 * starts a block (the orbits of a first row of labels; part = 1: one block,
 * orc_canonize with lenStab 0).  Columns never leave their block.  Returns 1
 * (canonical form in out), 0 (max_iters tests did not reach it), -1. */
int orc_canonize_part(const uint8_t *mat, int nrows, int ncols, uint64_t part, int max_iters, uint8_t *out,
                      uint64_t *aut, uint32_t *iters);

#ifdef __cplusplus
}
#endif
#endif
