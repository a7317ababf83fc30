// ise_random.h -- the seeded random valid extension walk, shared by the host
// threads and the GPU generator (one thread per candidate) so that both write
// the same candidates for the same seed.
//
// A candidate of depth t is a partial incidence matrix whose rows each take
// r columns, one at a time, uniformly among the columns whose weight is below
// k and that keep every pairwise row intersection <= lambda; a row that
// cannot be drawn in 4 tries makes the previous row be drawn again
// (randomized backtracking).  A walk stops after 256 draws.  With attempts >
// 1, a walk that falls short of the depth asked starts again from the empty
// matrix, up to `attempts` walks: rejection sampling of the depth, for the
// near-complete depths that few walks reach.  The deepest matrix met is kept,
// its columns sorted in decreasing lexicographic order (row 0 most
// significant), so that column orbits are contiguous with ones first.
#pragma once
#include <stdint.h>
#include "ise.h"

#define ISE_HD __host__ __device__ inline

namespace ise_rw {

struct Params {
    int v, b, r, k, lambda;
};

ISE_HD uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

ISE_HD uint32_t below(uint64_t &s, uint32_t n) { return (uint32_t)(splitmix(s) % n); }

ISE_HD int popc(uint64_t x) {
    int c = 0;
    for (; x; x &= x - 1) c++;
    return c;
}

ISE_HD int ctz(uint64_t x) {
    int c = 0;
    while (!((x >> c) & 1ull)) c++;
    return c;
}

// One row for position t; false when it gets stuck.  `blocked` collects the
// rows whose intersection with the new row reached lambda.
ISE_HD bool draw_row(const Params &p, const uint64_t *rows, const int *colw, int t, uint64_t &rng,
                     uint64_t *out) {
    int inter[ISE_MAX_ROWS];
    for (int u = 0; u < t; u++) inter[u] = 0;
    uint64_t open = 0, row = 0, blocked = 0;
    for (int c = 0; c < p.b; c++)
        if (colw[c] < p.k) open |= 1ull << c;
    for (int picked = 0; picked < p.r; picked++) {
        uint64_t elig = open & ~row & ~blocked;
        if (!elig) return false;
        // the j-th eligible column in increasing order
        for (uint32_t j = below(rng, (uint32_t)popc(elig)); j; j--) elig &= elig - 1;
        const int c = ctz(elig);
        row |= 1ull << c;
        for (int u = 0; u < t; u++)
            if (((rows[u] >> c) & 1) && ++inter[u] >= p.lambda) blocked |= rows[u];
    }
    *out = row;
    return true;
}

// Candidate of up to `depth` rows into rows_out (depth words); returns the
// depth reached.
ISE_HD int random_candidate(const Params &p, int depth, uint32_t attempts, uint64_t &rng, uint64_t *rows_out) {
    uint64_t best[ISE_MAX_ROWS], rows[ISE_MAX_ROWS];
    int best_depth = 0;
    for (uint32_t attempt = 0; attempt < attempts && best_depth < depth; attempt++) {
        int colw[ISE_MAX_COLS];
        int fails[ISE_MAX_ROWS + 1];
        for (int c = 0; c < p.b; c++) colw[c] = 0;
        fails[0] = 0;
        int t = 0;
        for (int budget = 256; budget > 0 && best_depth < depth; budget--) {
            uint64_t row;
            if (draw_row(p, rows, colw, t, rng, &row)) {
                rows[t] = row;
                for (int c = 0; c < p.b; c++)
                    if ((row >> c) & 1) colw[c]++;
                fails[++t] = 0;
                if (t > best_depth) {
                    best_depth = t;
                    for (int i = 0; i < t; i++) best[i] = rows[i];
                }
            } else if (++fails[t] >= 4 && t > 0) {
                --t;   // redraw the previous row
                for (int c = 0; c < p.b; c++)
                    if ((rows[t] >> c) & 1) colw[c]--;
            }
        }
    }
    const int d = best_depth;
    // columns in decreasing lexicographic order: keys sorted by insertion
    uint32_t key[ISE_MAX_COLS];
    for (int c = 0; c < p.b; c++) {
        uint32_t x = 0;
        for (int i = 0; i < d; i++) x = (x << 1) | (uint32_t)((best[i] >> c) & 1);
        int j = c;
        while (j > 0 && key[j - 1] < x) {
            key[j] = key[j - 1];
            j--;
        }
        key[j] = x;
    }
    for (int i = 0; i < d; i++) {
        uint64_t w = 0;
        for (int c = 0; c < p.b; c++)
            if ((key[c] >> (d - 1 - i)) & 1) w |= 1ull << c;
        rows_out[i] = w;
    }
    return d;
}

// Candidate i of a batch: its own generator state and uniform depth in
// [dmin, dmax]; writes stride words (zero past the depth) and its meta word.
ISE_HD void batch_candidate(const Params &p, uint64_t i, int dmin, int dmax, uint64_t seed, uint32_t attempts,
                            uint32_t stride, uint64_t *rows_out, uint32_t *meta_out) {
    uint64_t rng = seed ^ (0xd1342543de82ef95ull * (i + 1));
    splitmix(rng);
    const int target = dmin + (int)below(rng, (uint32_t)(dmax - dmin + 1));
    uint64_t *dst = rows_out + i * stride;
    for (uint32_t w = 0; w < stride; w++) dst[w] = 0;
    const int depth = random_candidate(p, target, attempts, rng, dst);
    const uint32_t oi = depth == p.v ? (ISE_SAVE_ROW_PERMUTATIONS | ISE_SAVE_ROW_TO_CHANGE) : 0u;
    meta_out[i] = (uint32_t)depth | ((uint32_t)p.b << 8) | (oi << 24);
}

} // namespace ise_rw

struct ise_ctx;
namespace ise {
// The batch generated on the context's device, one thread per candidate
// (host output buffers).
int random_extensions_device(ise_ctx *c, const ise_rw::Params &p, uint64_t n, int dmin, int dmax, uint64_t seed,
                             uint32_t attempts, uint32_t stride, uint64_t *rows_out, uint32_t *meta_out);
} // namespace ise
