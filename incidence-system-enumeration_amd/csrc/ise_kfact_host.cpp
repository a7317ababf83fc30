// ise_kfact_host.cpp -- host side of the 1-factorization path: the
// pair-slot arrangement tables and the seeded random candidate producer.
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <thread>
#include <vector>
#include "ise_kfact.h"

namespace ise {

static uint32_t factorial(uint32_t k) {
    uint32_t f = 1;
    for (uint32_t i = 2; i <= k; i++) f *= i;
    return f;
}

size_t kf_table_words() {
    size_t w = 0;
    for (uint32_t k = 0; k <= KF_MAX_SLOTS; k++) w += factorial(k);
    return 2 * w;   // arrangements, then their inverses
}

// Fuchs' countdown QuickPerm (the order canonizator visits the arrangements
// of the pair slots in, k-Factorization/TripleSys/Source/cnv.cpp:283-315).
void kf_build_tables(uint64_t *words, uint32_t *offset, uint32_t *count) {
    const size_t half = kf_table_words() / 2;
    uint32_t at = 0;
    for (uint32_t k = 0; k <= KF_MAX_SLOTS; k++) {
        offset[k] = at;
        count[k] = factorial(k);
        uint8_t a[KF_MAX_SLOTS + 1] = {0}, p[KF_MAX_SLOTS + 1] = {0};
        for (uint32_t i = 0; i < k; i++) { a[i] = (uint8_t)(2 * i); p[i] = (uint8_t)i; }
        p[k] = (uint8_t)k;
        auto emit = [&]() {
            uint64_t w = 0, v = 0;
            for (uint32_t i = 0; i < k; i++) w |= (uint64_t)a[i] << (8 * i);
            for (uint32_t j = 0; j < KF_MAX_SLOTS; j++) v |= (uint64_t)KF_PERM_ZERO << (8 * j);
            for (uint32_t i = 0; i < k; i++) {
                const uint32_t j = a[i] / 2;
                v = (v & ~(0xffull << (8 * j))) | ((uint64_t)i << (8 * j));
            }
            words[at] = w;
            words[half + at] = v;
            at++;
        };
        emit();
        uint32_t i = 1;
        while (i < k) {
            p[i]--;
            const uint32_t j = (i % 2) ? p[i] : 0;
            std::swap(a[i], a[j]);
            emit();
            i = 0;
            while (!p[++i]) p[i] = (uint8_t)i;
        }
    }
}

namespace {

inline uint64_t mix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// One random partial factorization of up to nr rounds of n players as pair
// lists; returns the rounds reached.  A stuck round is redrawn (64 tries), a
// matrix whose round cannot be placed is restarted (64 restarts); when every
// restart fails the deepest matrix met is kept, so the call always ends.
int random_rounds(int n, int nr, uint64_t &rng, uint8_t rounds[][KF_MAX_PLAYERS]) {
    uint8_t cur[KF_MAX_ROUNDS][KF_MAX_PLAYERS];
    int best = 0;
    for (int restart = 0; restart < 64 && best < nr; restart++) {
        uint32_t used[KF_MAX_PLAYERS] = {0};
        int r = 0;
        for (; r < nr; r++) {
            bool placed = false;
            for (int tries = 0; tries < 64 && !placed; tries++) {
                uint32_t freep = (1u << n) - 1;
                int k = 0;
                bool stuck = false;
                while (freep) {
                    const int x = __builtin_ctz(freep);
                    freep &= ~(1u << x);
                    const uint32_t cand = freep & ~used[x];
                    if (!cand) { stuck = true; break; }
                    int pick = (int)(mix(rng) % (uint64_t)__builtin_popcount(cand));
                    uint32_t m = cand;
                    while (pick--) m &= m - 1;
                    const int y = __builtin_ctz(m);
                    freep &= ~(1u << y);
                    cur[r][k++] = (uint8_t)x;
                    cur[r][k++] = (uint8_t)y;
                }
                if (!stuck) {
                    for (int q = 0; q < n; q += 2) {
                        used[cur[r][q]] |= 1u << cur[r][q + 1];
                        used[cur[r][q + 1]] |= 1u << cur[r][q];
                    }
                    placed = true;
                }
            }
            if (!placed) break;
        }
        if (r > best) {
            best = r;
            memcpy(rounds, cur, sizeof(cur[0]) * r);
        }
    }
    return best;
}

// Relabel so that round 0 becomes (0 1)(2 3)..., then normal form.
void normalize(int n, int nr, uint8_t rounds[][KF_MAX_PLAYERS], uint64_t *out) {
    uint8_t lab[KF_MAX_PLAYERS];
    for (int q = 0; q < n; q++) lab[rounds[0][q]] = (uint8_t)q;
    struct Row { uint8_t v[KF_MAX_PLAYERS]; };
    std::vector<Row> rows(nr);
    for (int r = 0; r < nr; r++) {
        uint8_t partner[KF_MAX_PLAYERS];
        for (int q = 0; q < n; q += 2) {
            const uint8_t x = lab[rounds[r][q]], y = lab[rounds[r][q + 1]];
            partner[x] = y;
            partner[y] = x;
        }
        int k = 0;
        for (int v = 0; v < n; v++)
            if (partner[v] > v) { rows[r].v[k++] = (uint8_t)v; rows[r].v[k++] = partner[v]; }
    }
    std::sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.v[1] < b.v[1]; });
    for (int r = 0; r < nr; r++) {
        uint64_t w = 0;
        for (int q = 0; q < n; q++) w |= (uint64_t)rows[r].v[q] << (4 * q);
        out[r] = w;
    }
}

} // namespace
} // namespace ise

extern "C" int ise_kf_random(int players, int rmin, int rmax, uint64_t count, uint64_t seed, uint32_t stride,
                             uint64_t *rounds_out, uint32_t *meta_out) {
    using namespace ise;
    if (players < 4 || players > (int)KF_MAX_PLAYERS || (players & 1) || rmin < 1 || rmax < rmin ||
        rmax > players - 1 || rmax > (int)KF_MAX_ROUNDS || (uint32_t)rmax > stride || !rounds_out || !meta_out)
        return ISE_ERR_ARG;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
    if (count < 1024) nt = 1;
    auto work = [&](unsigned t) {
        uint8_t rounds[KF_MAX_ROUNDS][KF_MAX_PLAYERS];
        for (uint64_t i = t; i < count; i += nt) {
            uint64_t rng = seed ^ (0xa0761d6478bd642full * (i + 1));
            mix(rng);
            const int target = rmin + (int)(mix(rng) % (uint64_t)(rmax - rmin + 1));
            const int nr = random_rounds(players, target, rng, rounds);
            uint64_t *dst = rounds_out + i * stride;
            memset(dst, 0, stride * sizeof(uint64_t));
            normalize(players, nr, rounds, dst);
            meta_out[i] = (uint32_t)nr | ((uint32_t)players << 8);
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return ISE_OK;
}

// Seeded relabelled prefixes of perfect 1-factorizations: for candidate i a
// source matrix, an ordered pair of its rounds (r0, r1) and a start player k
// are drawn; the cycle of rounds r0, r1 from k is sent onto the cycle of the
// source's rounds 0, 1 from player 0, the matrix is put in normal form and its
// first nr rounds (nr uniform in [rmin, rounds of the source]) are kept.  With
// sources whose rounds 0, 1 are the standard pair and whose round r pairs 0
// with r + 1 (as the reference prints them), every candidate has that form too.
extern "C" int ise_kf_relabel_p1f(const uint64_t *src, const uint32_t *src_meta, uint32_t nsrc, uint32_t src_stride,
                                  uint64_t count, int rmin, uint64_t seed, uint32_t stride, uint64_t *rounds_out,
                                  uint32_t *meta_out) {
    using namespace ise;
    if (!src || !src_meta || nsrc == 0 || src_stride < 1 || rmin < 1 || !rounds_out || !meta_out) return ISE_ERR_ARG;
    for (uint32_t s = 0; s < nsrc; s++) {
        const int nr = (int)(src_meta[s] & 0xff), n = (int)((src_meta[s] >> 8) & 0xff);
        if (n < 4 || n > (int)KF_MAX_PLAYERS || (n & 1) || nr < 2 || nr > n - 1 || nr < rmin ||
            (uint32_t)nr > src_stride || (uint32_t)nr > stride)
            return ISE_ERR_ARG;
    }
    for (uint64_t i = 0; i < count; i++) {
        uint64_t rng = seed ^ (0xe7037ed1a0b428dbull * (i + 1));
        mix(rng);
        const uint32_t s = (uint32_t)(mix(rng) % nsrc);
        const int nr = (int)(src_meta[s] & 0xff), n = (int)((src_meta[s] >> 8) & 0xff);
        uint8_t part[KF_MAX_ROUNDS][KF_MAX_PLAYERS];
        for (int r = 0; r < nr; r++) {
            const uint64_t w = src[(uint64_t)s * src_stride + r];
            uint32_t cover = 0;
            for (int q = 0; q < n; q += 2) {
                const uint8_t x = (uint8_t)((w >> (4 * q)) & 0xf), y = (uint8_t)((w >> (4 * q + 4)) & 0xf);
                if (x >= n || y >= n) return ISE_ERR_ARG;
                part[r][x] = y;
                part[r][y] = x;
                cover |= (1u << x) | (1u << y);
            }
            if (cover != (1u << n) - 1) return ISE_ERR_ARG;   // not a perfect matching
        }
        const int r0 = (int)(mix(rng) % (uint64_t)nr);
        int r1 = (int)(mix(rng) % (uint64_t)(nr - 1));
        if (r1 >= r0) r1++;
        int v1 = (int)(mix(rng) % (uint64_t)n), it = 0;
        uint8_t tr[KF_MAX_PLAYERS];
        uint32_t seen = 0;
        for (int st = 0; st < n; st += 2) {   // the walk of create2P1FTr
            tr[v1] = (uint8_t)it;
            const int v2 = part[r0][v1];
            it = part[0][it];
            tr[v2] = (uint8_t)it;
            it = part[1][it];
            seen |= (1u << v1) | (1u << v2);
            v1 = part[r1][v2];
        }
        if (seen != (1u << n) - 1) return ISE_ERR_ARG;   // rounds r0, r1 not one cycle
        // normal form of the relabelled matrix: pairs sorted, pairs by first
        // player, rounds by the partner of player 0
        struct Row { uint8_t v[KF_MAX_PLAYERS]; };
        Row rows[KF_MAX_ROUNDS];
        for (int r = 0; r < nr; r++) {
            uint8_t q[KF_MAX_PLAYERS];
            for (int x = 0; x < n; x++) q[tr[x]] = tr[part[r][x]];
            int k = 0;
            for (int v = 0; v < n; v++)
                if (q[v] > v) { rows[r].v[k++] = (uint8_t)v; rows[r].v[k++] = q[v]; }
        }
        std::sort(rows, rows + nr, [](const Row &x, const Row &y) { return x.v[1] < y.v[1]; });
        uint64_t full[KF_MAX_ROUNDS];
        for (int r = 0; r < nr; r++) {
            uint64_t w = 0;
            for (int q = 0; q < n; q++) w |= (uint64_t)rows[r].v[q] << (4 * q);
            full[r] = w;
        }
        const int keep = rmin + (int)(mix(rng) % (uint64_t)(nr - rmin + 1));
        uint64_t *dst = rounds_out + i * stride;
        memset(dst, 0, stride * sizeof(uint64_t));
        for (int r = 0; r < keep; r++) dst[r] = full[r];
        meta_out[i] = (uint32_t)keep | ((uint32_t)n << 8);
    }
    return ISE_OK;
}

// Seeded relabelled prefixes of uniform 1-factorizations: as
// ise_kf_relabel_p1f, with the drawn pair of rounds sent onto the source's
// rounds 0, 1 cycle by cycle (cycles of one length in a random assignment, a
// random start of the right parity and a random direction per cycle).  Every
// pair of rounds of a source must have the cycle lengths of its rounds 0, 1.
extern "C" int ise_kf_relabel_u1f(const uint64_t *src, const uint32_t *src_meta, uint32_t nsrc, uint32_t src_stride,
                                  uint64_t count, int rmin, uint64_t seed, uint32_t stride, uint64_t *rounds_out,
                                  uint32_t *meta_out) {
    using namespace ise;
    if (!src || !src_meta || nsrc == 0 || src_stride < 1 || rmin < 1 || !rounds_out || !meta_out) return ISE_ERR_ARG;
    for (uint32_t s = 0; s < nsrc; s++) {
        const int nr = (int)(src_meta[s] & 0xff), n = (int)((src_meta[s] >> 8) & 0xff);
        if (n < 4 || n > (int)KF_MAX_PLAYERS || (n & 1) || nr < 2 || nr > n - 1 || nr < rmin ||
            (uint32_t)nr > src_stride || (uint32_t)nr > stride)
            return ISE_ERR_ARG;
    }
    // cycles of rounds a, b: players in order from the smallest one not yet met, a-partner first
    struct Cyc { int len; uint8_t p[KF_MAX_PLAYERS]; };
    auto cycles = [](const uint8_t *pa, const uint8_t *pb, int n, Cyc *out) {
        uint32_t seen = 0;
        int nc = 0;
        for (int k0 = 0; k0 < n; k0++) {
            if ((seen >> k0) & 1u) continue;
            Cyc c{0, {0}};
            int x = k0;
            do {
                const int y = pa[x];
                c.p[c.len++] = (uint8_t)x;
                c.p[c.len++] = (uint8_t)y;
                seen |= (1u << x) | (1u << y);
                x = pb[y];
            } while (x != k0);
            int pos = nc;   // stable by length
            while (pos > 0 && out[pos - 1].len > c.len) { out[pos] = out[pos - 1]; pos--; }
            out[pos] = c;
            nc++;
        }
        return nc;
    };
    for (uint64_t i = 0; i < count; i++) {
        uint64_t rng = seed ^ (0xc2b2ae3d27d4eb4full * (i + 1));
        mix(rng);
        const uint32_t s = (uint32_t)(mix(rng) % nsrc);
        const int nr = (int)(src_meta[s] & 0xff), n = (int)((src_meta[s] >> 8) & 0xff);
        uint8_t part[KF_MAX_ROUNDS][KF_MAX_PLAYERS];
        for (int r = 0; r < nr; r++) {
            const uint64_t w = src[(uint64_t)s * src_stride + r];
            uint32_t cover = 0;
            for (int q = 0; q < n; q += 2) {
                const uint8_t x = (uint8_t)((w >> (4 * q)) & 0xf), y = (uint8_t)((w >> (4 * q + 4)) & 0xf);
                if (x >= n || y >= n) return ISE_ERR_ARG;
                part[r][x] = y;
                part[r][y] = x;
                cover |= (1u << x) | (1u << y);
            }
            if (cover != (1u << n) - 1) return ISE_ERR_ARG;
        }
        const int r0 = (int)(mix(rng) % (uint64_t)nr);
        int r1 = (int)(mix(rng) % (uint64_t)(nr - 1));
        if (r1 >= r0) r1++;
        Cyc ref[KF_MAX_PLAYERS / 4 + 1], cyc[KF_MAX_PLAYERS / 4 + 1];
        const int nc = cycles(part[0], part[1], n, ref);
        if (cycles(part[r0], part[r1], n, cyc) != nc) return ISE_ERR_ARG;
        for (int c = 0; c < nc; c++)
            if (cyc[c].len != ref[c].len) return ISE_ERR_ARG;
        // shuffle the cycles inside each group of equal length
        for (int g0 = 0, g1; g0 < nc; g0 = g1) {
            for (g1 = g0; g1 < nc && cyc[g1].len == cyc[g0].len; g1++) {}
            for (int j = g1 - 1; j > g0; j--) std::swap(cyc[j], cyc[g0 + (int)(mix(rng) % (uint64_t)(j - g0 + 1))]);
        }
        uint8_t tr[KF_MAX_PLAYERS];
        for (int c = 0; c < nc; c++) {
            const int L = ref[c].len;
            const int d = 2 * (int)(mix(rng) % (uint64_t)(L / 2));
            const bool back = mix(rng) & 1u;
            for (int m = 0; m < L; m++) tr[cyc[c].p[back ? ((d + 1 - m) % L + L) % L : (d + m) % L]] = ref[c].p[m];
        }
        struct Row { uint8_t v[KF_MAX_PLAYERS]; };
        Row rows[KF_MAX_ROUNDS];
        for (int r = 0; r < nr; r++) {
            uint8_t q[KF_MAX_PLAYERS];
            for (int x = 0; x < n; x++) q[tr[x]] = tr[part[r][x]];
            int k = 0;
            for (int v = 0; v < n; v++)
                if (q[v] > v) { rows[r].v[k++] = (uint8_t)v; rows[r].v[k++] = q[v]; }
        }
        std::sort(rows, rows + nr, [](const Row &x, const Row &y) { return x.v[1] < y.v[1]; });
        const int keep = rmin + (int)(mix(rng) % (uint64_t)(nr - rmin + 1));
        uint64_t *dst = rounds_out + i * stride;
        memset(dst, 0, stride * sizeof(uint64_t));
        for (int r = 0; r < keep; r++) {
            uint64_t w = 0;
            for (int q = 0; q < n; q++) w |= (uint64_t)rows[r].v[q] << (4 * q);
            dst[r] = w;
        }
        meta_out[i] = (uint32_t)keep | ((uint32_t)n << 8);
    }
    return ISE_OK;
}
