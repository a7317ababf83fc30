// ise_enum.cpp -- host-side candidate producers that feed the GPU checker.
//
//  * A level-synchronous row-by-row BIBD enumerator.  The reference extends a
//    matrix depth-first and tests canonicity after every row
//    (CEnumerator::Enumerate, Source/Enumerator.cpp:540-1024;
//    ProcessPartiallyConstructedMatrix :435-501,
//    ProcessFullyConstructedMatrix :350-433).  Here every canonical matrix of
//    depth d is extended by all solutions of the next-row system, and the
//    whole next level is checked as GPU batches.  Since every prefix of a
//    canonical matrix is canonical, both walks reach the same canonical
//    designs.
//  * The next-row system (the reference builds and solves it with
//    C_InSysEnumerator::MakeSystem / CInSysSolver, Source/InsSysEnumerator.h:140-471,
//    Source/InSysSolver.h:92-207): one unknown x_o per column orbit of the
//    current matrix, 0 <= x_o <= |o|, sum x_o = r, and for every earlier row i
//    the orbits holding a one in row i receive exactly lambda ones.  The row
//    is written with the x_o ones left-justified in each orbit, as
//    CMatrixCanonChecker::MakeRow does (Source/MatrixCanonChecker.cpp:14-93).
//  * Seeded random valid extensions for stress batches.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>
#include "ise_internal.h"
#include "ise_random.h"

namespace {

struct Design {
    int v, b, r, k, lambda;
};

bool derive(int v, int k, int lambda, Design *d) {
    if (v < 2 || k < 2 || k > v || lambda < 1) return false;
    if ((lambda * (v - 1)) % (k - 1)) return false;
    const int r = lambda * (v - 1) / (k - 1);
    if ((v * r) % k) return false;
    const int b = v * r / k;
    if (v > ISE_MAX_ROWS || b > ISE_MAX_COLS) return false;
    *d = Design{v, b, r, k, lambda};
    return true;
}

inline uint64_t low_bits(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

// Solves the next-row system of one parent and appends every solution row.
class RowSolver {
  public:
    RowSolver(const Design &ds) : ds_(ds) {}

    // parent: rows[0..depth-1]; appends solution rows to `out`.
    void solve(const uint64_t *rows, int depth, std::vector<uint64_t> &out) {
        depth_ = depth;
        rows_ = rows;
        out_ = &out;
        // column orbits of row `depth`
        uint64_t starts = 1;
        for (int i = 0; i < depth; i++) starts |= rows[i] ^ (rows[i] << 1);
        starts &= low_bits(ds_.b);
        norb_ = 0;
        const int after = ds_.v - depth - 1;   // rows still to come after this one
        for (uint64_t m = starts; m; m &= m - 1) {
            const int s = __builtin_ctzll(m);
            const uint64_t rest = m & (m - 1);
            const int e = rest ? __builtin_ctzll(rest) : ds_.b;
            Orb &o = orb_[norb_++];
            o.start = s;
            o.len = e - s;
            o.rowmask = 0;
            int w = 0;
            for (int i = 0; i < depth; i++)
                if ((rows[i] >> s) & 1) { w++; o.rowmask |= 1u << i; }
            const int need = ds_.k - w;
            if (need < 0 || need > after + 1) return;     // dead parent
            o.hi = need > 0 ? o.len : 0;
            o.lo = need > after ? o.len : 0;
        }
        // suffix capacities for pruning
        cap_r_[norb_] = 0;
        for (int i = 0; i < depth; i++) cap_l_[norb_][i] = 0;
        for (int o = norb_ - 1; o >= 0; o--) {
            cap_r_[o] = cap_r_[o + 1] + orb_[o].hi;
            for (int i = 0; i < depth; i++)
                cap_l_[o][i] = cap_l_[o + 1][i] + (((orb_[o].rowmask >> i) & 1) ? orb_[o].hi : 0);
        }
        for (int i = 0; i < depth; i++) lam_[i] = 0;
        place(0, 0, 0ull);
    }

  private:
    struct Orb {
        int start, len, lo, hi;
        uint32_t rowmask;
    };

    void place(int o, int sum, uint64_t row) {
        if (o == norb_) {
            if (sum != ds_.r) return;
            for (int i = 0; i < depth_; i++)
                if (lam_[i] != ds_.lambda) return;
            out_->push_back(row);
            return;
        }
        const Orb &ob = orb_[o];
        for (int x = ob.hi; x >= ob.lo; x--) {
            const int ns = sum + x;
            if (ns > ds_.r || ns + cap_r_[o + 1] < ds_.r) continue;
            bool ok = true;
            for (int i = 0; i < depth_ && ok; i++) {
                const int nl = lam_[i] + (((ob.rowmask >> i) & 1) ? x : 0);
                ok = nl <= ds_.lambda && nl + cap_l_[o + 1][i] >= ds_.lambda;
            }
            if (!ok) continue;
            for (int i = 0; i < depth_; i++)
                if ((ob.rowmask >> i) & 1) lam_[i] += x;
            place(o + 1, ns, row | (low_bits(x) << ob.start));
            for (int i = 0; i < depth_; i++)
                if ((ob.rowmask >> i) & 1) lam_[i] -= x;
        }
    }

    Design ds_;
    int depth_ = 0, norb_ = 0;
    const uint64_t *rows_ = nullptr;
    std::vector<uint64_t> *out_ = nullptr;
    Orb orb_[ISE_MAX_COLS];
    int cap_r_[ISE_MAX_COLS + 1];
    int cap_l_[ISE_MAX_COLS + 1][ISE_MAX_ROWS];
    int lam_[ISE_MAX_ROWS];
};

unsigned host_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return t > 16 ? 16 : t;
}

// Solution-orbit pruning (CRowSolution::sortSolutionsByGroup,
// Source/RowSolution.h:325-383, with CPermutStorage::constructGroup /
// findSolutionIndex, Source/PermutStorage.h:183-325): the automorphisms of a
// canonical parent permute the column orbits of its next row, so a solution
// (ones per orbit) and its images give isomorphic children with the same
// first rows, and among those only the child whose solution is the
// lexicographically largest image can be canonical.  The other solutions are
// dropped before their children are built.  gens: n_gens permutations of the
// next row's orbits (entry o = the orbit that orbit o goes to); the orbit of a
// solution is walked breadth-first under them (image[o] = sol[g[o]]).
constexpr uint32_t PRUNE_GEN_CAP = 32;

struct ParentGroup {
    uint32_t n_gens = 0;
    const uint8_t *gens = nullptr;   // [n_gens][ISE_MAX_COLS]
};

bool largest_in_orbit(const std::string &sol, const ParentGroup &g) {
    if (!g.n_gens) return true;
    std::unordered_set<std::string> seen{sol};
    std::vector<std::string> todo{sol};
    std::string img(sol.size(), '\0');
    while (!todo.empty()) {
        const std::string x = todo.back();
        todo.pop_back();
        for (uint32_t q = 0; q < g.n_gens; q++) {
            const uint8_t *perm = g.gens + (size_t)q * ISE_MAX_COLS;
            for (size_t o = 0; o < x.size(); o++) img[o] = perm[o] < x.size() ? x[perm[o]] : x[o];
            if (img > sol) return false;
            if (seen.insert(img).second) todo.push_back(img);
        }
    }
    return true;
}

// Ones of `row` per column orbit of row `depth` of the parent.
std::string solution_counts(const uint64_t *parent, int depth, int b, uint64_t row) {
    uint64_t starts = 1;
    for (int i = 0; i < depth; i++) starts |= parent[i] ^ (parent[i] << 1);
    starts &= low_bits(b);
    std::string out;
    for (uint64_t m = starts; m; m &= m - 1) {
        const int s = __builtin_ctzll(m);
        const uint64_t rest = m & (m - 1);
        const int e = rest ? __builtin_ctzll(rest) : b;
        out.push_back((char)__builtin_popcountll(row & (low_bits(e - s) << s)));
    }
    return out;
}

// Children of every canonical parent (rows packed with `stride` words each);
// with groups (one per parent), only the children of orbit-largest solutions.
void extend_level(const Design &ds, const std::vector<uint64_t> &parents, int depth, uint32_t stride,
                  std::vector<uint64_t> &children, const std::vector<ParentGroup> *groups = nullptr) {
    const size_t np = parents.size() / stride;
    const unsigned nt = np < 64 ? 1u : host_threads();
    std::vector<std::vector<uint64_t>> part(nt);
    auto work = [&](unsigned t) {
        RowSolver solver(ds);
        std::vector<uint64_t> sol;
        const size_t lo = np * t / nt, hi = np * (t + 1) / nt;
        for (size_t p = lo; p < hi; p++) {
            sol.clear();
            const uint64_t *pr = &parents[p * stride];
            solver.solve(pr, depth, sol);
            for (uint64_t row : sol) {
                if (groups && !largest_in_orbit(solution_counts(pr, depth, ds.b, row), (*groups)[p]))
                    continue;
                const size_t at = part[t].size();
                part[t].resize(at + stride, 0ull);
                memcpy(&part[t][at], pr, depth * sizeof(uint64_t));
                part[t][at + depth] = row;
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    // contiguous parent ranges per thread: children stay in parent order
    children.clear();
    for (auto &p : part) children.insert(children.end(), p.begin(), p.end());
}

struct EnumSink {
    uint64_t *designs = nullptr, *auts = nullptr;
    long max_designs = 0, ndesigns = 0;
    int dmin = 0, dmax = -1;
    uint64_t *cand_rows = nullptr;
    uint32_t *cand_meta = nullptr;
    long cand_cap = 0, ncand = 0;
    uint32_t cand_stride = 0;
};

inline uint64_t splitmix(uint64_t &s);

// Keep `beam` of the canonical candidates of a level (beam 0 keeps all):
// partial Fisher-Yates over their indices driven by splitmix64, the chosen
// indices then restored to level order.  The generator state carries over.
void sample_beam(std::vector<uint64_t> &keep, uint32_t stride, uint32_t beam, uint64_t &rng) {
    const size_t n = keep.size() / stride;
    if (beam == 0 || n <= beam) return;
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
    for (uint32_t t = 0; t < beam; t++) {
        const size_t j = t + (size_t)(splitmix(rng) % (uint64_t)(n - t));
        std::swap(idx[t], idx[j]);
    }
    idx.resize(beam);
    std::sort(idx.begin(), idx.end());
    std::vector<uint64_t> out((size_t)beam * stride);
    for (uint32_t t = 0; t < beam; t++)
        memcpy(&out[(size_t)t * stride], &keep[(size_t)idx[t] * stride], stride * sizeof(uint64_t));
    keep.swap(out);
}

long run_enumeration(ise_ctx *ctx, const Design &ds, EnumSink &sink, uint64_t *stats,
                     uint64_t *level_counts, uint32_t stride, uint32_t capacity, uint32_t beam = 0,
                     uint64_t seed = 0, bool prune = false) {
    uint64_t rng = seed;
    uint64_t st[5] = {0, 0, 0, 0, 0};
    double kernel_ms = 0.0;
    if (level_counts)
        for (int i = 0; i <= ds.v; i++) level_counts[i] = 0;
    std::vector<uint64_t> level, next, keep;
    // only verdicts and |Aut| are read back: no column permutations
    struct ColumnsOff {
        ise_ctx *c;
        bool prev;
        ~ColumnsOff() { ise::ctx_report_columns(c, prev); }
    } columns_off{ctx, ise::ctx_report_columns(ctx, false)};
    // depth 0 -> candidates of depth 1
    std::vector<uint64_t> root(stride, 0ull);
    extend_level(ds, root, 0, stride, level);
    std::vector<uint32_t> meta;
    std::vector<uint8_t> verdict;
    std::vector<uint64_t> order;
    // with pruning: the generators of every canonical candidate of the level
    std::vector<uint16_t> ngen;
    std::vector<uint8_t> gens, keep_gens;
    std::vector<ParentGroup> groups;
    for (int depth = 1; depth <= ds.v && !level.empty(); depth++) {
        const size_t n = level.size() / stride;
        const uint32_t out_info = depth == ds.v ? (ISE_SAVE_ROW_PERMUTATIONS | ISE_SAVE_ROW_TO_CHANGE) : 0u;
        const uint32_t m = (uint32_t)depth | ((uint32_t)ds.b << 8) | (out_info << 24);
        keep.clear();
        keep_gens.clear();
        const bool want_gens = prune && depth < ds.v;
        if (level_counts) level_counts[depth] = n;
        st[0] += n;
        if (depth == ds.v) st[2] += n;
        if (sink.cand_rows && depth >= sink.dmin && depth <= sink.dmax) {
            for (size_t c = 0; c < n && sink.ncand < sink.cand_cap; c++, sink.ncand++) {
                uint64_t *dst = sink.cand_rows + (size_t)sink.ncand * sink.cand_stride;
                memset(dst, 0, sink.cand_stride * sizeof(uint64_t));
                memcpy(dst, &level[c * stride], depth * sizeof(uint64_t));
                sink.cand_meta[sink.ncand] = m;
            }
        }
        for (size_t off = 0; off < n; off += capacity) {
            const uint32_t cnt = (uint32_t)std::min<size_t>(capacity, n - off);
            meta.assign(cnt, m);
            verdict.resize(cnt);
            order.resize(cnt);
            const uint32_t gen_cap = want_gens ? PRUNE_GEN_CAP : 0u;
            int rc = ise_submit(ctx, cnt, stride, &level[off * stride], nullptr, meta.data(), gen_cap);
            if (rc) return rc;
            ise_results h{};
            h.verdict = verdict.data();
            h.group_order = order.data();
            if (want_gens) {
                ngen.resize(cnt);
                gens.resize((size_t)cnt * gen_cap * ISE_MAX_COLS);
                h.n_gens = ngen.data();
                h.gens = gens.data();
                h.gen_cap = gen_cap;
            }
            rc = ise_collect(ctx, &h);
            if (rc) return rc;
            kernel_ms += ise_last_kernel_ms(ctx);
            for (uint32_t c = 0; c < cnt; c++) {
                if (verdict[c] == ISE_VERDICT_CANONICAL) {
                    const uint64_t *src = &level[(off + c) * stride];
                    if (depth == ds.v) {
                        if (sink.designs && sink.ndesigns < sink.max_designs) {
                            memcpy(sink.designs + (size_t)sink.ndesigns * ds.v, src, ds.v * sizeof(uint64_t));
                            if (sink.auts) sink.auts[sink.ndesigns] = order[c];
                        }
                        sink.ndesigns++;
                        st[3]++;
                    } else {
                        keep.insert(keep.end(), src, src + stride);
                        if (want_gens) {
                            // this parent's generators, as many as were stored
                            const uint32_t ng = std::min<uint32_t>(ngen[c], PRUNE_GEN_CAP);
                            const uint8_t *gp = &gens[(size_t)c * PRUNE_GEN_CAP * ISE_MAX_COLS];
                            keep_gens.push_back((uint8_t)ng);
                            keep_gens.insert(keep_gens.end(), gp, gp + (size_t)PRUNE_GEN_CAP * ISE_MAX_COLS);
                        }
                        st[1]++;
                    }
                } else if (verdict[c] != ISE_VERDICT_NONCANONICAL) {
                    return ISE_ERR_ARG;   // search limit or invalid record: refuse to guess
                }
            }
        }
        sample_beam(keep, stride, beam, rng);
        if (depth < ds.v) {
            if (want_gens && beam == 0) {
                const size_t rec = 1 + (size_t)PRUNE_GEN_CAP * ISE_MAX_COLS;
                groups.assign(keep.size() / stride, ParentGroup{});
                for (size_t p = 0; p < groups.size(); p++)
                    groups[p] = ParentGroup{keep_gens[p * rec], &keep_gens[p * rec + 1]};
                extend_level(ds, keep, depth, stride, next, &groups);
            } else {
                extend_level(ds, keep, depth, stride, next);
            }
        } else {
            next.clear();
        }
        level.swap(next);
    }
    st[4] = (uint64_t)(kernel_ms * 1e6);
    if (stats) memcpy(stats, st, sizeof st);
    return sink.ndesigns;
}

// ---- seeded random valid extensions (the walk: ise_random.h) ---------------

inline uint64_t splitmix(uint64_t &s) { return ise_rw::splitmix(s); }

} // namespace

extern "C" {

long ise_enumerate_bibd(ise_ctx *ctx, int v, int k, int lambda, uint64_t *designs,
                        uint64_t *auts, long max_designs, uint64_t *stats, uint64_t *level_counts) {
    Design ds;
    if (!ctx || !derive(v, k, lambda, &ds)) return ISE_ERR_ARG;
    EnumSink sink;
    sink.designs = designs;
    sink.auts = auts;
    sink.max_designs = max_designs;
    const uint32_t stride = ise::ctx_stride(ctx);
    if ((int)stride < v) return ISE_ERR_ARG;
    return run_enumeration(ctx, ds, sink, stats, level_counts, stride, ise::ctx_capacity(ctx));
}

long ise_enumerate_bibd_pruned(ise_ctx *ctx, int v, int k, int lambda, uint64_t *designs, uint64_t *auts,
                               long max_designs, uint64_t *stats, uint64_t *level_counts) {
    Design ds;
    if (!ctx || !derive(v, k, lambda, &ds)) return ISE_ERR_ARG;
    EnumSink sink;
    sink.designs = designs;
    sink.auts = auts;
    sink.max_designs = max_designs;
    const uint32_t stride = ise::ctx_stride(ctx);
    if ((int)stride < v) return ISE_ERR_ARG;
    return run_enumeration(ctx, ds, sink, stats, level_counts, stride, ise::ctx_capacity(ctx), 0, 0, true);
}

long ise_enumerate_bibd_candidates(ise_ctx *ctx, int v, int k, int lambda, int dmin, int dmax,
                                   uint32_t stride, uint64_t *rows_out, uint32_t *meta_out,
                                   long cap, long *n_out) {
    Design ds;
    if (!ctx || !derive(v, k, lambda, &ds) || !rows_out || !meta_out || !n_out ||
        stride < (uint32_t)dmax || dmin < 1 || dmax < dmin)
        return ISE_ERR_ARG;
    const uint32_t cstride = ise::ctx_stride(ctx);
    if ((int)cstride < v) return ISE_ERR_ARG;
    EnumSink sink;
    sink.dmin = dmin;
    sink.dmax = dmax;
    sink.cand_rows = rows_out;
    sink.cand_meta = meta_out;
    sink.cand_cap = cap;
    sink.cand_stride = stride;
    long rc = run_enumeration(ctx, ds, sink, nullptr, nullptr, cstride, ise::ctx_capacity(ctx));
    *n_out = sink.ncand;
    return rc;
}

long ise_enumerate_bibd_sampled(ise_ctx *ctx, int v, int k, int lambda, uint32_t beam, uint64_t seed, int dmin,
                                int dmax, uint32_t stride, uint64_t *rows_out, uint32_t *meta_out, long cap,
                                long *n_out, uint64_t *level_counts) {
    Design ds;
    if (!ctx || !derive(v, k, lambda, &ds) || !rows_out || !meta_out || !n_out || beam == 0 ||
        stride < (uint32_t)dmax || dmin < 1 || dmax < dmin)
        return ISE_ERR_ARG;
    const uint32_t cstride = ise::ctx_stride(ctx);
    if ((int)cstride < v) return ISE_ERR_ARG;
    EnumSink sink;
    sink.dmin = dmin;
    sink.dmax = dmax;
    sink.cand_rows = rows_out;
    sink.cand_meta = meta_out;
    sink.cand_cap = cap;
    sink.cand_stride = stride;
    long rc = run_enumeration(ctx, ds, sink, nullptr, level_counts, cstride, ise::ctx_capacity(ctx), beam, seed);
    *n_out = sink.ncand;
    return rc < 0 ? rc : sink.ncand;
}

int ise_random_extensions_ex(ise_ctx *ctx, int v, int k, int lambda, uint64_t n, int dmin, int dmax,
                             uint64_t seed, uint32_t attempts, uint32_t stride, uint64_t *rows_out,
                             uint32_t *meta_out) {
    Design ds;
    if (!derive(v, k, lambda, &ds) || !rows_out || !meta_out || dmin < 1 || dmax < dmin ||
        dmax > v || (uint32_t)dmax > stride || attempts < 1 || attempts > ISE_RANDOM_MAX_ATTEMPTS)
        return ISE_ERR_ARG;
    const ise_rw::Params p{ds.v, ds.b, ds.r, ds.k, ds.lambda};
    if (ctx) return ise::random_extensions_device(ctx, p, n, dmin, dmax, seed, attempts, stride, rows_out, meta_out);
    const unsigned nt = n < 1024 ? 1u : host_threads();
    auto work = [&](unsigned t) {
        for (uint64_t i = t; i < n; i += nt)
            ise_rw::batch_candidate(p, i, dmin, dmax, seed, attempts, stride, rows_out, meta_out);
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return ISE_OK;
}

int ise_random_extensions(int v, int k, int lambda, uint64_t n, int dmin, int dmax, uint64_t seed,
                          uint32_t stride, uint64_t *rows_out, uint32_t *meta_out) {
    return ise_random_extensions_ex(nullptr, v, k, lambda, n, dmin, dmax, seed, 1u, stride, rows_out, meta_out);
}

} // extern "C"
