// ise_cdtools.hip -- the k-Sys binary canonizer on the GPU (createCanonizer /
// runCanonizer / releaseCanonizer, Source/CDTools.cpp:5-51).
//
// runCanonizer builds the binary matrix of a k-system (C_InSys::
// convertToBinaryMatrix, Source/matrix.h:198-210) under a first row of day
// labels and calls CMatrixCanonChecker::CanonizeMatrix(k, NULL, numClasses)
// (MatrixCanonChecker.cpp:420-564).  The label row makes every day's columns
// a column orbit; here it is the initial column partition of the batched
// canonical-form loop (ise_canonize_batch_device_part), so the matrices hold
// the v player rows only.  With every day present the reference also walks
// the day permutations, canonizing each and keeping the largest matrix
// (the walk starts again whenever the best matrix grows, so it ends on the
// largest one over all day orders); here the days! day orders of a k-system
// are canonized in one batch and a reduction keeps the largest.
//
// Per chunk of candidates (one candidate = a k-system in one day order):
//   cd_expand_kernel        one wave per candidate: lane c builds column c
//                           (day j = c / (v/k), group c % (v/k) of the day in
//                           position j) as a mask of players, and v ballots
//                           give the packed rows;
//   canonical-form loop     ise_canonize_batch_device_part, day partition;
//   cd_segment_best_kernel  one workgroup per run of <= CD_SEG candidates of
//                           one k-system: the largest canonical matrix (rows
//                           compared in order, each with column 0 first) and
//                           how many candidates of the run equal it;
//   cd_merge_kernel         one thread per k-system: the run winners against
//                           the k-system's best so far (kept across chunks),
//                           with the count of day orders that reach it (the
//                           order of the day-permutation group CanonizeMatrix
//                           stores in extraGroupOrder) and its |Aut| with the
//                           days fixed.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <vector>
#include "ise.h"
#include "ise_internal.h"

struct ise_cd_canonizer {
    ise_ctx *ctx = nullptr;
    int device = 0;
    uint32_t v = 0, k = 0, days_max = 0, groups = 0, chunk = 0;
    uint8_t *d_mats = nullptr;       // host batches staged here
    size_t mats_cap = 0;             // bytes
    uint64_t *d_rows = nullptr, *d_canon = nullptr;
    uint32_t *d_meta = nullptr;
    uint8_t *d_status = nullptr;
    uint64_t *d_aut = nullptr;       // [chunk] |Aut| of each canonical matrix, days fixed
    uint32_t *d_seg = nullptr;       // [nseg][4]: first candidate (in the chunk), length, winner, ties
    size_t seg_cap = 0;
    uint32_t *d_mseg = nullptr;      // [nm][3]: k-system, first run, runs
    size_t mseg_cap = 0;
    uint64_t *d_best = nullptr;      // [n][v] best matrix of every k-system of a host batch
    size_t best_cap = 0;             // k-systems
    uint32_t *d_have = nullptr;      // [n] a best matrix is set
    size_t have_cap = 0;
    uint64_t *d_orders = nullptr;    // [2][n] day orders reaching the best matrix, its |Aut| days fixed
    size_t orders_cap = 0;           // k-systems
    uint64_t last_day_order = 1, last_fixed_order = 0;
    uint32_t *d_flags = nullptr;     // [0] a player id >= v, [1] a canonization did not end
    std::vector<uint64_t> packed;
    std::vector<uint8_t> out;
    std::string err;
};

namespace {

constexpr uint32_t CD_CHUNK = 1u << 20;   // candidates per canonical-form batch, at most
constexpr uint32_t CD_CHUNK_MIN = 1u << 12;   // ... and at first: the batch grows with the runs
constexpr uint32_t CD_SEG = 1024u;        // candidates per reduction workgroup
constexpr uint32_t CD_NONE = 0xffffffffu;

constexpr uint32_t SEG_WORDS = 4;

__device__ inline bool cd_equal(const uint64_t *x, const uint64_t *y, uint32_t v) {
    for (uint32_t i = 0; i < v; i++)
        if (x[i] != y[i]) return false;
    return true;
}

// x > y for two v-row matrices (rows in order, each with column 0 first).
__device__ inline bool cd_larger(const uint64_t *x, const uint64_t *y, uint32_t v) {
    for (uint32_t i = 0; i < v; i++) {
        const uint64_t a = x[i], b = y[i];
        if (a != b) {
            const uint64_t diff = a ^ b;
            return (a & (diff & (0ull - diff))) != 0ull;
        }
    }
    return false;
}

// Candidate c0 + w of the chunk: k-system c / P in day order c % P (the
// factorial number system, most significant position first; P = 1: the
// given order).  Day positions are packed 4 bits each (days <= 10 for a walk).
__global__ __launch_bounds__(256) void cd_expand_kernel(const uint8_t *mats, uint32_t v, uint32_t k, uint32_t days,
                                                        uint32_t walk, uint64_t P, uint64_t c0, uint32_t cnt,
                                                        uint32_t stride, uint64_t *rows, uint32_t *meta,
                                                        uint32_t *flags) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= cnt) return;
    const uint64_t c = c0 + w;
    const uint64_t mat = c / P;
    uint32_t t = (uint32_t)(c - mat * P);
    uint64_t order = 0;
    if (walk) {
        uint32_t used = 0, f = 1;
        for (uint32_t i = 2; i < days; i++) f *= i;          // (days - 1)!
        for (uint32_t i = 0; i < days; i++) {
            const uint32_t q = t / f;
            t -= q * f;
            uint32_t pick = 0, seen = 0;
            for (uint32_t dd = 0; dd < days; dd++) {
                if ((used >> dd) & 1u) continue;
                if (seen == q) { pick = dd; break; }
                seen++;
            }
            used |= 1u << pick;
            order |= (uint64_t)pick << (4 * i);
            if (days - 1 - i) f /= days - 1 - i;
        }
    }
    const uint32_t G = v / k, b = days * G;
    uint32_t col = 0;
    if (lane < b) {
        const uint32_t j = lane / G, g = lane - j * G;
        const uint32_t od = walk ? (uint32_t)((order >> (4 * j)) & 15u) : j;
        const uint8_t *grp = mats + (mat * days + od) * v + (uint64_t)g * k;
        for (uint32_t n = 0; n < k; n++) {
            const uint32_t p = grp[n];
            if (p < v) col |= 1u << p;
            else atomicOr(&flags[0], 1u);
        }
    }
    uint64_t word = 0;
    for (uint32_t r = 0; r < v; r++) {
        const uint64_t bw = __builtin_amdgcn_ballot_w64(((col >> r) & 1u) != 0u);
        if (lane == r) word = bw;
    }
    // rows in decreasing order, column 0 most significant (the reference sorts
    // the rows before its loop, sortRowsUpdateColumnOrbits,
    // MatrixCanonChecker.cpp:394-418): the same canonical form, fewer rounds
    const uint64_t key = __builtin_bitreverse64(word);
    uint32_t rank = 0;
    for (uint32_t j = 0; j < v; j++) {
        const uint64_t kj = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(key >> 32), (int)j) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, (int)j);
        rank += kj > key || (kj == key && j < lane);
    }
    if (lane < v) rows[(uint64_t)w * stride + rank] = word;
    else if (lane < stride) rows[(uint64_t)w * stride + lane] = 0ull;
    if (lane == 0) meta[w] = v | (b << 8);
}

__global__ __launch_bounds__(256) void cd_segment_best_kernel(const uint64_t *canon, const uint8_t *status,
                                                              uint32_t stride, uint32_t v, uint32_t *seg,
                                                              uint32_t *flags) {
    __shared__ uint32_t best[256];
    __shared__ uint32_t ties;
    const uint32_t s = blockIdx.x, start = seg[SEG_WORDS * s], len = seg[SEG_WORDS * s + 1];
    uint32_t mine = CD_NONE;
    if (threadIdx.x == 0) ties = 0;
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
        const uint32_t c = start + i;
        if (status[c] != ISE_VERDICT_CANONICAL) {
            atomicOr(&flags[1], 1u);
            continue;
        }
        if (mine == CD_NONE || cd_larger(canon + (uint64_t)c * stride, canon + (uint64_t)mine * stride, v)) mine = c;
    }
    best[threadIdx.x] = mine;
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const uint32_t o = best[threadIdx.x + h], m = best[threadIdx.x];
            if (o != CD_NONE &&
                (m == CD_NONE || cd_larger(canon + (uint64_t)o * stride, canon + (uint64_t)m * stride, v)))
                best[threadIdx.x] = o;
        }
        __syncthreads();
    }
    const uint32_t win = best[0];
    // candidates of the run equal to the winner (day orders that reach it)
    uint32_t eq = 0;
    if (win != CD_NONE)
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
            const uint32_t c = start + i;
            eq += status[c] == ISE_VERDICT_CANONICAL &&
                  cd_equal(canon + (uint64_t)c * stride, canon + (uint64_t)win * stride, v);
        }
    if (eq) atomicAdd(&ties, eq);
    __syncthreads();
    if (threadIdx.x == 0) {
        seg[SEG_WORDS * s + 2] = win;
        seg[SEG_WORDS * s + 3] = ties;
    }
}

__global__ void cd_merge_kernel(const uint64_t *canon, const uint64_t *aut, uint32_t stride, uint32_t v,
                                const uint32_t *seg, const uint32_t *mseg, uint32_t nm, uint64_t *best,
                                uint32_t *have, uint64_t *day_orders, uint64_t *fixed_orders) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nm) return;
    const uint32_t mat = mseg[3 * e], s0 = mseg[3 * e + 1], ns = mseg[3 * e + 2];
    uint64_t *bp = best + (uint64_t)mat * v;
    bool h = have[mat] != 0u;
    uint64_t reach = h ? day_orders[mat] : 0ull, fixed = h ? fixed_orders[mat] : 0ull;
    for (uint32_t s = s0; s < s0 + ns; s++) {
        const uint32_t c = seg[SEG_WORDS * s + 2];
        if (c == CD_NONE) continue;
        const uint64_t *cp = canon + (uint64_t)c * stride;
        if (!h || cd_larger(cp, bp, v)) {
            for (uint32_t i = 0; i < v; i++) bp[i] = cp[i];
            h = true;
            reach = seg[SEG_WORDS * s + 3];
            fixed = aut[c];
        } else if (cd_equal(cp, bp, v)) {
            reach += seg[SEG_WORDS * s + 3];
        }
    }
    have[mat] = h ? 1u : 0u;
    day_orders[mat] = reach;
    fixed_orders[mat] = fixed;
}

// Without a walk every candidate is its own k-system's result.
__global__ void cd_copy_kernel(const uint64_t *canon, const uint8_t *status, const uint64_t *aut, uint32_t stride,
                               uint32_t v, uint64_t m0, uint32_t cnt, uint64_t *best, uint64_t *day_orders,
                               uint64_t *fixed_orders, uint32_t *flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)cnt * v) return;
    const uint64_t c = i / v, w = i - c * v;
    best[(m0 + c) * v + w] = canon[c * stride + w];
    if (w == 0) {
        if (status[c] != ISE_VERDICT_CANONICAL) atomicOr(&flags[1], 1u);
        day_orders[m0 + c] = 1ull;
        fixed_orders[m0 + c] = aut[c];
    }
}

int fail(ise_cd_canonizer *h, int rc, const std::string &msg) {
    h->err = msg;
    return rc;
}

int hip_fail(ise_cd_canonizer *h, hipError_t e, const char *where) {
    h->err = std::string(where) + ": " + hipGetErrorString(e);
    return ISE_ERR_HIP;
}

#define CDCHK(h, call)                                      \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #call); \
    } while (0)

template <typename T> hipError_t grow(T **p, size_t *cap, size_t need) {
    if (*cap >= need) return hipSuccess;
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return e;
        *p = nullptr;
        *cap = 0;
    }
    hipError_t e = hipMalloc((void **)p, need * sizeof(T));
    if (e == hipSuccess) *cap = need;
    return e;
}

uint64_t factorial(uint32_t n) {
    uint64_t f = 1;
    for (uint32_t i = 2; i <= n; i++) f *= i;
    return f;
}

// Days of a call: `day` <= 0 means all of them.
int resolve_days(ise_cd_canonizer *h, int day, uint32_t *days, uint32_t *b, bool *walk) {
    const uint32_t d = day <= 0 ? h->days_max : (uint32_t)day;
    if (d > h->days_max) return fail(h, ISE_ERR_ARG, "runCanonizer: more days than the k-system has");
    const uint32_t cols = d * h->groups;
    if (cols > ISE_MAX_COLS) return fail(h, ISE_ERR_ARG, "runCanonizer: more than 64 columns");
    *walk = d == h->days_max;
    if (*walk && d > ISE_CD_MAX_WALK_DAYS)
        return fail(h, ISE_ERR_ARG, "runCanonizer: too many days for the walk over day orders");
    *days = d;
    *b = cols;
    return ISE_OK;
}

// The canonical-form batch size: a context for `chunk` candidates and the
// staging buffers of that many.  The new ones are made first, so a failure
// leaves the canonizer as it was.
int set_chunk(ise_cd_canonizer *h, uint32_t chunk) {
    ise_ctx *ctx = nullptr;
    int rc = ise_open(&ctx, h->device, chunk, h->v);
    if (rc) return fail(h, rc, "runCanonizer: cannot open a context for the batch");
    const size_t words = (size_t)chunk * h->v;
    uint64_t *rows = nullptr, *canon = nullptr, *aut = nullptr;
    uint32_t *meta = nullptr;
    uint8_t *status = nullptr;
    hipError_t e = hipSetDevice(h->device);
    if (e == hipSuccess) e = hipMalloc((void **)&rows, words * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&canon, words * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&meta, (size_t)chunk * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&status, chunk);
    if (e == hipSuccess) e = hipMalloc((void **)&aut, (size_t)chunk * 8);
    if (e != hipSuccess) {
        void *made[] = {rows, canon, meta, status, aut};
        for (void *p : made)
            if (p) (void)hipFree(p);
        ise_close(ctx);
        return hip_fail(h, e, "runCanonizer: batch buffers");
    }
    if (h->ctx) {
        (void)hipStreamSynchronize(ise::ctx_stream(h->ctx));
        void *old[] = {h->d_rows, h->d_canon, h->d_meta, h->d_status, h->d_aut};
        for (void *p : old)
            if (p) (void)hipFree(p);
        ise_close(h->ctx);
    }
    h->ctx = ctx;
    h->chunk = chunk;
    h->d_rows = rows;
    h->d_canon = canon;
    h->d_meta = meta;
    h->d_status = status;
    h->d_aut = aut;
    return ISE_OK;
}

} // namespace

extern "C" {

int ise_cd_create_canonizer(ise_cd_canonizer **out, int device, uint32_t v, uint32_t k) {
    if (!out) return ISE_ERR_ARG;
    *out = nullptr;
    if (k < 2 || v < k || v % k || (v - 1) % (k - 1) || v > ISE_MAX_ROWS) return ISE_ERR_ARG;
    auto *h = new ise_cd_canonizer();
    h->device = device;
    h->v = v;
    h->k = k;
    h->days_max = (v - 1) / (k - 1);
    h->groups = v / k;
    int rc = set_chunk(h, CD_CHUNK_MIN);   // a small batch first (a missing device fails here)
    if (rc) {
        delete h;
        return rc;
    }
    if (hipMalloc((void **)&h->d_flags, 64) != hipSuccess) {
        ise_cd_release_canonizer(h);
        return ISE_ERR_HIP;
    }
    *out = h;
    return ISE_OK;
}

int ise_cd_release_canonizer(ise_cd_canonizer *h) {
    if (!h) return ISE_ERR_ARG;
    (void)hipSetDevice(h->device);
    if (h->ctx) (void)hipStreamSynchronize(ise::ctx_stream(h->ctx));
    void *bufs[] = {h->d_mats, h->d_rows, h->d_canon, h->d_meta, h->d_status, h->d_aut,
                    h->d_seg,  h->d_mseg, h->d_best,  h->d_have, h->d_orders, h->d_flags};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (h->ctx) ise_close(h->ctx);
    delete h;
    return ISE_OK;
}

const char *ise_cd_error(const ise_cd_canonizer *h) {
    return h ? h->err.c_str() : "null canonizer";
}

int ise_cd_group_order(const ise_cd_canonizer *h, uint64_t *day_order, uint64_t *fixed_order) {
    if (!h) return ISE_ERR_ARG;
    if (day_order) *day_order = h->last_day_order;
    if (fixed_order) *fixed_order = h->last_fixed_order;
    return ISE_OK;
}

int ise_cd_run_canonizer_device(ise_cd_canonizer *h, uint32_t n, const uint8_t *d_matrices, int day,
                                uint64_t *d_canon, uint64_t *d_day_orders, uint64_t *d_fixed_orders) {
    if (!h) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!d_matrices || !d_canon) return fail(h, ISE_ERR_ARG, "runCanonizer: null buffer");
    uint32_t days = 0, b = 0;
    bool walk = false;
    int rc = resolve_days(h, day, &days, &b, &walk);
    if (rc) return rc;
    const uint32_t v = h->v, stride = v;
    const uint64_t P = walk ? factorial(days) : 1ull, total = (uint64_t)n * P;
    uint64_t part = 0;
    for (uint32_t j = 0; j < days; j++) part |= 1ull << (j * h->groups);
    CDCHK(h, hipSetDevice(h->device));
    // the batch: the whole run if it fits in CD_CHUNK, in powers of two
    uint32_t want = CD_CHUNK_MIN;
    while (want < CD_CHUNK && want < total) want <<= 1;
    if (want > h->chunk && (rc = set_chunk(h, want)) != ISE_OK) return rc;
    hipStream_t s = ise::ctx_stream(h->ctx);
    CDCHK(h, grow(&h->d_have, &h->have_cap, (size_t)n));
    if (!d_day_orders || !d_fixed_orders) {
        CDCHK(h, grow(&h->d_orders, &h->orders_cap, 2 * (size_t)n));
        if (!d_day_orders) d_day_orders = h->d_orders;
        if (!d_fixed_orders) d_fixed_orders = h->d_orders + n;
    }
    CDCHK(h, hipMemsetAsync(h->d_have, 0, (size_t)n * 4, s));
    CDCHK(h, hipMemsetAsync(h->d_flags, 0, 64, s));
    std::vector<uint32_t> seg, mseg;
    for (uint64_t c0 = 0; c0 < total; c0 += h->chunk) {
        const uint32_t cnt = (uint32_t)std::min<uint64_t>(h->chunk, total - c0);
        hipLaunchKernelGGL(cd_expand_kernel, dim3((cnt + 3) / 4), dim3(256), 0, s, d_matrices, v, h->k, days,
                           walk ? 1u : 0u, P, c0, cnt, stride, h->d_rows, h->d_meta, h->d_flags);
        CDCHK(h, hipGetLastError());
        rc = ise_canonize_batch_device_part(h->ctx, cnt, stride, h->d_rows, h->d_meta, part, h->d_canon,
                                            h->d_status, h->d_aut, nullptr, ISE_CANONIZE_MAX_ITERS);
        if (rc) return fail(h, rc, std::string("runCanonizer: ") + ise_last_error(h->ctx));
        if (!walk) {
            const uint64_t items = (uint64_t)cnt * v;
            hipLaunchKernelGGL(cd_copy_kernel, dim3((uint32_t)((items + 255) / 256)), dim3(256), 0, s, h->d_canon,
                               h->d_status, h->d_aut, stride, v, c0, cnt, d_canon, d_day_orders, d_fixed_orders,
                               h->d_flags);
            CDCHK(h, hipGetLastError());
            continue;
        }
        // runs of <= CD_SEG candidates of one k-system, and every k-system's runs
        seg.clear();
        mseg.clear();
        const uint64_t m_first = c0 / P, m_last = (c0 + cnt - 1) / P;
        for (uint64_t m = m_first; m <= m_last; m++) {
            const uint64_t lo = std::max<uint64_t>(c0, m * P), hi = std::min<uint64_t>(c0 + cnt, (m + 1) * P);
            mseg.push_back((uint32_t)m);
            mseg.push_back((uint32_t)(seg.size() / SEG_WORDS));
            uint32_t runs = 0;
            for (uint64_t a = lo; a < hi; a += CD_SEG, runs++) {
                seg.push_back((uint32_t)(a - c0));
                seg.push_back((uint32_t)std::min<uint64_t>(CD_SEG, hi - a));
                seg.push_back(CD_NONE);
                seg.push_back(0u);
            }
            mseg.push_back(runs);
        }
        const uint32_t nseg = (uint32_t)(seg.size() / SEG_WORDS), nm = (uint32_t)(mseg.size() / 3);
        CDCHK(h, grow(&h->d_seg, &h->seg_cap, seg.size()));
        CDCHK(h, grow(&h->d_mseg, &h->mseg_cap, mseg.size()));
        CDCHK(h, hipMemcpyAsync(h->d_seg, seg.data(), seg.size() * 4, hipMemcpyHostToDevice, s));
        CDCHK(h, hipMemcpyAsync(h->d_mseg, mseg.data(), mseg.size() * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(cd_segment_best_kernel, dim3(nseg), dim3(256), 0, s, h->d_canon, h->d_status, stride, v,
                           h->d_seg, h->d_flags);
        CDCHK(h, hipGetLastError());
        hipLaunchKernelGGL(cd_merge_kernel, dim3((nm + 63) / 64), dim3(64), 0, s, h->d_canon, h->d_aut, stride, v,
                           h->d_seg, h->d_mseg, nm, d_canon, h->d_have, d_day_orders, d_fixed_orders);
        CDCHK(h, hipGetLastError());
        // the host run lists are rewritten next chunk: their copies must be done
        CDCHK(h, hipStreamSynchronize(s));
    }
    uint32_t flags[2] = {0, 0};
    CDCHK(h, hipMemcpyAsync(flags, h->d_flags, sizeof flags, hipMemcpyDeviceToHost, s));
    CDCHK(h, hipStreamSynchronize(s));
    if (flags[0]) return fail(h, ISE_ERR_ARG, "runCanonizer: a player number is not below v");
    if (flags[1]) return fail(h, ISE_ERR_STATE, "runCanonizer: a canonization did not reach its canonical form");
    return ISE_OK;
}

int ise_cd_run_canonizer_batch(ise_cd_canonizer *h, uint32_t n, const uint8_t *matrices, int day, uint8_t *out,
                               uint64_t *day_orders, uint64_t *fixed_orders) {
    if (!h) return ISE_ERR_ARG;
    if (n == 0) return ISE_OK;
    if (!matrices || !out) return fail(h, ISE_ERR_ARG, "runCanonizer: null buffer");
    uint32_t days = 0, b = 0;
    bool walk = false;
    int rc = resolve_days(h, day, &days, &b, &walk);
    if (rc) return rc;
    const uint32_t v = h->v;
    const size_t in_bytes = (size_t)n * days * v;
    for (size_t i = 0; i < in_bytes; i++)
        if (matrices[i] >= v) return fail(h, ISE_ERR_ARG, "runCanonizer: a player number is not below v");
    CDCHK(h, hipSetDevice(h->device));
    hipStream_t s = ise::ctx_stream(h->ctx);
    CDCHK(h, grow(&h->d_mats, &h->mats_cap, in_bytes));
    CDCHK(h, grow(&h->d_best, &h->best_cap, (size_t)n * v));
    CDCHK(h, grow(&h->d_orders, &h->orders_cap, 2 * (size_t)n));
    CDCHK(h, hipMemcpyAsync(h->d_mats, matrices, in_bytes, hipMemcpyHostToDevice, s));
    rc = ise_cd_run_canonizer_device(h, n, h->d_mats, day, h->d_best, h->d_orders, h->d_orders + n);
    if (rc) return rc;
    s = ise::ctx_stream(h->ctx);   // a larger batch brings a new context (and stream)
    h->packed.resize((size_t)n * v);
    CDCHK(h, hipMemcpyAsync(h->packed.data(), h->d_best, h->packed.size() * 8, hipMemcpyDeviceToHost, s));
    if (day_orders) CDCHK(h, hipMemcpyAsync(day_orders, h->d_orders, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    if (fixed_orders) CDCHK(h, hipMemcpyAsync(fixed_orders, h->d_orders + n, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    CDCHK(h, hipStreamSynchronize(s));
    for (uint32_t m = 0; m < n; m++)
        for (uint32_t i = 0; i < v; i++) {
            const uint64_t r = h->packed[(size_t)m * v + i];
            uint8_t *o = out + ((size_t)m * v + i) * b;
            for (uint32_t q = 0; q < b; q++) o[q] = (uint8_t)((r >> q) & 1ull);
        }
    return ISE_OK;
}

int ise_cd_run_canonizer(ise_cd_canonizer *h, const uint8_t *matrix, uint32_t k, int day, const uint8_t **out,
                         uint32_t *ncols) {
    if (!h) return ISE_ERR_ARG;
    if (!matrix || !out) return fail(h, ISE_ERR_ARG, "runCanonizer: null buffer");
    if (k != h->k) return fail(h, ISE_ERR_ARG, "runCanonizer: group size differs from the canonizer's");
    uint32_t days = 0, b = 0;
    bool walk = false;
    int rc = resolve_days(h, day, &days, &b, &walk);
    if (rc) return rc;
    h->out.resize((size_t)h->v * b);
    uint64_t dord = 1, ford = 0;
    rc = ise_cd_run_canonizer_batch(h, 1, matrix, day, h->out.data(), &dord, &ford);
    if (rc) return rc;
    h->last_day_order = dord;
    h->last_fixed_order = ford;
    *out = h->out.data();
    if (ncols) *ncols = b;
    return ISE_OK;
}

} // extern "C"
