// ise_solver.hip -- the row-extension step of the enumerator on the GPU, and
// a device-resident level-by-level enumeration built on it.
//
// The next row of a canonical partial matrix is a choice, for every column
// orbit of that row (orbits induced by the rows above), of how many of its
// columns get a one (ones left-justified inside the orbit, as
// CMatrixCanonChecker::MakeRow writes them, Source/MatrixCanonChecker.cpp:
// 14-93).  The counts must sum to r, give every earlier row an intersection
// of exactly lambda, and leave every column able to reach k ones by the last
// row.  The reference solves this system with CInSysSolver / CRowSolution
// (Source/InSysSolver.h:92-286, Source/RowSolution.h:216-323); here one GPU
// wave per parent runs the same depth-first search over the orbits, with
// the counts tried from the largest down, so the solutions come out in the
// order of the host solver (ise_enum.cpp) and of the oracle.
//
// Two passes: one counts the solutions of every parent, an exclusive scan
// turns the counts into output offsets, the second writes the children in
// parent order.  A level of the enumeration is then: solve (two kernels +
// scan) -> canonicity kernel -> order-preserving compaction of the canonical
// children (flag kernel + scan + scatter) -> next level.  The host only reads
// the level sizes.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "ise_internal.h"

namespace ise {

struct SolverDesign {
    int v, b, r, k, lambda;
};

__device__ __forceinline__ uint64_t low_bits_d(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

// Solutions of the row system of one parent (rows[0..depth-1]), one wave
// per parent: lane i < depth holds earlier row i and its intersection so far
// (lam), lane c < b is column c, lane o < norb holds orbit o's first column,
// end, count range and current count.
// The lambda conditions of every earlier row are checked at once (a ballot);
// the ones row i can still collect after orbit o are the ones of row i in
// the columns after it that still need a one -- a popcount.  Solutions come
// out in the thread version's order.  Writes the children of the parent
// (parent rows, then the solution row, stride words each) when `children`
// is non-null and returns the number of solutions (every lane).
__device__ uint32_t solve_parent_wave(const uint64_t *rows, int depth, uint32_t stride, const SolverDesign ds,
                                      uint64_t *children) {
    const int l = (int)__lane_id();
    const uint64_t bmask = low_bits_d(ds.b);
    const uint64_t prow = l < depth ? rows[l] : 0ull;          // lane i: row i
    // column orbits: column c starts one when an earlier row changes value at c
    uint64_t starts = 1, w_col = 0;
    for (int i = 0; i < depth; i++) {
        const uint64_t ri = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(prow >> 32), i) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)prow, i);
        starts |= ri ^ (ri << 1);
        w_col += (ri >> l) & 1ull;                            // lane c: ones of column c so far
    }
    starts &= bmask;
    const int norb = __popcll(starts);
    const int after = ds.v - depth - 1;                       // rows still to come after this one
    const int need = ds.k - (int)w_col;
    // a column that can no longer reach k ones (or must take more than the rows left)
    if (__builtin_amdgcn_ballot_w64(l < ds.b && (need < 0 || need > after + 1)))
        return 0;
    // columns that may take a one in this row, and those that must
    const uint64_t HI = __builtin_amdgcn_ballot_w64(l < ds.b && need > 0);
    const uint64_t LO = __builtin_amdgcn_ballot_w64(l < ds.b && need > after);
    // lane o: orbit o
    int o_start = 0, o_end = 0, o_hi = 0, o_lo = 0;
    if (l < norb) {
        uint64_t m = starts;
        for (int t = 0; t < l; t++) m &= m - 1;
        o_start = __builtin_ctzll(m);
        const uint64_t rest = m & (m - 1);
        o_end = rest ? __builtin_ctzll(rest) : ds.b;
        o_hi = ((HI >> o_start) & 1ull) ? o_end - o_start : 0;
        o_lo = ((LO >> o_start) & 1ull) ? o_end - o_start : 0;
    }
    // lane c: its orbit index and first column (to write a solution row)
    const uint64_t le = l == 63 ? ~0ull : ((2ull << l) - 1);
    const int c_orb = __popcll(starts & le) - 1;
    const int c_start = 63 - __builtin_clzll(starts & le);
    int lam = 0;                                              // lane i < depth
    int cur = 0;                                              // lane o: count applied (hi + 1: none)
    uint32_t found = 0;
    int sum = 0, o = 0;
    if (l == 0) cur = __builtin_amdgcn_readlane(o_hi, 0) + 1;
    for (;;) {
        o = __builtin_amdgcn_readfirstlane(o);
        sum = __builtin_amdgcn_readfirstlane(sum);
        if (o == norb) {
            const bool all = sum == ds.r &&
                             !__builtin_amdgcn_ballot_w64(l < depth && lam != ds.lambda);
            if (all) {
                if (children) {
                    const int cc = __builtin_amdgcn_ds_bpermute((c_orb < 0 ? 0 : c_orb) << 2, cur);
                    const uint64_t row = __builtin_amdgcn_ballot_w64(l < ds.b && l - c_start < cc);
                    uint64_t *ch = children + (uint64_t)found * stride;
                    if (l < (int)stride) ch[l] = l < depth ? prow : (l == depth ? row : 0ull);
                }
                found++;
            }
            if (--o < 0) break;
            continue;
        }
        const int hi = __builtin_amdgcn_readlane(o_hi, o), lo = __builtin_amdgcn_readlane(o_lo, o);
        const int s0 = __builtin_amdgcn_readlane(o_start, o), e0 = __builtin_amdgcn_readlane(o_end, o);
        const int in_row = (int)((prow >> s0) & 1ull);       // lane i: orbit o inside row i
        int x = __builtin_amdgcn_readlane(cur, o);
        if (x <= hi) {                                        // undo the count applied at this orbit
            sum -= x;
            lam -= in_row * x;
        }
        const uint64_t later = bmask & ~low_bits_d(e0);       // columns after orbit o
        const int cap_r = __popcll(HI & later);
        const int cap_l = __popcll(prow & HI & later);
        for (x = x - 1; x >= lo; x--) {
            const int ns = sum + x;
            if (ns > ds.r || ns + cap_r < ds.r) continue;
            const int nl = lam + in_row * x;
            if (!__builtin_amdgcn_ballot_w64(l < depth && (nl > ds.lambda || nl + cap_l < ds.lambda)))
                break;
        }
        if (x < lo) {                                         // orbit exhausted: back to the previous one
            if (l == o) cur = hi + 1;
            if (--o < 0) break;
            continue;
        }
        if (l == o) cur = x;
        sum += x;
        lam += in_row * x;
        if (++o < norb && l == o) cur = __builtin_amdgcn_readlane(o_hi, o) + 1;
    }
    return found;
}

__global__ __launch_bounds__(256) void solve_count_kernel(uint32_t n, uint32_t stride,
                                                          const uint64_t *__restrict__ parents, int depth,
                                                          SolverDesign ds, uint32_t *__restrict__ counts) {
    const uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (p >= n) return;
    const uint32_t cnt = solve_parent_wave(parents + (uint64_t)p * stride, depth, stride, ds, nullptr);
    if (__lane_id() == 0) counts[p] = cnt;
}

// children[offsets[p] + t] = parent p with solution t as row `depth`
__global__ __launch_bounds__(256) void solve_write_kernel(uint32_t n, uint32_t stride,
                                                          const uint64_t *__restrict__ parents, int depth,
                                                          SolverDesign ds, const uint32_t *__restrict__ offsets,
                                                          uint64_t *__restrict__ children) {
    const uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (p >= n) return;
    solve_parent_wave(parents + (uint64_t)p * stride, depth, stride, ds,
                      children + (uint64_t)offsets[p] * stride);
}

__global__ void fill_meta_kernel(uint32_t n, uint32_t m, uint32_t *__restrict__ meta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) meta[i] = m;
}

__global__ void canon_flag_kernel(uint32_t n, const uint8_t *__restrict__ verdict, uint32_t *__restrict__ flag,
                                  uint32_t *__restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t v = verdict[i];
    flag[i] = v == ISE_VERDICT_CANONICAL ? 1u : 0u;
    if (v != ISE_VERDICT_CANONICAL && v != ISE_VERDICT_NONCANONICAL) atomicOr(bad, 1u);
}

__global__ void compact_kernel(uint32_t n, uint32_t stride, const uint64_t *__restrict__ rows,
                               const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                               const uint64_t *__restrict__ order_in, uint64_t *__restrict__ out,
                               uint64_t *__restrict__ order_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t d = pos[i];
    for (uint32_t w = 0; w < stride; w++) out[(uint64_t)d * stride + w] = rows[(uint64_t)i * stride + w];
    if (order_out) order_out[d] = order_in[i];
}

// ---- host side -------------------------------------------------------------

// Growable device buffer.
template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = n + n / 2 + 64;
        hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
        if (e == hipSuccess) cap = want;
        return e;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct Scanner {
    DevBuf<uint8_t> tmp;
    hipError_t exclusive(const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s) {
        size_t bytes = 0;
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s);
        if (e == hipSuccess) e = tmp.reserve(bytes);
        if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, s);
        return e;
    }
};

inline unsigned blocks256(size_t n) { return (unsigned)((n + 255) / 256); }
// one wave per parent, 4 per workgroup
inline unsigned wave_blocks(size_t n) { return (unsigned)((n + 3) / 4); }

} // namespace ise

using namespace ise;

extern "C" {

int ise_solve_rows_device(ise_ctx *ctx, int v, int k, int lambda, uint32_t n, uint32_t stride,
                          const uint64_t *d_parents, int depth, uint32_t *d_counts, uint32_t *d_offsets,
                          uint64_t *d_children, uint64_t children_cap, uint64_t *n_children) {
    if (!ctx || !n_children || v < 2 || k < 2 || lambda < 1 || depth < 0 || depth >= v || (int)stride <= depth ||
        stride > ISE_MAX_ROWS || (n && (!d_parents || !d_counts || !d_offsets)))
        return ISE_ERR_ARG;
    if ((lambda * (v - 1)) % (k - 1)) return ISE_ERR_ARG;
    const int r = lambda * (v - 1) / (k - 1);
    if ((v * r) % k) return ISE_ERR_ARG;
    const SolverDesign ds{v, v * r / k, r, k, lambda};
    if (ds.b > ISE_MAX_COLS || v > ISE_MAX_ROWS) return ISE_ERR_ARG;
    *n_children = 0;
    if (n == 0) return ISE_OK;
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ISE_ERR_HIP;
    hipStream_t s = ctx_stream(ctx);
    Scanner scan;
    hipLaunchKernelGGL(solve_count_kernel, dim3(wave_blocks(n)), dim3(256), 0, s, n, stride, d_parents, depth, ds,
                       d_counts);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = scan.exclusive(d_counts, d_offsets, n, s);
    uint32_t last[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&last[0], d_offsets + n - 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], d_counts + n - 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ISE_ERR_HIP;
    const uint64_t total = (uint64_t)last[0] + last[1];
    *n_children = total;
    if (total == 0) return ISE_OK;
    if (!d_children || total > children_cap) return ISE_ERR_CAPACITY;
    hipLaunchKernelGGL(solve_write_kernel, dim3(wave_blocks(n)), dim3(256), 0, s, n, stride, d_parents, depth, ds,
                       d_offsets, d_children);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? ISE_OK : ISE_ERR_HIP;
}

long ise_enumerate_bibd_device(ise_ctx *ctx, int v, int k, int lambda, uint64_t *designs, uint64_t *auts,
                               long max_designs, uint64_t *stats, uint64_t *level_counts) {
    if (!ctx || v < 2 || k < 2 || lambda < 1 || v > ISE_MAX_ROWS) return ISE_ERR_ARG;
    if ((lambda * (v - 1)) % (k - 1)) return ISE_ERR_ARG;
    const int r = lambda * (v - 1) / (k - 1);
    if ((v * r) % k) return ISE_ERR_ARG;
    const SolverDesign ds{v, v * r / k, r, k, lambda};
    if (ds.b > ISE_MAX_COLS) return ISE_ERR_ARG;
    const uint32_t stride = (uint32_t)v;   // rows of one candidate, packed
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ISE_ERR_HIP;
    hipStream_t s = ctx_stream(ctx);
    DevBuf<uint64_t> par, kids, order_all, order_kept;
    DevBuf<uint32_t> counts, offs, meta, flag, pos, head, bad;
    DevBuf<uint8_t> verdict;
    Scanner scan;
    uint64_t st[5] = {0, 0, 0, 0, 0};
    if (level_counts)
        for (int i = 0; i <= v; i++) level_counts[i] = 0;
    long ndesigns = 0;
    hipError_t e = par.reserve(stride);
    if (e == hipSuccess) e = hipMemsetAsync(par.p, 0, stride * sizeof(uint64_t), s);
    if (e == hipSuccess) e = head.reserve(ISE_QUEUE_WORDS);
    if (e == hipSuccess) e = bad.reserve(1);
    uint64_t npar = 1;
    std::vector<uint64_t> host_rows, host_order;
    auto fail = [&](int rc) -> long {
        if (rc == ISE_ERR_HIP) ctx_set_error(ctx, "device enumeration failed");
        return rc;
    };
    if (e != hipSuccess) return fail(ISE_ERR_HIP);
    // the check launch shape: the smallest lane group holding v rows and b columns
    ise_launch_cfg saved;
    ise_get_launch(ctx, &saved);
    ise_launch_cfg cfg = saved;
    cfg.group_lanes = (uint32_t)fit_group_lanes((uint32_t)v, (uint32_t)ds.b);
    ise_set_launch(ctx, &cfg);
    long result = 0;
    for (int depth = 0; depth < v && npar > 0; depth++) {
        // ---- children of every kept parent
        if ((e = counts.reserve(npar)) != hipSuccess || (e = offs.reserve(npar)) != hipSuccess) break;
        hipLaunchKernelGGL(solve_count_kernel, dim3(wave_blocks(npar)), dim3(256), 0, s, (uint32_t)npar, stride, par.p,
                           depth, ds, counts.p);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = scan.exclusive(counts.p, offs.p, (uint32_t)npar, s)) != hipSuccess) break;
        uint32_t last[2];
        if ((e = hipMemcpyAsync(&last[0], offs.p + npar - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(&last[1], counts.p + npar - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
        const uint64_t nkid = (uint64_t)last[0] + last[1];
        const int nd = depth + 1;
        if (level_counts) level_counts[nd] = nkid;
        if (nkid == 0) break;
        if (nkid > 0xffffffffull) { result = ISE_ERR_CAPACITY; break; }
        if ((e = kids.reserve(nkid * stride)) != hipSuccess) break;
        hipLaunchKernelGGL(solve_write_kernel, dim3(wave_blocks(npar)), dim3(256), 0, s, (uint32_t)npar, stride,
                           par.p, depth, ds, offs.p, kids.p);
        if ((e = hipGetLastError()) != hipSuccess) break;
        // ---- canonicity of every child
        const uint32_t out_info = nd == v ? (ISE_SAVE_ROW_PERMUTATIONS | ISE_SAVE_ROW_TO_CHANGE) : 0u;
        const uint32_t m = (uint32_t)nd | ((uint32_t)ds.b << 8) | (out_info << 24);
        if ((e = meta.reserve(nkid)) != hipSuccess || (e = verdict.reserve(nkid)) != hipSuccess ||
            (e = order_all.reserve(nkid)) != hipSuccess)
            break;
        hipLaunchKernelGGL(fill_meta_kernel, dim3(blocks256(nkid)), dim3(256), 0, s, (uint32_t)nkid, m, meta.p);
        if ((e = hipGetLastError()) != hipSuccess) break;
        ise_results dres{};
        dres.verdict = verdict.p;
        dres.group_order = order_all.p;
        int rc = ise_check_batch_device(ctx, (uint32_t)nkid, stride, kids.p, nullptr, meta.p, &dres, head.p, s);
        if (rc) { result = rc; break; }
        st[0] += nkid;
        if (nd == v) st[2] += nkid;
        // ---- order-preserving compaction of the canonical children
        if ((e = flag.reserve(nkid)) != hipSuccess || (e = pos.reserve(nkid)) != hipSuccess) break;
        if ((e = hipMemsetAsync(bad.p, 0, 4, s)) != hipSuccess) break;
        hipLaunchKernelGGL(canon_flag_kernel, dim3(blocks256(nkid)), dim3(256), 0, s, (uint32_t)nkid, verdict.p,
                           flag.p, bad.p);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = scan.exclusive(flag.p, pos.p, (uint32_t)nkid, s)) != hipSuccess) break;
        uint32_t tail[3];
        if ((e = hipMemcpyAsync(&tail[0], pos.p + nkid - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(&tail[1], flag.p + nkid - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipMemcpyAsync(&tail[2], bad.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
        if (tail[2]) { result = ISE_ERR_ARG; break; }   // search limit or invalid record: refuse to guess
        const uint64_t ncan = (uint64_t)tail[0] + tail[1];
        if (ncan == 0) { npar = 0; break; }
        if ((e = order_kept.reserve(ncan)) != hipSuccess) break;
        DevBuf<uint64_t> next;
        if ((e = next.reserve(ncan * stride)) != hipSuccess) break;
        hipLaunchKernelGGL(compact_kernel, dim3(blocks256(nkid)), dim3(256), 0, s, (uint32_t)nkid, stride, kids.p,
                           flag.p, pos.p, order_all.p, next.p, order_kept.p);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if (nd == v) {
            // full designs: copy the canonical ones out
            host_rows.resize(ncan * stride);
            host_order.resize(ncan);
            if ((e = hipMemcpyAsync(host_rows.data(), next.p, ncan * stride * 8, hipMemcpyDeviceToHost, s)) !=
                hipSuccess)
                break;
            if ((e = hipMemcpyAsync(host_order.data(), order_kept.p, ncan * 8, hipMemcpyDeviceToHost, s)) !=
                hipSuccess)
                break;
            if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
            for (uint64_t c = 0; c < ncan; c++, ndesigns++) {
                if (designs && ndesigns < max_designs) {
                    memcpy(designs + (size_t)ndesigns * v, &host_rows[c * stride], v * sizeof(uint64_t));
                    if (auts) auts[ndesigns] = host_order[c];
                }
            }
            st[3] += ncan;
            npar = 0;
        } else {
            st[1] += ncan;
            std::swap(par.p, next.p);
            std::swap(par.cap, next.cap);
            npar = ncan;
        }
    }
    ise_set_launch(ctx, &saved);
    if (e != hipSuccess) return fail(ISE_ERR_HIP);
    if (result < 0) return fail((int)result);
    if (stats) memcpy(stats, st, sizeof st);
    return ndesigns;
}

} // extern "C"
