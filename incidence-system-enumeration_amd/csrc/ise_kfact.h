// ise_kfact.h -- shared declarations of the 1-factorization canonicity path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ise.h"

namespace ise {

constexpr uint32_t KF_MAX_PLAYERS = 16;   // labels fit in 4 bits
constexpr uint32_t KF_MAX_ROUNDS = 15;
constexpr uint32_t KF_MAX_SLOTS = 8;      // pairs per round

// QuickPerm tables of the pair-slot arrangements for every slot count 2..8:
// entry h of table k is a uint64 whose byte i is the base label (2*slot) put
// in slot i by the h-th arrangement.
struct KfTables {
    const uint64_t *base;
    uint32_t offset[KF_MAX_SLOTS + 1];
    uint32_t count[KF_MAX_SLOTS + 1];
};

struct KfOut {
    uint8_t *verdict;
    uint64_t *witness, *n_auts, *n_tried;
};

// Queue words of a split launch, each on its own 128-byte line: head[0] the
// first pass's candidate queue, head[KF_HEAD_WORD] the number of candidates
// listed for the ranges, head[2 * KF_HEAD_WORD] the range queue.
constexpr uint32_t KF_HEAD_WORD = 32;

// Context buffers of a split launch: list, part and best hold one entry per
// listed candidate (candidate index, automorphisms of the first pass, least
// witness published so far); ranges holds range_words results.
struct KfSplit {
    uint32_t *list;
    uint64_t *part;
    uint32_t *best;
    uint64_t *ranges;
    uint32_t range_words, max_ranges;
};

// Host side: build the tables (Fuchs' countdown QuickPerm order,
// k-Factorization/TripleSys/Source/cnv.cpp:283-315).
size_t kf_table_words();
void kf_build_tables(uint64_t *words, uint32_t *offset, uint32_t *count);

// limit = 0: one pass, every candidate scanned to its decision by one wave.
// limit > 0: first pass up to `limit` relabellings, then the ranges and the
// combine (sp must hold at least ncand list entries and range_words >= ncand).
hipError_t kf_launch(uint32_t blocks, uint32_t range_blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                     const uint32_t *meta, const KfOut &o, uint32_t *head, const KfTables &qp,
                     uint32_t limit, const KfSplit &sp, hipStream_t stream);
hipError_t kf_occupancy(int *blocks_per_cu);

} // namespace ise
