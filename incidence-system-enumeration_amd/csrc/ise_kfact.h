// ise_kfact.h -- shared declarations of the 1-factorization canonicity path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ise.h"

namespace ise {

constexpr uint32_t KF_MAX_PLAYERS = 16;   // labels fit in 4 bits
constexpr uint32_t KF_MAX_ROUNDS = 15;
constexpr uint32_t KF_MAX_SLOTS = 8;      // pairs per round

// Byte selector of v_perm_b32 that yields 0x00.
constexpr uint32_t KF_PERM_ZERO = 0x0c;

// QuickPerm tables of the pair-slot arrangements for every slot count 2..8:
// entry h of table k is a uint64 whose byte i is the base label (2*slot) put
// in slot i by the h-th arrangement.  The inverse table (same offsets) holds
// in byte j the slot whose pair receives labels 2j, 2j+1 (KF_PERM_ZERO for
// j >= k).
struct KfTables {
    const uint64_t *base;
    const uint64_t *inv;
    uint32_t offset[KF_MAX_SLOTS + 1];
    uint32_t count[KF_MAX_SLOTS + 1];
};

struct KfOut {
    uint8_t *verdict;
    uint64_t *witness, *n_auts, *n_tried;
    uint16_t *player_index = nullptr;
};

// Split launch: after the first pass the relabellings of the candidates still
// undecided are taken in windows [W_p, W_p+1), W_0 = the first-pass limit and
// W_p+1 = KF_WINDOW_GROWTH * W_p, each window cut into ranges over many
// waves; a candidate decided in a window leaves the list.  The window that
// starts at KF_LAST_WINDOW or later runs to the end of the index space, and
// so does window KF_MAX_WINDOWS - 1 (from a limit of 64 the window from
// 2^20 is window 4).
constexpr uint32_t KF_WINDOW_GROWTH = 16;
constexpr uint32_t KF_MAX_WINDOWS = 6;
constexpr uint64_t KF_LAST_WINDOW = 1ull << 20;

// Queue words, each on its own 128-byte line: head[0] the candidate queue of
// a one-pass launch, head[KF_HEAD_WORD * (1 + p)] the length of the list of
// candidates entering window p.
constexpr uint32_t KF_HEAD_WORD = 32;
// head[KF_QUAD_WORD]: the number of candidates the first look lists for the
// first pass (in list[1], which window 0 refills only after the first pass).
constexpr uint32_t KF_QUAD_WORD = 16;
static_assert(KF_HEAD_WORD * (KF_MAX_WINDOWS + 1) < ISE_QUEUE_WORDS, "queue words of a split launch");

// Context buffers of a split launch.  Window p reads list[p & 1] / part[p & 1]
// (candidate index, automorphisms counted before the window) and appends the
// candidates it leaves undecided to the other pair; best holds per entry the
// least witness published in the current window; ranges holds range_words
// range results.
struct KfSplit {
    uint32_t *list[2];
    uint64_t *part[2];
    uint32_t *best;
    uint64_t *ranges;
    uint32_t range_words, max_ranges;
};

// Host side: build the tables (Fuchs' countdown QuickPerm order,
// k-Factorization/TripleSys/Source/cnv.cpp:283-315); kf_table_words() words,
// the arrangements in the first half, their inverses in the second.
size_t kf_table_words();
void kf_build_tables(uint64_t *words, uint32_t *offset, uint32_t *count);

// limit = 0: one pass, every candidate scanned to its decision by one wave.
// limit > 0: first pass up to `limit` relabellings, then the ranges and the
// combine (sp must hold at least ncand list entries and range_words >= ncand).
hipError_t kf_launch(uint32_t blocks, uint32_t range_blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                     const uint32_t *meta, const KfOut &o, uint32_t *head, const KfTables &qp,
                     uint32_t limit, const KfSplit &sp, hipStream_t stream);
// Two-row canonizer of perfect 1-factorizations (cnvCheck2P1F): one wave per
// candidate, candidates by grid stride.
hipError_t kf2_launch(uint32_t blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                      const uint32_t *meta, const KfOut &o, hipStream_t stream);
// Two-row canonizer of uniform 1-factorizations (cnvCheck2U1F, md.reps = 1)
// or cnvCheck3U1F for groups of 2 (md.reps = 32, the cycle types the run
// allows in md.types: (cycles << 16) | half lengths a nibble each, ascending;
// none: one cycle through every player): one wave per candidate, candidates by
// grid stride.
constexpr uint32_t KF_U1F_MAX_TYPES = 4;
struct KfU1fMode {
    uint32_t reps, ntypes;
    uint32_t types[KF_U1F_MAX_TYPES];
};
hipError_t kf_u1f_launch(uint32_t blocks, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                         const uint32_t *meta, const KfOut &o, const KfU1fMode &md, hipStream_t stream);
hipError_t kf_occupancy(int *blocks_per_cu);
// The k-Sys backjump position of every candidate (o.player_index) from the
// first witness the canonizer of `mode` (0 general, 1 two-row P1F, 2 two-row
// U1F) wrote to o.witness; one lane per candidate.
hipError_t kf_backjump_launch(int mode, uint32_t ncand, uint32_t stride, const uint64_t *rounds,
                              const uint32_t *meta, const KfOut &o, const KfTables &qp, hipStream_t stream);

} // namespace ise
