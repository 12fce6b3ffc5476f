// gx_common.h -- structures shared by the host driver and the HIP kernels.
//
// Device state representation ("key"): a state vector of K = firstbit[N]
// bits is stored as W = ceil(K/63) 64-bit words; word j carries bits
// [63j, 63j+63) of the little-endian state bit string in bits 0..62 and has
// bit 63 set.  A visited-set slot is the same W words (plus, with the cycle
// proviso, a 32-bit level stamp appended after bit K), so a key is CASed into
// a slot as is, and an all-zero word marks an empty slot.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace gx {

constexpr int kMaxProcs = 32;  // rules are N-bit masks packed in 32-bit ints (cu:205)
constexpr int kMaxKeyWords = 5;  // 9 x 32 bits (cu:92) + 32-bit stamp fit in 5 x 63
constexpr uint64_t kFlag = 1ull << 63;
constexpr uint64_t kPayload = ~kFlag;
// Buckets a key may probe before the table is full.  The reference tries 8
// (NR_HASH_FUNCTIONS, cu:88) with buckets of 10 three-word slots; a 64-byte
// bucket here holds 4 two-word slots, and with 8 probes a table reported
// full at 37 % load (a 3.7e8-state model, 1e9 slots: a few keys in 1e6 found
// 8 full buckets), so 16.  Extra probes are taken only by keys whose earlier
// buckets are full.
constexpr int kHashFunctions = 16;

// Compiled model: offsets (in 32-bit words) of the arrays inside one
// contiguous buffer that a workgroup stages into LDS.
struct ModelHdr {
  uint32_t nprocs;      // N
  uint32_t key_words;   // W
  uint32_t state_bits;  // K
  uint32_t nlab;        // labels with a rule table: 0 .. nlab-1
  uint32_t total_words; // size of the buffer
  uint32_t nrows;       // sum of rows over processes
  // per process
  uint32_t off_floc;    // field location: chunk | offset << 8 | width << 16
  uint32_t off_rowbase; // first row of process p
  uint32_t off_rowcnt;  // rows of process p (states with offsets)
  // per row (p, s)
  uint32_t off_loc_start;  // nrows + 1, CSR into loc_tgt
  uint32_t off_syn_start;  // nrows + 1, CSR into syn_* (distinct labels, ascending)
  uint32_t off_rowflags;   // nrows: bit0 = has a local entry, bit1 = bad offsets
  uint32_t off_loc_tgt;    // decoded local target state ids
  // per (row, label)
  uint32_t off_syn_lab;
  uint32_t off_syn_tstart; // into syn_tgt
  uint32_t off_syn_tcnt;
  uint32_t off_syn_tgt;    // concatenated targets T_p(a), decoded ids
  // per label
  uint32_t off_rule_start; // nlab + 1, CSR into rules
  uint32_t off_rules;      // N-bit process masks
  uint32_t off_rowlab;     // nrows x 2 words: mask of the synchronising labels of the row (nlab <= 64)
  uint32_t lab64;          // 1 when every label fits the 64-bit row masks
  // packed copies for the state-per-lane kernels (16-byte aligned offsets)
  uint32_t off_rowrec;     // (nrows + 1) x 8 words: loc_start, loc_end, syn_start, flags, rowlab lo, hi,
                           // and four 16-bit single-target delta indices (the model's first four
                           // synchronising labels; kNoSyn16 when none or several)
  uint32_t off_synrec;     // per (row, label) entry x 2 words: tstart, tcnt
  uint32_t off_fdesc;      // per process x 4 words: key word | shift << 8, value mask, first
                           // row, rows (row records have one more, a bad one, at nrows)
  uint32_t off_locx;       // per loc_tgt entry: key XOR this is the successor (the row's
                           // state and the target, both placed in the process's field, XORed;
                           // key_words 64-bit words as lo, hi pairs)
  uint32_t off_synx;       // per syn_tgt entry: the same delta for synchronising targets
  uint32_t synlab_lo, synlab_hi;  // labels < 64 with a synchronising entry in some row
  uint32_t tps_begin;      // first word of the section the state-per-lane kernels read
                           // (rowbase .. synx, 4-word aligned, up to total_words)
};


// kRowMultiSync: a synchronising entry of the row has a target count other than 1
enum RowFlags : uint32_t { kRowHasLocal = 1u, kRowBad = 2u, kRowMultiSync = 4u };
// In the packed row records' flag word only: bit kRowLabShift + k is set
// when the row has an entry for the model's k-th synchronising label (k < 4,
// labels in ascending order).
constexpr uint32_t kRowLabShift = 8;
// no single-target synchronising delta for that label in the row record
constexpr uint32_t kNoSyn16 = 0xFFFFu;

// Exploration modes of the successor kernel.
enum ExpandMode : int { kModePlain = 0, kModePor = 1, kModePorProviso = 2 };

// Device-resident control block of one exploration.  Kernels read the work
// sizes from here, so BFS levels can be enqueued (or graph-replayed) without
// a host round trip.
struct Ctrl {
  uint64_t front_count[2];  // frontier sizes; [cur] is being expanded
  uint64_t cand_count;      // successors in the candidate buffer
  uint64_t retry_count[2];  // frontier indices deferred when the candidate buffer filled
  uint64_t lookups;         // FIND-OR-PUT operations issued
  uint64_t inserts;         // new states
  uint64_t transitions;     // successors generated
  uint64_t expanded;        // states expanded
  uint32_t cur;             // current frontier buffer
  uint32_t retry_cur;       // retry list being consumed when use_retry
  uint32_t use_retry;       // expand the retry list instead of the frontier
  uint32_t level;           // BFS level being expanded
  uint32_t violation;       // property violated at this level
  uint32_t table_full;      // a key found no free slot in its kHashFunctions buckets
  uint32_t overflow;        // frontier overflow (fatal)
  uint32_t bad_state;       // a reached local state has no offsets (fatal)
  uint32_t spin_timeout;    // a publish wait did not finish (fatal)
  uint32_t stop;            // no further level runs
  uint32_t done;            // frontier empty
  uint32_t levels_done;     // levels completed
  uint32_t truncated;       // stopped by max_states
  uint64_t max_states;      // stop after the level that reaches this many states (0: no bound)
  // Partial counters, folded into inserts, lookups, transitions and expanded
  // by k_advance (or k_fold): a workgroup adds into slot blockIdx % 64, the
  // slots 128 B apart, instead of every wave of the device hitting one
  // address at its end.  Kept last: the host copies the fields before it.
  uint64_t part[64 * 16];  // kCounterSlots x 16
};
constexpr uint32_t kCounterSlots = 64;
// bytes of Ctrl before the partial counters: what the host reads back
constexpr size_t kCtrlHead = offsetof(Ctrl, part);
enum CounterIdx : uint32_t { kCntInserts = 0, kCntLookups = 1, kCntTransitions = 2, kCntExpanded = 3 };

}  // namespace gx
