// gx_batch.h -- device structures of the multi-instance path (batching and
// serving) and the launches of one round.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gx_common.h"

namespace gx {

// One instance (model exploration) of a batch or a server slot.  A plain
// instance is expanded in chunks of 64 states (one state per lane); a POR
// instance (mode kModePor or kModePorProviso) in chunks of G = floor(64/N)
// states, one process-group layout pass (gx_group.h) per chunk.  The two
// kinds have their own chunk prefixes, so each expand kernel walks only its
// own chunks.
struct BatchInst {
  ModelHdr hdr;
  uint32_t model_off;   // into the concatenated model buffer (words)
  int32_t property;
  uint32_t mode;        // ExpandMode: plain, POR, POR with the cycle proviso
  uint32_t por_method;  // GX_POR_* under POR
  uint64_t tb, nb;      // visited-set partition: first bucket, buckets
  uint64_t fo, fcap;    // frontier region: first entry, capacity
  uint64_t front_count[2];
  uint64_t inserts, transitions, expanded;
  uint64_t chunk_prefix;      // first plain chunk of this instance in the round
  uint64_t chunk_prefix_por;  // first POR chunk of this instance in the round
  uint32_t cur, level;
  uint32_t violation, table_full, overflow, bad_state, spin_timeout, stop;
  uint32_t cand_overflow;  // a 64-state chunk of this instance emits more successors than the
                           // whole candidate buffer holds: the instance stops (more visited-set
                           // pages cannot help it)
  uint32_t pad;
};

struct BatchCtrl {
  uint64_t cand_count;
  uint64_t lookups;
  uint64_t total_chunks;      // plain chunks of the round
  uint64_t total_chunks_por;  // POR chunks of the round
  uint64_t deferred_count[2];
  uint32_t deferred_cur;
  uint32_t use_deferred;
  uint32_t active;       // instances not stopped
  uint32_t rounds;
  uint32_t n;
  uint32_t cand_overflow;  // some instance stopped with a chunk larger than the candidate buffer
  uint32_t prefix_only;  // next advance only recomputes the chunk prefix
  uint32_t pad;
  uint64_t cand_entries;  // candidate-buffer entries of all rounds (candidates + released gaps)
};

struct BatchArgs {
  BatchInst *inst;
  BatchCtrl *ctrl;
  const uint32_t *models;
  uint64_t *front[2];   // pools, W words per entry
  uint64_t *cand;
  uint32_t *cand_blk;   // instance of each 64-entry block of the candidate buffer
  uint64_t cand_cap;
  uint64_t *deferred[2];  // chunk descriptors: instance << 40 | first state
  uint64_t *table;
  uint64_t seed;
  uint32_t bucket_words;  // 64-bit words per bucket
  uint32_t max_procs;
  uint32_t rows16;   // every model has fewer than 0xFFFF rows: 16-bit row indices in LDS
  uint32_t por;      // POR instances may run: the POR expand kernel is part of the round
  uint32_t stamps;   // instances with the cycle proviso may run: slots carry level stamps
  uint32_t por_blocks, por_threads;  // launch of the POR expand kernel
  uint32_t chunk_states;  // states per plain workgroup chunk: 64 x the plain expand kernel's waves
  uint32_t expand_lds;    // dynamic LDS bytes of the plain expand kernel (batch_expand_lds)
};

// Mode of an instance from its job options (gx_job_options).
inline uint32_t batch_mode(int por, int cycle_proviso) {
  return por ? (cycle_proviso ? kModePorProviso : kModePor) : kModePlain;
}

// expand (plain, and POR when a.por) + FIND-OR-PUT + advance of every
// unfinished instance (one BFS level)
// Events recorded before the successor launches, between them and the
// FIND-OR-PUT launch, and after it (kernel timing of a batch round).
struct RoundEvents {
  hipEvent_t e0, e1, e2;
};

hipError_t launch_round_w(const BatchArgs &a, int W, int BW, uint32_t blocks, uint32_t threads,
                          uint32_t ins_blocks, hipStream_t s, const RoundEvents *ev = nullptr);
// LDS of the plain batch expand kernel: 40 KiB (four workgroups per CU, as
// the single-model kernel gets on a 20-process model), or the waves' row
// tables for the largest model when they need more.  A model is staged when
// its section and its own row tables fit.
constexpr size_t kBatchExpandLds = 40 * 1024;
inline size_t batch_rows_lds(uint32_t threads, uint32_t max_procs, bool rows16) {
  return static_cast<size_t>(threads / 64) * max_procs * 64 * (rows16 ? 2 : 4);
}
inline size_t batch_expand_lds(uint32_t threads, uint32_t max_procs, bool rows16) {
  const size_t r = batch_rows_lds(threads, max_procs, rows16);
  return r > kBatchExpandLds ? r : kBatchExpandLds;
}
// the advance step alone (with prefix_only set: chunk prefix after admissions)
hipError_t launch_batch_advance(const BatchArgs &a, hipStream_t s);
// the initial state into a zeroed partition (first bucket tb, nb buckets)
// and as entry fo of frontier pool front[0]
hipError_t launch_place_initial(const BatchArgs &a, int W, int BW, uint64_t tb, uint64_t nb, uint64_t fo,
                                hipStream_t s);
// the same for instances [0, n) of a.inst at once
hipError_t launch_place_initial_all(const BatchArgs &a, int W, uint32_t n, hipStream_t s);

}  // namespace gx
