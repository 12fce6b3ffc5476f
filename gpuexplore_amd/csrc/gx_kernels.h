// gx_kernels.h -- kernel argument blocks and launch helpers shared by the
// single-model driver and the table API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gx_common.h"
#include "gx_device.h"

namespace gx {

struct ExpandArgs {
  const uint32_t *model;  // compiled model (global copy)
  ModelHdr hdr;
  Ctrl *ctrl;
  const uint64_t *front[2];
  uint64_t *cand;
  uint64_t cand_cap;
  uint64_t *retry[2];
  int property;
  int bucket_words;  // FIND under the cycle proviso
  Table table;
  int per_state;     // plain mode: one state per lane (k_expand_tps)
  int por_method;    // GX_POR_* under POR (cample, ample, stubborn)
  int packed;        // successors as 12-byte packed candidates (see pack12)
};

// 12-byte candidate of a 2-word key with at most 95 state bits: payload bits
// [0, 95) in three 32-bit words, bit 31 of the third set (a zero third word
// is a released entry).  Word j of a key holds payload bits [63j, 63j+63)
// and the flag bit 63.  Successor generation writes each successor in its
// state's range of the wave's reservation, so a 128-byte line fills over a
// batch's emission steps and the L2 merges the pieces; with 12 bytes instead
// of 16 per successor the open lines of the waves in flight fit the L2
// better: on the 20-process ring successor generation took 31.6-32.0 ms
// instead of 36.5-37.0 ms, and FIND-OR-PUT, which reads the candidates
// once, 1.5-3 % less.
constexpr uint32_t kPackedMaxBits = 95;
__device__ __forceinline__ void pack12(const uint64_t (&k)[2], uint32_t *dst) {
  const uint32_t u0 = static_cast<uint32_t>(k[0]);
  const uint32_t u1 = (static_cast<uint32_t>(k[0] >> 32) & 0x7FFFFFFFu) | (static_cast<uint32_t>(k[1]) << 31);
  const uint32_t u2 = (static_cast<uint32_t>(k[1] >> 1) & 0x7FFFFFFFu) | 0x80000000u;
  dst[0] = u0;
  dst[1] = u1;
  dst[2] = u2;
}
__device__ __forceinline__ void unpack12(uint32_t u0, uint32_t u1, uint32_t u2, uint64_t (&k)[2]) {
  k[0] = static_cast<uint64_t>(u0) | (static_cast<uint64_t>(u1 & 0x7FFFFFFFu) << 32) | kFlag;
  k[1] = static_cast<uint64_t>(u1 >> 31) | (static_cast<uint64_t>(u2 & 0x7FFFFFFFu) << 1) | kFlag;
}

// LDS bytes k_expand_tps needs besides the staged model: per wave a
// [process][lane] table of row indices, 2 bytes each when the model has fewer
// than 0xFFFF rows.
inline size_t expand_tps_scratch_bytes(uint32_t threads, uint32_t nprocs, bool rows16) {
  return static_cast<size_t>(threads / 64) * nprocs * 64 * (rows16 ? 2 : 4) + 16;
}

struct InsertArgs {
  Ctrl *ctrl;
  const uint64_t *cand;
  uint64_t cand_cap;
  uint64_t *front[2];
  uint64_t front_cap;
  Table table;
  uint32_t state_bits;
  uint8_t *is_new;  // optional: 1 per candidate that was inserted
  int append;       // append new keys to front[cur ^ 1]
  int set_flag;     // keys come from a caller: set bit 63 of every word (a caller's
                    // zero word would otherwise read as a released or unpublished one)
  int dedup;        // skip candidates equal to an earlier one of the same or the previous
                    // 64-candidate step of the wave (needs append and no is_new)
  int packed;       // candidates are 12-byte packed (pack12); 2-word keys only
};

// compile-time options of the FIND-OR-PUT kernel (InsertArgs.dedup, .set_flag)
constexpr int kInsDedup = 1;
constexpr int kInsSetFlag = 2;
constexpr int kInsPacked = 4;

// LDS tag table of the FIND-OR-PUT duplicate filter: per wave, kDedupSlots
// words of (step << 6 | lane), indexed by the key hash's low bits
constexpr uint32_t kDedupSlots = 128;

// Resets a control block for a FIND-OR-PUT pass of n candidates with a
// kernel (no host copy, no memset): what a caller that cannot allocate or
// copy from pageable memory needs.
hipError_t launch_ctrl_reset(Ctrl *c, uint64_t n, hipStream_t s);

// FIND-OR-PUT launch geometry: 256-thread workgroups, 64 per CU by default
constexpr uint32_t kInsertThreads = 256;
constexpr uint32_t kInsertBlocksPerCU = 64;

// The FIND-OR-PUT lookups are latency bound, so resident waves matter:
// eight per SIMD are asked for where slots are at most two words and carry
// no level stamp (the hashed steps otherwise take about 106 SGPRs and seven
// waves; asking costs a few SGPR spills to VGPR lanes and runs 4 % faster).
// Wider or stamped slots would spill VGPRs to scratch, so they ask nothing.
template <int WT, bool STAMP>
constexpr int kInsertMinWaves = (WT <= 2 && !STAMP) ? 8 : 1;

// Candidate entries a wave reserves per atomic (at most; see CandChunk).
// All waves of the device share one counter, so every reservation is a
// same-address atomic.  With one per 64-state batch (a ring of 20 16-state
// processes emits about 1,300 successors per batch) the counter serialised
// successor generation: 65 ms per exploration of that ring, against 44 ms
// at 16,384 (4,096 and 65,536 measured alike).
constexpr uint64_t kChunk = 16384;

// Iterations in a row that one FIND-OR-PUT wave takes from the candidate
// buffer (`stride` = candidates of one iteration of all waves).  A pass that
// gives every wave fewer than kInsertSpread iterations is latency bound and
// spreads over the whole grid; a larger one goes in runs of kInsertRun, even
// if that leaves waves idle: fewer streams keep a run's recent buckets in
// the caches until its duplicate lookups arrive (a ring of 20 16-state
// processes, 295 M states, and 64 models of 4-14 processes measured best at
// 8 / 64; runs of 64 to 256 are alike).
constexpr uint64_t kInsertRun = 64;
constexpr uint64_t kInsertSpread = 8;
__device__ __forceinline__ uint64_t insert_run(uint64_t n, uint64_t stride) {
  const uint64_t iters = (n + stride - 1) / stride;
  return iters < 1 ? 1 : (iters < kInsertSpread ? iters : kInsertRun);
}

// A wave's current chunk of the candidate buffer.  Space is reserved with one
// atomicAdd per chunk, sized for the wave's remaining batches (at most kChunk
// entries unless one batch needs more); unused chunk tails are released by
// zeroing their first words (a zero first word is never a key: keys carry
// bit 63), or the third words of 12-byte entries (pack12), which the
// FIND-OR-PUT kernel skips.  All members are wave-uniform.
struct CandChunk {
  uint64_t *cand;
  uint64_t cap;
  uint64_t *count;
  uint64_t base = 0, left = 0;
  bool full = false;
  bool packed = false;  // 12-byte entries: the third word marks them
  __device__ CandChunk(uint64_t *c, uint64_t cp, uint64_t *cnt) : cand(c), cap(cp), count(cnt) {}
  template <int W>
  __device__ void release(uint64_t b, uint64_t n) {
    const uint64_t end = (b + n < cap) ? b + n : cap;
    if (packed) {
      uint32_t *c32 = reinterpret_cast<uint32_t *>(cand);
      for (uint64_t i = b + lane_id(); i < end; i += 64) c32[3 * i + 2] = 0;
      return;
    }
    for (uint64_t i = b + lane_id(); i < end; i += 64) cand[i * W] = 0;
  }
  template <int W>
  __device__ void release_all() {
    release<W>(base, left);
    left = 0;
  }
  // Reserves `total` entries for this batch (`batches` batches left for the
  // wave).  Returns 1 and *at on success, 0 if the buffer is full (defer the
  // batch), -1 if one batch can never fit.
  template <int W>
  __device__ int reserve(uint64_t total, uint64_t batches, uint64_t *at) {
    if (!total) return 1;
    if (full) return 0;
    if (total > left) {
      release<W>(base, left);
      uint64_t want = total * batches;
      want = want < kChunk ? want : kChunk;
      want = want > total ? want : total;
      want = (want + 63) & ~63ull;  // whole 64-entry blocks: chunks start block-aligned
      uint64_t got = 0;
      if (lane_id() == 0)
        got = atomicAdd(reinterpret_cast<unsigned long long *>(count), static_cast<unsigned long long>(want));
      got = __shfl(got, 0, 64);
      const uint64_t avail = got < cap ? cap - got : 0;
      if (avail >= total) {
        base = got;
        left = want < avail ? want : avail;
      } else {
        release<W>(got, want);
        base = 0;
        left = 0;
        full = true;
        return total > cap ? -1 : 0;
      }
    }
    *at = base;
    base += total;
    left -= total;
    return 1;
  }
};

// Launch one successor-generation pass (dispatches on key width, mode and
// whether the compiled model fits in LDS).  Returns a hipError_t.
hipError_t launch_expand(const ExpandArgs &a, int mode, int key_words, int slot_words, bool lds,
                         uint32_t blocks, uint32_t threads, hipStream_t s);

// Launch one FIND-OR-PUT pass over the candidate buffer.
hipError_t launch_insert(const InsertArgs &a, int key_words, int slot_words, int bucket_words,
                         bool stamp, uint32_t blocks, uint32_t threads, hipStream_t s);

hipError_t launch_advance(Ctrl *c, hipStream_t s);
// folds the partial counters into Ctrl's counters (k_advance also does)
hipError_t launch_fold(Ctrl *c, hipStream_t s);

hipError_t launch_count(const Table &t, int slot_words, int bucket_words, uint64_t *dev_count,
                        uint32_t blocks, hipStream_t s);

// Streaming device-to-device copy of n 16-byte elements (nt: non-temporal
// hints), streaming read of n 16-byte elements, and n random 64-byte bucket
// reads from a table of nbuckets 64-byte buckets (measurement kernels).
hipError_t launch_copy(const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks, bool nt, hipStream_t s);
hipError_t launch_read(const uint4 *src, uint64_t n, uint32_t *sink, uint32_t blocks, hipStream_t s);
hipError_t launch_random_read(const uint64_t *table, uint64_t nbuckets, uint64_t n, uint64_t seed, uint32_t *sink,
                              uint32_t blocks, hipStream_t s);

// LDS bytes the expand kernel needs besides the staged model.
size_t expand_scratch_bytes(uint32_t threads);

}  // namespace gx
