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
};

// LDS bytes k_expand_tps needs besides the staged model.
inline size_t expand_tps_scratch_bytes(uint32_t threads, uint32_t nprocs) {
  return static_cast<size_t>(threads / 64) * nprocs * 64 * 4;
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
};

// Launch one successor-generation pass (dispatches on key width, mode and
// whether the compiled model fits in LDS).  Returns a hipError_t.
hipError_t launch_expand(const ExpandArgs &a, int mode, int key_words, int slot_words, bool lds,
                         uint32_t blocks, uint32_t threads, hipStream_t s);

// Launch one FIND-OR-PUT pass over the candidate buffer.
hipError_t launch_insert(const InsertArgs &a, int key_words, int slot_words, int bucket_words,
                         bool stamp, uint32_t blocks, uint32_t threads, hipStream_t s);

hipError_t launch_advance(Ctrl *c, hipStream_t s);

hipError_t launch_count(const Table &t, int slot_words, int bucket_words, uint64_t *dev_count,
                        uint32_t blocks, hipStream_t s);

// Streaming device-to-device copy of n 16-byte elements.
hipError_t launch_copy(const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks, hipStream_t s);

// LDS bytes the expand kernel needs besides the staged model.
size_t expand_scratch_bytes(uint32_t threads);

}  // namespace gx
