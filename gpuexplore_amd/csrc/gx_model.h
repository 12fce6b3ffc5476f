// gx_model.h -- gpf network-of-LTSs model: loader and compiler.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "gx_common.h"

namespace gx {

// The gpf fields in parse order (cu:1885-1983).
struct Gpf {
  uint32_t liveness = 0;
  uint32_t nr_procs = 0;
  uint32_t bits_act = 0;
  uint32_t proc_nrstates = 0;
  uint32_t bits_statevector = 0;  // +1 under POR (cu:1912)
  std::vector<uint32_t> firstbit;   // N+1
  std::vector<uint32_t> bits_state; // N
  uint32_t nbits_offset = 0;
  uint32_t max_buf_ints = 0;
  std::vector<uint32_t> proc_offsets_start;  // N+1
  std::vector<uint32_t> proc_offsets;
  std::vector<uint32_t> proc_trans;
  uint32_t nbits_syncbits_offset = 0;
  std::vector<uint32_t> syncbits_offsets;
  std::vector<uint32_t> syncbits;
  uint32_t sv_nints() const { return (bits_statevector + 31) / 32; }
};

// Parses gpf text.  Returns GX_OK or a GX_E_* status.
int parse_gpf(const std::string &text, bool por, Gpf *out);

// Field location of process p inside the 63-bit-chunk key.
struct FieldLoc {
  uint32_t chunk, offset, width;
};

struct Compiled {
  ModelHdr hdr{};
  std::vector<uint32_t> words;  // the buffer described by hdr
  std::vector<FieldLoc> floc;
};

// Decodes the packed gpf tables into flat per-(process, state) arrays.
int compile_model(const Gpf &g, Compiled *out);

// Host-side key codec (the device mirrors it in gx_device.h).
void key_from_words32(const Compiled &c, const uint32_t *vec32, uint64_t *key);
void key_to_words32(const Compiled &c, const uint64_t *key, uint32_t *vec32, uint32_t nwords32);
uint32_t key_get(const Compiled &c, const uint64_t *key, uint32_t p);
void key_set(const Compiled &c, uint64_t *key, uint32_t p, uint32_t v);

}  // namespace gx

struct gx_model {
  gx::Gpf gpf;
  gx::Compiled comp;
  bool por = false;
};
