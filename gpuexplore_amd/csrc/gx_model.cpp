// gx_model.cpp -- gpf loader and model compiler.
//
// The loader reads the fields in the reference's order (cu:1885-1983) with
// the reference's line semantics: records of at most 255 characters (fgets
// into a 256-byte buffer) converted like glibc atoi, i.e. a 64-bit strtol
// truncated to 32 bits, so values up to 2^32-1 round-trip.
//
// The compiler decodes the packed tables once on the host, following the
// reference kernel's bit arithmetic: transition offsets (cu:883-897),
// transition entries (cu:143, 157-159), synchronisation-rule offsets
// (cu:997-1009) and rule fields (cu:205, 1480-1487).  The device then reads
// flat arrays instead of re-decoding packed fields per state.
#include "gx_model.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>

#include "../../include/gpuexplore.h"

namespace gx {

namespace {

struct LineReader {
  const std::string &t;
  size_t pos = 0;
  explicit LineReader(const std::string &text) : t(text) {}
  // One fgets(buf, 256, fp) + atoi step; false at end of input.
  bool next(uint32_t *v) {
    if (pos >= t.size()) return false;
    size_t end = pos;
    while (end < t.size() && end - pos < 255) {
      if (t[end++] == '\n') break;
    }
    std::string rec = t.substr(pos, end - pos);
    pos = end;
    *v = static_cast<uint32_t>(strtol(rec.c_str(), nullptr, 10));
    return true;
  }
  bool array(size_t n, std::vector<uint32_t> *out) {
    out->resize(n);
    for (size_t i = 0; i < n; i++)
      if (!next(&(*out)[i])) return false;
    return true;
  }
};

inline uint32_t bits(uint32_t x, uint32_t start, uint32_t len) {
  if (len == 0 || start >= 32) return 0;
  uint64_t mask = (len >= 32) ? 0xFFFFFFFFull : ((1ull << len) - 1ull);
  return static_cast<uint32_t>((static_cast<uint64_t>(x) >> start) & mask);
}

}  // namespace

int parse_gpf(const std::string &text, bool por, Gpf *g) {
  LineReader r(text);
  uint32_t n = 0;
  if (!r.next(&g->liveness)) return GX_E_FORMAT;       // cu:1888
  if (!r.next(&g->nr_procs)) return GX_E_FORMAT;       // cu:1899
  if (!r.next(&g->bits_act)) return GX_E_FORMAT;       // cu:1903
  if (!r.next(&g->proc_nrstates)) return GX_E_FORMAT;  // cu:1907
  if (!r.next(&g->bits_statevector)) return GX_E_FORMAT;  // cu:1911
  g->bits_statevector += por ? 1u : 0u;                   // cu:1912
  n = g->nr_procs;
  if (n == 0 || n > static_cast<uint32_t>(kMaxProcs)) return GX_E_LIMIT;
  if (!r.array(n + 1, &g->firstbit)) return GX_E_FORMAT;   // cu:1915-1921
  if (!r.array(n, &g->bits_state)) return GX_E_FORMAT;     // cu:1924-1930
  if (!r.next(&g->nbits_offset)) return GX_E_FORMAT;       // cu:1931
  if (!r.next(&g->max_buf_ints)) return GX_E_FORMAT;       // cu:1935
  if (!r.array(n + 1, &g->proc_offsets_start)) return GX_E_FORMAT;  // cu:1939-1944
  if (!r.array(g->proc_offsets_start[n], &g->proc_offsets)) return GX_E_FORMAT;  // cu:1945-1950
  uint32_t nr_trans = 0;
  if (!r.next(&nr_trans)) return GX_E_FORMAT;               // cu:1951
  if (!r.array(nr_trans, &g->proc_trans)) return GX_E_FORMAT;  // cu:1955-1960
  if (!r.next(&g->nbits_syncbits_offset)) return GX_E_FORMAT;  // cu:1962
  uint32_t nso = 0, nsb = 0;
  if (!r.next(&nso)) return GX_E_FORMAT;                    // cu:1965
  if (!r.array(nso, &g->syncbits_offsets)) return GX_E_FORMAT;  // cu:1968-1973
  if (!r.next(&nsb)) return GX_E_FORMAT;                    // cu:1974
  if (!r.array(nsb, &g->syncbits)) return GX_E_FORMAT;      // cu:1977-1982
  return GX_OK;
}

int compile_model(const Gpf &g, Compiled *out) {
  const uint32_t N = g.nr_procs;
  if (N == 0 || N > static_cast<uint32_t>(kMaxProcs)) return GX_E_LIMIT;
  for (uint32_t p = 0; p < N; p++) {
    if (g.firstbit[p + 1] < g.firstbit[p] || g.firstbit[p + 1] - g.firstbit[p] > 32) return GX_E_RANGE;
    if (g.bits_state[p] == 0 || g.bits_state[p] > 31) return GX_E_RANGE;
  }
  if (g.firstbit[0] != 0) return GX_E_RANGE;
  if (g.nbits_offset == 0 || g.nbits_offset > 32) return GX_E_RANGE;
  if (g.nbits_syncbits_offset == 0 || g.nbits_syncbits_offset > 32) return GX_E_RANGE;
  if (g.bits_act > 30) return GX_E_RANGE;
  const uint32_t K = g.firstbit[N];
  if (K > 32u * 9u) return GX_E_LIMIT;  // MAX_SIZE (cu:92)
  ModelHdr h{};
  h.nprocs = N;
  h.state_bits = K;
  h.key_words = std::max<uint32_t>(1, (K + 62) / 63);

  out->floc.resize(N);
  std::vector<uint32_t> floc(N), rowbase(N), rowcnt(N);
  uint32_t nrows = 0;
  const uint32_t k = 32 / g.nbits_offset;
  for (uint32_t p = 0; p < N; p++) {
    const uint32_t lo = g.firstbit[p], w = g.firstbit[p + 1] - lo;
    out->floc[p] = FieldLoc{lo / 63, lo % 63, w};
    floc[p] = (lo / 63) | ((lo % 63) << 8) | (w << 16);
    // a state s has offsets iff fields s and s+1 of p's offset ints exist
    uint64_t nfields = static_cast<uint64_t>(g.proc_offsets_start[p + 1] - g.proc_offsets_start[p]) * k;
    if (g.proc_offsets_start[p + 1] < g.proc_offsets_start[p]) return GX_E_RANGE;
    uint64_t rows = nfields ? nfields - 1 : 0;
    uint64_t domain = (w >= 32) ? (1ull << 32) : (1ull << w);
    rows = std::min<uint64_t>(rows, domain);
    if (rows > (1u << 24)) return GX_E_LIMIT;
    rowbase[p] = nrows;
    rowcnt[p] = static_cast<uint32_t>(rows);
    nrows += static_cast<uint32_t>(rows);
  }
  h.nrows = nrows;

  std::vector<uint32_t> loc_start, syn_start, rowflags, loc_tgt, syn_lab, syn_tstart, syn_tcnt, syn_tgt;
  std::vector<uint32_t> rowlab;  // per row: 64-bit label mask as two words
  loc_start.reserve(nrows + 1);
  syn_start.reserve(nrows + 1);
  rowflags.reserve(nrows);
  uint32_t max_label = 0;
  bool any_sync = false;
  uint64_t synlab_all = 0;
  const uint32_t nt = static_cast<uint32_t>(g.proc_trans.size());
  for (uint32_t p = 0; p < N; p++) {
    const uint32_t bs = g.bits_state[p];
    const uint32_t cap = (31 - g.bits_act) / bs;  // cu:143
    for (uint32_t s = 0; s < rowcnt[p]; s++) {
      loc_start.push_back(static_cast<uint32_t>(loc_tgt.size()));
      syn_start.push_back(static_cast<uint32_t>(syn_lab.size()));
      // transition offsets (cu:887-897)
      uint32_t index = s / k, pos = s - index * k;
      uint32_t at = g.proc_offsets_start[p] + index;
      uint32_t tmp = g.proc_offsets[at];
      uint32_t o1 = bits(tmp, pos * g.nbits_offset, g.nbits_offset), o2;
      if (pos == k - 1)
        o2 = bits(g.proc_offsets[at + 1], 0, g.nbits_offset);
      else
        o2 = bits(tmp, (pos + 1) * g.nbits_offset, g.nbits_offset);
      if (o1 > o2 || o2 > nt) {
        rowflags.push_back(kRowBad);
        rowlab.push_back(0);
        rowlab.push_back(0);
        continue;
      }
      uint32_t flags = 0;
      std::map<uint32_t, std::vector<uint32_t>> by_label;
      for (uint32_t e = o1; e < o2; e++) {
        const uint32_t t = g.proc_trans[e];
        const bool sync = t & 1u;                       // cu:158
        const uint32_t lab = bits(t, 1, g.bits_act);    // cu:157
        std::vector<uint32_t> *dst = nullptr;
        if (sync) {
          dst = &by_label[lab];
        } else {
          flags |= kRowHasLocal;  // cu:910-913
        }
        for (uint32_t i = 0; i < cap; i++) {           // cu:159, 921-943
          uint32_t v = bits(t, 1 + g.bits_act + i * bs, bs);
          if (v == 0) break;
          if (sync)
            dst->push_back(v - 1);
          else
            loc_tgt.push_back(v - 1);
        }
      }
      uint64_t lm = 0;
      for (auto &kv : by_label) {  // ascending labels
        if (kv.second.size() != 1) flags |= kRowMultiSync;  // products other than 1
        if (kv.first < 64) lm |= 1ull << kv.first;
        syn_lab.push_back(kv.first);
        syn_tstart.push_back(static_cast<uint32_t>(syn_tgt.size()));
        syn_tcnt.push_back(static_cast<uint32_t>(kv.second.size()));
        syn_tgt.insert(syn_tgt.end(), kv.second.begin(), kv.second.end());
        max_label = std::max(max_label, kv.first);
        any_sync = true;
      }
      rowflags.push_back(flags);
      synlab_all |= lm;
      rowlab.push_back(static_cast<uint32_t>(lm));
      rowlab.push_back(static_cast<uint32_t>(lm >> 32));
    }
  }
  loc_start.push_back(static_cast<uint32_t>(loc_tgt.size()));
  syn_start.push_back(static_cast<uint32_t>(syn_lab.size()));

  // rules per label (cu:997-1009, 1466-1488)
  const uint32_t nlab = any_sync ? max_label + 1 : 0;
  h.nlab = nlab;
  h.synlab_lo = static_cast<uint32_t>(synlab_all);
  h.synlab_hi = static_cast<uint32_t>(synlab_all >> 32);
  h.lab64 = nlab <= 64 ? 1u : 0u;
  std::vector<uint32_t> rule_start, rules;
  const uint32_t k2 = 32 / g.nbits_syncbits_offset, per = 32 / N;
  for (uint32_t a = 0; a < nlab; a++) {
    rule_start.push_back(static_cast<uint32_t>(rules.size()));
    uint32_t i = a / k2, pos = a - i * k2;
    if (i >= g.syncbits_offsets.size()) return GX_E_RANGE;
    uint32_t l = g.syncbits_offsets[i];
    uint32_t so1 = bits(l, pos * g.nbits_syncbits_offset, g.nbits_syncbits_offset), so2;
    if (pos == k2 - 1) {
      if (i + 1 >= g.syncbits_offsets.size()) return GX_E_RANGE;
      so2 = bits(g.syncbits_offsets[i + 1], 0, g.nbits_syncbits_offset);
    } else {
      so2 = bits(l, (pos + 1) * g.nbits_syncbits_offset, g.nbits_syncbits_offset);
    }
    if (so1 > so2 || so2 > g.syncbits.size()) return GX_E_RANGE;
    for (uint32_t j = so1; j < so2; j++)
      for (uint32_t f = 0; f < per; f++) {
        uint32_t r = bits(g.syncbits[j], f * N, N);  // cu:205
        if (r) rules.push_back(r);
      }
  }
  rule_start.push_back(static_cast<uint32_t>(rules.size()));

  // lay out the buffer
  std::vector<uint32_t> &B = out->words;
  B.clear();
  auto put = [&](const std::vector<uint32_t> &v) {
    uint32_t off = static_cast<uint32_t>(B.size());
    B.insert(B.end(), v.begin(), v.end());
    return off;
  };
  // arrays only the process-group kernel reads come first; everything the
  // state-per-lane kernels read forms one section from tps_begin to the end,
  // which is all they stage in LDS
  h.off_floc = put(floc);
  h.off_loc_start = put(loc_start);
  h.off_syn_start = put(syn_start);
  h.off_rowflags = put(rowflags);
  h.off_syn_lab = put(syn_lab);
  h.off_syn_tstart = put(syn_tstart);
  h.off_syn_tcnt = put(syn_tcnt);
  h.off_rowlab = put(rowlab);
  while (B.size() % 4) B.push_back(0);
  h.tps_begin = static_cast<uint32_t>(B.size());
  h.off_rowbase = put(rowbase);
  h.off_rowcnt = put(rowcnt);
  h.off_loc_tgt = put(loc_tgt);
  h.off_syn_tgt = put(syn_tgt);
  h.off_rule_start = put(rule_start);
  h.off_rules = put(rules);
  // packed records (uint4 / uint2 reads in the state-per-lane kernels)
  std::vector<uint32_t> rowrec, synrec;
  const uint32_t nrec = static_cast<uint32_t>(rowflags.size());
  rowrec.reserve(8u * nrec);
  uint32_t first_labs[4];  // the model's first four synchronising labels
  {
    uint64_t L = synlab_all;
    for (int k = 0; k < 4; k++) {
      first_labs[k] = L ? static_cast<uint32_t>(__builtin_ctzll(L)) : 64u;
      L &= L ? L - 1 : 0;
    }
  }
  for (uint32_t r = 0; r < nrec; r++) {
    const uint64_t rl = static_cast<uint64_t>(rowlab[2 * r]) | (static_cast<uint64_t>(rowlab[2 * r + 1]) << 32);
    uint32_t labbits = 0;
    for (int k = 0; k < 4; k++)
      if (first_labs[k] < 64u && ((rl >> first_labs[k]) & 1u)) labbits |= 1u << k;
    // words 6-7: for the model's k-th synchronising label (k < 4), the
    // index of the row's single target delta when its entry for the label
    // has exactly one target (16 bits each, kNoSyn16 otherwise)
    uint32_t one[4] = {kNoSyn16, kNoSyn16, kNoSyn16, kNoSyn16};
    if (syn_tgt.size() < kNoSyn16) {
      for (uint32_t e = syn_start[r]; e < syn_start[r + 1]; e++)
        for (int k = 0; k < 4; k++)
          if (syn_lab[e] == first_labs[k] && syn_tcnt[e] == 1) one[k] = syn_tstart[e];
    }
    rowrec.insert(rowrec.end(), {loc_start[r], loc_start[r + 1], syn_start[r], rowflags[r] | (labbits << kRowLabShift),
                                 rowlab[2 * r], rowlab[2 * r + 1], one[0] | (one[1] << 16), one[2] | (one[3] << 16)});
  }
  // row record nrows: a local state outside a process's rows reads this one
  // (no entries, flagged bad) instead of branching around the read
  rowrec.insert(rowrec.end(), {0u, 0u, 0u, static_cast<uint32_t>(kRowBad), 0u, 0u, 0u, 0u});
  for (size_t e = 0; e < syn_tstart.size(); e++) {
    synrec.push_back(syn_tstart[e]);
    synrec.push_back(syn_tcnt[e]);
  }
  while (B.size() % 4) B.push_back(0);
  h.off_rowrec = put(rowrec);
  while (B.size() % 4) B.push_back(0);
  h.off_synrec = put(synrec);
  std::vector<uint32_t> fdesc;
  for (uint32_t p = 0; p < N; p++) {
    const FieldLoc &f = out->floc[p];
    const uint64_t vmask = f.width >= 32 ? 0xFFFFFFFFull : ((1ull << f.width) - 1ull);
    fdesc.insert(fdesc.end(), {f.chunk | (f.offset << 8), static_cast<uint32_t>(vmask), rowbase[p], rowcnt[p]});
  }
  while (B.size() % 4) B.push_back(0);
  h.off_fdesc = put(fdesc);
  // key-shaped forms (W 64-bit words as lo, hi pairs) for successor
  // emission: a target value already placed in its process's field bits,
  // and the field bits of a process and of a rule's members to clear first
  const uint32_t W = h.key_words;
  auto field_bits = [&](uint32_t p, std::vector<uint64_t> &k) {
    const FieldLoc &f = out->floc[p];
    if (f.width == 0) return;
    const uint64_t m = (f.width >= 32) ? 0xFFFFFFFFull : ((1ull << f.width) - 1);
    k[f.chunk] |= (m << f.offset) & kPayload;
    if (f.offset + f.width > 63) k[f.chunk + 1] |= (1ull << (f.offset + f.width - 63)) - 1;
  };
  auto put_key = [&](std::vector<uint32_t> &dst, const std::vector<uint64_t> &k) {
    for (uint32_t j = 0; j < W; j++) {
      dst.push_back(static_cast<uint32_t>(k[j]));
      dst.push_back(static_cast<uint32_t>(k[j] >> 32));
    }
  };
  auto placed = [&](uint32_t p, uint32_t v) {
    std::vector<uint64_t> k(W, 0);
    key_set(*out, k.data(), p, v);
    return k;
  };
  // A move of process p from local state s (its row's) to target v changes
  // only p's field: the successor is key XOR (s placed ^ v placed).  One such
  // delta per local and per synchronising target entry.
  std::vector<uint32_t> locx, synx;
  std::vector<uint32_t> row_proc(nrec, 0), row_state(nrec, 0);
  for (uint32_t p = 0; p < N; p++)
    for (uint32_t r = rowbase[p]; r < rowbase[p] + rowcnt[p]; r++) {
      row_proc[r] = p;
      row_state[r] = r - rowbase[p];
    }
  auto delta = [&](uint32_t r, uint32_t v) {
    std::vector<uint64_t> a = placed(row_proc[r], row_state[r]), b = placed(row_proc[r], v);
    for (uint32_t j = 0; j < W; j++) a[j] ^= b[j];
    return a;
  };
  std::vector<uint32_t> syn_row(syn_tstart.size(), 0);
  for (uint32_t r = 0; r < nrec; r++) {
    for (uint32_t t = loc_start[r]; t < loc_start[r + 1]; t++) put_key(locx, delta(r, loc_tgt[t]));
    for (uint32_t e = syn_start[r]; e < syn_start[r + 1]; e++) syn_row[e] = r;
  }
  std::vector<uint32_t> tgt_row(syn_tgt.size(), 0);
  for (size_t e = 0; e < syn_tstart.size(); e++)
    for (uint32_t t = syn_tstart[e]; t < syn_tstart[e] + syn_tcnt[e]; t++) tgt_row[t] = syn_row[e];
  for (size_t t = 0; t < syn_tgt.size(); t++) put_key(synx, delta(tgt_row[t], syn_tgt[t]));
  while (B.size() % 4) B.push_back(0);
  h.off_locx = put(locx);
  while (B.size() % 4) B.push_back(0);
  h.off_synx = put(synx);

  while (B.size() % 4) B.push_back(0);
  h.total_words = static_cast<uint32_t>(B.size());
  out->hdr = h;
  return GX_OK;
}

// ----------------------------------------------------------------------
// host key codec: bit i of the state string lives in word i/63, bit i%63
void key_from_words32(const Compiled &c, const uint32_t *v, uint64_t *key) {
  const uint32_t K = c.hdr.state_bits, W = c.hdr.key_words;
  for (uint32_t j = 0; j < W; j++) key[j] = kFlag;
  for (uint32_t i = 0; i < K; i++)
    if ((v[i / 32] >> (i % 32)) & 1u) key[i / 63] |= 1ull << (i % 63);
}

void key_to_words32(const Compiled &c, const uint64_t *key, uint32_t *v, uint32_t n32) {
  const uint32_t K = c.hdr.state_bits;
  for (uint32_t j = 0; j < n32; j++) v[j] = 0;
  for (uint32_t i = 0; i < K; i++)
    if ((key[i / 63] >> (i % 63)) & 1ull) v[i / 32] |= 1u << (i % 32);
}

uint32_t key_get(const Compiled &c, const uint64_t *key, uint32_t p) {
  const FieldLoc &f = c.floc[p];
  if (f.width == 0) return 0;
  uint64_t x = (key[f.chunk] & kPayload) >> f.offset;
  if (f.offset + f.width > 63) x |= (key[f.chunk + 1] & kPayload) << (63 - f.offset);
  return static_cast<uint32_t>(x & ((f.width >= 32) ? 0xFFFFFFFFull : ((1ull << f.width) - 1)));
}

void key_set(const Compiled &c, uint64_t *key, uint32_t p, uint32_t v) {
  const FieldLoc &f = c.floc[p];
  if (f.width == 0) return;
  const uint64_t m = (f.width >= 32) ? 0xFFFFFFFFull : ((1ull << f.width) - 1);
  const uint64_t x = v & m;
  key[f.chunk] = (key[f.chunk] & ~((m << f.offset) & kPayload)) | ((x << f.offset) & kPayload);
  if (f.offset + f.width > 63) {
    const uint32_t hi = f.offset + f.width - 63;
    const uint64_t mh = (1ull << hi) - 1;
    key[f.chunk + 1] = (key[f.chunk + 1] & ~mh) | (x >> (63 - f.offset));
  }
}

}  // namespace gx
