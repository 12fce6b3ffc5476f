// gx_tps.h -- successor generation for one state per lane.
//
// Shared by the single-model kernel (k_expand_tps, model staged in LDS) and
// the batched kernel (k_batch_expand, model read through the caches).  The
// per-process rows of the lane's state are kept in LDS as rows[p * 64 + lane]
// (one column per lane: conflict-free).  Semantics: local moves of every
// process (cu:904-949), then every enabled rule of every synchronising label
// in ascending order with the product of target lists, lowest process
// varying fastest (cu:951-1130); deadlock per cu:1134-1151; safety per
// cu:926-935, 1063-1069 checked on the successors that move the last process.
//
// Row data comes from the packed 8-word row records (one 16-byte and one
// 8-byte read per process) and (tstart, tcnt) pairs per (row, label) entry.
// The enabled sets of up to kTpsLabCache labels are computed in one extra
// pass over the processes, or in the first pass itself when the whole model
// has at most kTpsLabCache synchronising labels.  They are handed from the
// scan to the emission, together with the lane's enabled rules of each such
// label as a bit mask: the 64 lanes of a wave hold different states, and
// walking every rule of a label would make each lane pay for the rules
// enabled in any lane.  For the same reason each lane walks its own
// processes with local targets (a bit mask from the scan) instead of the
// wave stepping through every process.  A rule's product is 1 unless a
// member's row has a synchronising entry with a target count other than 1
// (kRowMultiSync), so members are visited only then.  Successors are the
// key XOR the model's key-shaped target deltas; a rule whose members all
// have one target finds each member's delta through the row record.
#pragma once
#include "gx_device.h"

namespace gx {

// The header with the state-per-lane section's offsets rebased onto a copy
// of words [tps_begin, total_words) (the LDS-staged section).
__device__ __forceinline__ ModelHdr tps_rebased(const ModelHdr &h) {
  ModelHdr r = h;
  const uint32_t b = h.tps_begin;
  r.off_rowbase -= b;
  r.off_rowcnt -= b;
  r.off_loc_tgt -= b;
  r.off_syn_tgt -= b;
  r.off_rule_start -= b;
  r.off_rules -= b;
  r.off_rowrec -= b;
  r.off_synrec -= b;
  r.off_fdesc -= b;
  r.off_locx -= b;
  r.off_synx -= b;
  return r;
}

constexpr uint32_t kNoRow = 0xFFFFFFFFu;

// rows[p * 64 + lane] holds row indices in RowT (uint16_t when the model has
// fewer than 0xFFFF rows, halving the LDS they take); the all-ones value of
// RowT stands for kNoRow
template <typename RowT>
__device__ __forceinline__ uint32_t row_at(const RowT *rows, int i) {
  const uint32_t v = rows[i];
  if constexpr (sizeof(RowT) == 2) return v == 0xFFFFu ? kNoRow : v;
  return v;
}
constexpr int kTpsLabCache = 4;

// loc_start, loc_end, syn_start, flags of a row
__device__ __forceinline__ uint4 tps_row(const uint32_t *M, const ModelHdr &H, uint32_t row) {
  return *reinterpret_cast<const uint4 *>(M + H.off_rowrec + 8u * row);
}

// synchronising labels of a row (labels < 64)
__device__ __forceinline__ uint64_t tps_rowlab(const uint32_t *M, const ModelHdr &H, uint32_t row) {
  const uint2 v = *reinterpret_cast<const uint2 *>(M + H.off_rowrec + 8u * row + 4u);
  return static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
}

// (tstart, tcnt) of the (row, l) entry: l is present in the row's label mask
__device__ __forceinline__ uint2 tps_syn(const uint32_t *M, const ModelHdr &H, uint32_t row, uint32_t l) {
  const uint64_t below = (1ull << l) - 1ull;
  const uint32_t e = M[H.off_rowrec + 8u * row + 2u] + static_cast<uint32_t>(__popcll(tps_rowlab(M, H, row) & below));
  return *reinterpret_cast<const uint2 *>(M + H.off_synrec + 2u * e);
}

// processes whose current row has an entry for label l
template <typename RowT>
__device__ __forceinline__ uint32_t tps_enabled(const uint32_t *M, const ModelHdr &H, const RowT *rows, int lane,
                                                int N, uint32_t l) {
  uint32_t en = 0;
  for (int p = 0; p < N; p++) {
    const uint32_t row = row_at(rows, p * 64 + lane);
    if (row != kNoRow && ((tps_rowlab(M, H, row) >> l) & 1ull)) en |= 1u << p;
  }
  return en;
}

struct TpsInfo {
  uint64_t labs;  // synchronising labels present in the state
  uint64_t cnt;   // successors
  uint32_t lmask; // processes whose row has local targets
  uint32_t multi; // processes whose row has a synchronising entry with other than one target
  uint32_t en[kTpsLabCache];     // enabled sets of the first labels of labs (when cached)
  uint64_t rmask[kTpsLabCache];  // enabled rules of those labels, bit k = rule start + k
  bool cached;    // labs has at most kTpsLabCache labels
  bool rcached;   // ... and each of them has at most 64 rules
  bool dead;      // no local entry and no enabled rule
  bool viol;      // a successor has the last process in local state 1
  bool bad;       // a local state without transition offsets was reached
};

__device__ __forceinline__ uint64_t tps_rmask_of(const TpsInfo &inf, int i) {
  return i == 0 ? inf.rmask[0] : i == 1 ? inf.rmask[1] : i == 2 ? inf.rmask[2] : inf.rmask[3];
}

// enabled set of the i-th label l of labs
template <typename RowT>
__device__ __forceinline__ uint32_t tps_en_of(const TpsInfo &inf, int i, const uint32_t *M, const ModelHdr &H,
                                              const RowT *rows, int lane, int N, uint32_t l) {
  if (!inf.cached) return tps_enabled(M, H, rows, lane, N, l);
  return i == 0 ? inf.en[0] : i == 1 ? inf.en[1] : i == 2 ? inf.en[2] : inf.en[3];
}

// A process's record for the scan: where its field sits in the key, its
// value mask, and its rows (ModelHdr::off_fdesc).
struct ProcRec {
  uint32_t c, o, vmask, rowbase, rowcnt;
};

__device__ __forceinline__ ProcRec proc_rec(const uint32_t *M, const ModelHdr &H, int p) {
  const uint4 a = *reinterpret_cast<const uint4 *>(M + H.off_fdesc + 4u * static_cast<uint32_t>(p));
  return ProcRec{a.x & 0xFFu, a.x >> 8, a.y, a.z, a.w};
}

// Local state of process p: the field's bits in word c from offset o, and
// from bit 0 of word c+1 when the field straddles the two.  A field that
// does not straddle has its width <= 63 - o, so whatever word c+1 adds
// lands above the width and the value mask drops it: no straddle branch.
template <int W>
__device__ __forceinline__ uint32_t proc_state(const uint64_t (&key)[W], const ProcRec &f) {
  uint64_t lo = 0, hi = 0;
#pragma unroll
  for (int j = 0; j < W; j++) {
    if (static_cast<uint32_t>(j) == f.c) lo = key[j];
    if (static_cast<uint32_t>(j) == f.c + 1) hi = key[j];
  }
  return static_cast<uint32_t>(((lo & kPayload) >> f.o) | ((hi & kPayload) << (63u - f.o))) & f.vmask;
}

// Pass 1 for the lane's state: rows, successor count, deadlock, safety.
// The per-process loop has no branch on lane data: a local state outside
// the process's rows reads the bad row record at index nrows, and lanes
// without a state (has == false, key all zero) run it on local states 0
// and have their results cleared at the end.
template <int W, typename RowT>
__device__ __forceinline__ TpsInfo tps_scan(const uint32_t *M, const ModelHdr &H, const uint64_t (&key)[W],
                                            bool has, RowT *rows, int lane, bool safety) {
  const int N = static_cast<int>(H.nprocs);
  const uint32_t fmask_last = proc_rec(M, H, N - 1).vmask;
  TpsInfo r{};
  bool anyloc = false, anyen = false, bad = false;
  uint32_t multi = 0;  // processes whose row has a multi-target synchronising entry
  uint32_t nloc = 0;   // local successors
  // A model with at most kTpsLabCache synchronising labels gets every
  // label's enabled set in this pass (men[k] for its k-th label), from the
  // label bits of the row records' flag words: the rows' 64-bit label masks
  // are not read at all.
  static_assert(kTpsLabCache == 4, "row records carry four label bits");
  const uint64_t msyn = static_cast<uint64_t>(H.synlab_lo) | (static_cast<uint64_t>(H.synlab_hi) << 32);
  const bool men_all = __popcll(msyn) <= kTpsLabCache;
  uint32_t mlab[kTpsLabCache], men[kTpsLabCache];
  {
    uint64_t L = msyn;
#pragma unroll
    for (int k = 0; k < kTpsLabCache; k++) {
      mlab[k] = L ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(L)) - 1) : 64u;
      L &= L - 1;
      men[k] = 0;
    }
  }
  for (int p = 0; p < N; p++) {
    const ProcRec f = proc_rec(M, H, p);
    const uint32_t sp = proc_state<W>(key, f);
    const uint32_t row = sp < f.rowcnt ? f.rowbase + sp : H.nrows;
    const uint4 ra = tps_row(M, H, row);
    const bool valid = (ra.w & kRowBad) == 0;
    bad |= !valid;
    rows[p * 64 + lane] = static_cast<RowT>(valid ? row : kNoRow);
    const uint32_t w = valid ? ra.w : 0u;
    const uint32_t nl = valid ? ra.y - ra.x : 0u;
    nloc += nl;
    r.lmask |= (nl != 0 ? 1u : 0u) << p;
    anyloc |= (w & kRowHasLocal) != 0;
    multi |= ((w & kRowMultiSync) != 0 ? 1u : 0u) << p;
    if (men_all) {  // the row's labels among the model's (at most four) from its flag word
#pragma unroll
      for (int k = 0; k < kTpsLabCache; k++) men[k] |= ((w >> (kRowLabShift + k)) & 1u) << p;
    } else if (valid) {
      r.labs |= tps_rowlab(M, H, row);
    }
    if (safety && p == N - 1 && nl)
      for (uint32_t t = ra.x; t < ra.y; t++)
        if ((M[H.off_loc_tgt + t] & fmask_last) == 1u) r.viol = true;
  }
  r.cnt = nloc;
  r.bad = bad;
  r.multi = multi;
  if (men_all) {  // every synchronising label of the model is one of mlab[]
#pragma unroll
    for (int k = 0; k < kTpsLabCache; k++)
      if (mlab[k] < 64u && men[k]) r.labs |= 1ull << mlab[k];
  }
  // enabled sets of the first labels in one pass over the processes
  uint32_t lab[kTpsLabCache];
  {
    uint64_t L = r.labs;
#pragma unroll
    for (int i = 0; i < kTpsLabCache; i++) {
      lab[i] = L ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(L)) - 1) : 64u;
      L &= L - 1;
      r.en[i] = 0;
    }
    r.cached = (L == 0);
  }
  if (r.cached && r.labs && men_all) {
#pragma unroll
    for (int i = 0; i < kTpsLabCache; i++) {
      if (lab[i] >= 64u) continue;
      const int k = __popcll(msyn & ((1ull << lab[i]) - 1ull));  // lab[i] is the model's k-th label
      r.en[i] = k == 0 ? men[0] : k == 1 ? men[1] : k == 2 ? men[2] : men[3];
    }
  } else if (r.cached && r.labs) {
    for (int p = 0; p < N; p++) {
      const uint32_t row = row_at(rows, p * 64 + lane);
      if (row == kNoRow) continue;
      const uint64_t rl = tps_rowlab(M, H, row);
#pragma unroll
      for (int i = 0; i < kTpsLabCache; i++)
        if (lab[i] < 64u && ((rl >> lab[i]) & 1ull)) r.en[i] |= 1u << p;
    }
  }
  r.rcached = r.cached;
  int i = 0;
  for (uint64_t L = r.labs; L; L &= L - 1, i++) {
    const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(L)) - 1);
    const uint32_t en = tps_en_of(r, i, M, H, rows, lane, N, l);
    const uint32_t rs = M[H.off_rule_start + l], re = M[H.off_rule_start + l + 1];
    if (re - rs > 64) r.rcached = false;
    uint64_t rm = 0;
    for (uint32_t j = rs; j < re; j++) {
      const uint32_t rule = M[H.off_rules + j];
      if (rule & ~en) continue;
      anyen = true;
      if (j - rs < 64) rm |= 1ull << (j - rs);
      uint64_t prod = 1;
      if (rule & multi) {
        for (uint32_t rr = rule; rr; rr &= rr - 1) {
          const int q = __ffs(rr) - 1;
          prod *= tps_syn(M, H, row_at(rows, q * 64 + lane), l).y;
        }
      }
      r.cnt += prod;
      if (safety && ((rule >> (N - 1)) & 1u)) {
        const uint2 e = tps_syn(M, H, row_at(rows, (N - 1) * 64 + lane), l);
        for (uint32_t t = 0; t < e.y; t++)
          if ((M[H.off_syn_tgt + e.x + t] & fmask_last) == 1u) r.viol = true;
      }
    }
    if (i == 0) r.rmask[0] = rm;
    if (i == 1) r.rmask[1] = rm;
    if (i == 2) r.rmask[2] = rm;
    if (i == 3) r.rmask[3] = rm;
  }
  r.dead = has && !anyloc && !anyen;
  if (!has) {
    r.cnt = 0;
    r.lmask = 0;
    r.labs = 0;
    r.viol = false;
    r.bad = false;
  }
  return r;
}

// The i-th key-shaped record at `off` as W key words.  Records hold the
// model's own key_words `kw` (lo, hi word pairs); a batch may run the model
// with wider keys (kw < W), whose upper words are then zero.  16-byte reads
// when kw == W is even: the records then start 16-byte aligned.
template <int W>
__device__ __forceinline__ void tps_keyrec(const uint32_t *M, uint32_t off, uint32_t i, uint32_t kw,
                                           uint64_t (&x)[W]) {
  const uint32_t *q = M + off + 2u * kw * i;
  if constexpr (W % 2 == 0) {
    if (kw == static_cast<uint32_t>(W)) {
#pragma unroll
      for (int j = 0; j < W; j += 2) {
        const uint4 v = *reinterpret_cast<const uint4 *>(q + 2 * j);
        x[j] = static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
        x[j + 1] = static_cast<uint64_t>(v.z) | (static_cast<uint64_t>(v.w) << 32);
      }
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < W; j++) {
    x[j] = 0;
    if (static_cast<uint32_t>(j) < kw) {
      const uint2 v = *reinterpret_cast<const uint2 *>(q + 2 * j);
      x[j] = static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
    }
  }
}

// Pass 2: calls emit(s) for every successor s of the lane's state, in the
// order counted by tps_scan.  A move of process p from its current local
// state to a target changes only p's field, so a successor is the key XOR
// the moving processes' target deltas (the model holds one per target
// entry: current and target state, both placed in p's field, XORed).
// Products are below 2^32 (they fit the successor buffer).
// EXACT: the model's key_words equal W (the single-model kernel), so the
// record width is a constant and the record reads carry no width branches.
template <int W, bool EXACT = false, typename RowT, typename Emit>
__device__ __forceinline__ void tps_emit(const uint32_t *M, const ModelHdr &H, const uint64_t (&key)[W],
                                         const RowT *rows, int lane, const TpsInfo &inf, Emit &&emit) {
  const int N = static_cast<int>(H.nprocs);
  const uint32_t kw = EXACT ? static_cast<uint32_t>(W) : H.key_words;
  // Local moves: each lane walks its own processes with local targets
  // (ascending, targets in order), so the wave takes as many steps as its
  // busiest lane has local successors, not the sum over processes of the
  // busiest lane per process.
  {
    uint32_t lm = inf.lmask, t = 0, te = 0;
    while (ballot(lm != 0 || t < te)) {
      if (t == te && lm) {
        const uint32_t p = static_cast<uint32_t>(__ffs(lm) - 1);
        lm &= lm - 1u;
        const uint4 ra = tps_row(M, H, row_at(rows, static_cast<int>(p) * 64 + lane));
        t = ra.x;
        te = ra.y;
      }
      if (t < te) {
        uint64_t s[W];
        tps_keyrec<W>(M, H.off_locx, t, kw, s);
#pragma unroll
        for (int j = 0; j < W; j++) s[j] ^= key[j];
        emit(s);
        t++;
      }
    }
  }
  const uint64_t msyn = static_cast<uint64_t>(H.synlab_lo) | (static_cast<uint64_t>(H.synlab_hi) << 32);
  const bool syn4 = __popcll(msyn) <= kTpsLabCache;  // row records hold single-target delta indices
  int i = 0;
  for (uint64_t L = inf.labs; L; L &= L - 1, i++) {
    const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(L)) - 1);
    const uint32_t rs = M[H.off_rule_start + l], re = M[H.off_rule_start + l + 1];
    // l is the model's k-th synchronising label: word 6 + k / 2, half k % 2
    const uint32_t k = static_cast<uint32_t>(__popcll(msyn & ((1ull << l) - 1ull)));
    // the lane's enabled rules: from the scan's mask, else by walking the label's rules
    const uint32_t en = inf.rcached ? 0u : tps_en_of(inf, i, M, H, rows, lane, N, l);
    uint64_t rm = inf.rcached ? tps_rmask_of(inf, i) : 0ull;
    for (uint32_t j = rs; inf.rcached ? rm != 0 : j < re; j++) {
      if (inf.rcached) {
        j = rs + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(rm)) - 1);
        rm &= rm - 1;
      }
      const uint32_t rule = M[H.off_rules + j];
      if (!inf.rcached && (rule & ~en)) continue;
      uint64_t s[W];
#pragma unroll
      for (int jj = 0; jj < W; jj++) s[jj] = key[jj];
      if (syn4 && (rule & inf.multi) == 0) {
        // every member has exactly one target: one successor, each member's
        // delta found from its row record
        for (uint32_t rr = rule; rr; rr &= rr - 1) {
          const int q = __ffs(rr) - 1;
          const uint32_t row = row_at(rows, q * 64 + lane);
          const uint32_t w = M[H.off_rowrec + 8u * row + 6u + (k >> 1)];
          uint32_t d = (w >> (16u * (k & 1u))) & 0xFFFFu;
          if (d == kNoSyn16) d = tps_syn(M, H, row, l).x;  // a model with 2^16 or more such targets
          uint64_t x[W];
          tps_keyrec<W>(M, H.off_synx, d, kw, x);
#pragma unroll
          for (int jj = 0; jj < W; jj++) s[jj] ^= x[jj];
        }
        emit(s);
        continue;
      }
      // combination 0 (every member at its first target) with the product
      uint32_t prod = 1;
      for (uint32_t rr = rule; rr; rr &= rr - 1) {
        const int q = __ffs(rr) - 1;
        const uint2 e = tps_syn(M, H, row_at(rows, q * 64 + lane), l);
        prod *= e.y;
        if (e.y) {
          uint64_t x[W];
          tps_keyrec<W>(M, H.off_synx, e.x, kw, x);
#pragma unroll
          for (int jj = 0; jj < W; jj++) s[jj] ^= x[jj];
        }
      }
      if (prod == 0) continue;  // a member without targets: the rule yields nothing
      emit(s);
      for (uint32_t c = 1; c < prod; c++) {  // lowest process varies fastest
#pragma unroll
        for (int jj = 0; jj < W; jj++) s[jj] = key[jj];
        uint32_t rem = c;
        for (uint32_t rr = rule; rr; rr &= rr - 1) {
          const int q = __ffs(rr) - 1;
          const uint2 e = tps_syn(M, H, row_at(rows, q * 64 + lane), l);
          const uint32_t d = rem % e.y;
          rem /= e.y;
          uint64_t x[W];
          tps_keyrec<W>(M, H.off_synx, e.x + d, kw, x);
#pragma unroll
          for (int jj = 0; jj < W; jj++) s[jj] ^= x[jj];
        }
        emit(s);
      }
    }
  }
}

}  // namespace gx
