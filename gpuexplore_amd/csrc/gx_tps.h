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
#pragma once
#include "gx_device.h"

namespace gx {

constexpr uint32_t kNoRow = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t tps_rowlab(const uint32_t *M, const ModelHdr &H, uint32_t row) {
  return static_cast<uint64_t>(M[H.off_rowlab + 2 * row]) |
         (static_cast<uint64_t>(M[H.off_rowlab + 2 * row + 1]) << 32);
}

// index of the (row, label l) entry: l is present in the row's label mask
__device__ __forceinline__ uint32_t tps_syn_entry(const uint32_t *M, const ModelHdr &H, uint32_t row,
                                                  uint32_t l) {
  const uint64_t below = (l >= 64) ? ~0ull : ((1ull << l) - 1ull);
  return M[H.off_syn_start + row] + static_cast<uint32_t>(__popcll(tps_rowlab(M, H, row) & below));
}

struct TpsInfo {
  uint64_t labs;  // synchronising labels present in the state
  uint64_t cnt;   // successors
  bool dead;      // no local entry and no enabled rule
  bool viol;      // a successor has the last process in local state 1
  bool bad;       // a local state without transition offsets was reached
};

// Pass 1 for the lane's state: rows, successor count, deadlock, safety.
template <int W>
__device__ __forceinline__ TpsInfo tps_scan(const uint32_t *M, const ModelHdr &H, const uint64_t (&key)[W],
                                            bool has, uint32_t *rows, int lane, bool safety) {
  const int N = static_cast<int>(H.nprocs);
  const uint32_t wlast = M[H.off_floc + N - 1] >> 16;
  const uint32_t fmask_last = (wlast >= 32) ? 0xFFFFFFFFu : ((1u << wlast) - 1u);
  TpsInfo r{0, 0, false, false, false};
  bool anyloc = false, anyen = false;
  for (int p = 0; p < N; p++) {
    uint32_t row = kNoRow;
    if (has) {
      const uint32_t sp = kget<W>(key, M[H.off_floc + p]);
      if (sp < M[H.off_rowcnt + p]) {
        row = M[H.off_rowbase + p] + sp;
        if (M[H.off_rowflags + row] & kRowBad) row = kNoRow;
      }
      if (row == kNoRow) r.bad = true;
    }
    rows[p * 64 + lane] = row;
    if (row != kNoRow) {
      const uint32_t l0 = M[H.off_loc_start + row], l1 = M[H.off_loc_start + row + 1];
      r.cnt += l1 - l0;
      anyloc |= (M[H.off_rowflags + row] & kRowHasLocal) != 0;
      r.labs |= tps_rowlab(M, H, row);
      if (safety && p == N - 1)
        for (uint32_t t = l0; t < l1; t++)
          if ((M[H.off_loc_tgt + t] & fmask_last) == 1u) r.viol = true;
    }
  }
  for (uint64_t L = r.labs; L; L &= L - 1) {
    const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(L)) - 1);
    uint32_t en = 0;
    for (int p = 0; p < N; p++) {
      const uint32_t row = rows[p * 64 + lane];
      if (row != kNoRow && ((tps_rowlab(M, H, row) >> l) & 1ull)) en |= 1u << p;
    }
    const uint32_t rs = M[H.off_rule_start + l], re = M[H.off_rule_start + l + 1];
    for (uint32_t j = rs; j < re; j++) {
      const uint32_t rule = M[H.off_rules + j];
      if (rule & ~en) continue;
      anyen = true;
      uint64_t prod = 1;
      for (uint32_t rr = rule; rr; rr &= rr - 1) {
        const int q = __ffs(rr) - 1;
        prod *= M[H.off_syn_tcnt + tps_syn_entry(M, H, rows[q * 64 + lane], l)];
      }
      r.cnt += prod;
      if (safety && ((rule >> (N - 1)) & 1u)) {
        const uint32_t e = tps_syn_entry(M, H, rows[(N - 1) * 64 + lane], l);
        const uint32_t ts = M[H.off_syn_tstart + e], tc = M[H.off_syn_tcnt + e];
        for (uint32_t i = 0; i < tc; i++)
          if ((M[H.off_syn_tgt + ts + i] & fmask_last) == 1u) r.viol = true;
      }
    }
  }
  r.dead = has && !anyloc && !anyen;
  if (!has) r.cnt = 0;
  return r;
}

// Pass 2: calls emit(s) for every successor s of the lane's state, in the
// order counted by tps_scan.  Products are below 2^32 (they fit the
// successor buffer).
template <int W, typename Emit>
__device__ __forceinline__ void tps_emit(const uint32_t *M, const ModelHdr &H, const uint64_t (&key)[W],
                                         const uint32_t *rows, int lane, uint64_t labs, Emit &&emit) {
  const int N = static_cast<int>(H.nprocs);
  for (int p = 0; p < N; p++) {
    const uint32_t row = rows[p * 64 + lane];
    if (row == kNoRow) continue;
    const uint32_t fl = M[H.off_floc + p];
    const uint32_t l0 = M[H.off_loc_start + row], l1 = M[H.off_loc_start + row + 1];
    for (uint32_t t = l0; t < l1; t++) {
      uint64_t s[W];
#pragma unroll
      for (int j = 0; j < W; j++) s[j] = key[j];
      kset<W>(s, fl, M[H.off_loc_tgt + t]);
      emit(s);
    }
  }
  for (uint64_t L = labs; L; L &= L - 1) {
    const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(L)) - 1);
    uint32_t en = 0;
    for (int p = 0; p < N; p++) {
      const uint32_t row = rows[p * 64 + lane];
      if (row != kNoRow && ((tps_rowlab(M, H, row) >> l) & 1ull)) en |= 1u << p;
    }
    const uint32_t rs = M[H.off_rule_start + l], re = M[H.off_rule_start + l + 1];
    for (uint32_t j = rs; j < re; j++) {
      const uint32_t rule = M[H.off_rules + j];
      if (rule & ~en) continue;
      uint32_t prod = 1;
      for (uint32_t rr = rule; rr; rr &= rr - 1)
        prod *= M[H.off_syn_tcnt + tps_syn_entry(M, H, rows[(__ffs(rr) - 1) * 64 + lane], l)];
      for (uint32_t k = 0; k < prod; k++) {
        uint64_t s[W];
#pragma unroll
        for (int jj = 0; jj < W; jj++) s[jj] = key[jj];
        uint32_t rem = k;
        for (uint32_t rr = rule; rr; rr &= rr - 1) {  // lowest process varies fastest
          const int q = __ffs(rr) - 1;
          const uint32_t e = tps_syn_entry(M, H, rows[q * 64 + lane], l);
          const uint32_t tc = M[H.off_syn_tcnt + e];
          const uint32_t d = rem % tc;
          rem /= tc;
          kset<W>(s, M[H.off_floc + q], M[H.off_syn_tgt + M[H.off_syn_tstart + e] + d]);
        }
        emit(s);
      }
    }
  }
}

}  // namespace gx
