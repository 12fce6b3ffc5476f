// gx_group.h -- successor generation in the process-group layout.
//
// A wave of 64 lanes expands G = floor(64/N) states at once: lane
// (g*N + p) handles process p of the g-th state (the reference uses groups
// of N lanes inside a 32-lane warp, cu:64-82).  Synchronising labels are
// walked in ascending order; the enabled set En(a) is a ballot over the
// group; rules of the label are split over the group's lanes and each lane
// enumerates the product of target lists of its enabled rules, lowest
// process varying fastest (cu:951-1130).
//
// Partial-order reduction (cu:1340-1696) selects a process set per state
// from the group's clusters (cample, cu:1616-1672), the single-process ample
// set or the stubborn closure; the cycle proviso (B6 of the semantics: a
// successor is new unless visited at a level <= the level being expanded)
// looks successors up in the visited set with find_thread.
//
// Every state is walked in a counting pass and again in an emitting pass:
// the counts give each lane its output offset (wave prefix sum) and the
// caller reserves the wave's whole output at once.  Shared by the
// single-model kernel (k_expand) and the batch's POR kernel
// (k_batch_expand_por), which call group_expand once per G states.
#pragma once
#include "../../include/gpuexplore.h"
#include "gx_device.h"

namespace gx {

constexpr uint32_t kNoLabel = 0xFFFFFFFFu;

// Per-lane constants of the layout for one model.
struct GroupLane {
  int N, G, g, p, base;
  bool in_group, last_proc;
  uint64_t nmask;
  GroupRot rot;
  uint32_t floc, rowbase, rowcnt, fmask, fmask_last;
};

__device__ __forceinline__ GroupLane group_lane(const uint32_t *M, const ModelHdr &H) {
  GroupLane L;
  L.N = static_cast<int>(H.nprocs);
  L.G = 64 / L.N;
  const int lane = lane_id();
  L.g = lane / L.N;
  L.p = lane - L.g * L.N;
  L.in_group = L.g < L.G;
  L.base = L.in_group ? L.g * L.N : 0;
  L.nmask = (L.N >= 64) ? ~0ull : ((1ull << L.N) - 1ull);
  L.rot = make_rot(L.N, L.base, L.p, L.in_group);
  L.floc = L.in_group ? M[H.off_floc + L.p] : 0u;
  L.rowbase = L.in_group ? M[H.off_rowbase + L.p] : 0u;
  L.rowcnt = L.in_group ? M[H.off_rowcnt + L.p] : 0u;
  const uint32_t fwidth = L.floc >> 16;
  L.fmask = (fwidth >= 32) ? 0xFFFFFFFFu : ((1u << fwidth) - 1u);
  L.last_proc = L.in_group && L.p == L.N - 1;
  const uint32_t wlast = M[H.off_floc + L.N - 1] >> 16;
  L.fmask_last = (wlast >= 32) ? 0xFFFFFFFFu : ((1u << wlast) - 1u);
  return L;
}

// Per-wave LDS exchange of the layout: 64 x uint2 + 2 x 64 x uint32.
struct GroupScratch {
  uint2 *xch;   // (tstart, tcnt) of each lane's entry for the current label
  uint32_t *xc; // per-lane cluster / stubborn relation
  uint32_t *xf; // per-lane final sets
};
constexpr size_t kGroupScratchBytesPerWave = 64 * (8 + 4 + 4);

// What a wave accumulates over its batches.
struct GroupAcc {
  uint64_t trans = 0, states = 0;
  bool viol = false, bad = false;
};

// Expands the states held by the wave's groups (group g: `key`, `has`).
// reserve(total, &obase) -> 1 (output at obase), 0 (defer these states) or
// -1 (they can never fit); emit(pos, s) stores successor s at position pos.
// Returns reserve's result; on 1 the successors are emitted and counted into
// acc.  `fmt` is the slot format of the visited set (level stamps under the
// proviso), `level` the level being expanded.
template <int W, int MODE, int WT, typename Reserve, typename Emit>
__device__ __forceinline__ int group_expand(const uint32_t *M, const ModelHdr &H, const GroupLane &L,
                                            const GroupScratch &X, const uint64_t (&key)[W], bool has,
                                            const SlotFmt<WT> &fmt, const Table &T, int bucket_words,
                                            uint32_t level, int property, int por_method, Reserve &&reserve,
                                            Emit &&emit, GroupAcc &acc) {
  const int N = L.N, p = L.p, base = L.base;
  const bool in_group = L.in_group;
  const uint64_t nmask = L.nmask;
  const int lane = lane_id();
  uint2 *xch = X.xch;
  uint32_t *xc = X.xc, *xf = X.xf;
  const bool safety = property == GX_PROP_SAFETY;
  bool viol = false;
  uint32_t l0 = 0, l1 = 0, e0 = 0, e1 = 0, rf = 0;
  if (has) {
    const uint32_t sp = kget<W>(key, L.floc);
    if (sp < L.rowcnt) {
      const uint32_t row = L.rowbase + sp;
      rf = M[H.off_rowflags + row];
      if (rf & kRowBad) {
        acc.bad = true;
      } else {
        l0 = M[H.off_loc_start + row];
        l1 = M[H.off_loc_start + row + 1];
        e0 = M[H.off_syn_start + row];
        e1 = M[H.off_syn_start + row + 1];
      }
    } else {
      acc.bad = true;
    }
  }

  // Walk the synchronising labels of the G states in ascending order.
  // on_label(cur, en, mine) runs on every lane of a group whose label is
  // `cur`; on_rule(r, cur) on the lane owning an enabled rule r.
  auto walk = [&](auto &&on_label, auto &&on_rule) {
    uint32_t e = e0;
    for (;;) {
      const uint32_t mylab = (e < e1) ? M[H.off_syn_lab + e] : kNoLabel;
      const uint32_t curl = group_min(mylab, L.rot);
      if (!ballot(in_group && curl != kNoLabel)) break;
      const bool mine = in_group && mylab == curl && curl != kNoLabel;
      const uint64_t en = (ballot(mine) >> base) & nmask;
      xch[lane] = mine ? make_uint2(M[H.off_syn_tstart + e], M[H.off_syn_tcnt + e]) : make_uint2(0u, 0u);
      wave_sync();
      if (in_group && curl != kNoLabel) {
        on_label(curl, en, mine);
        const uint32_t rs = M[H.off_rule_start + curl], re = M[H.off_rule_start + curl + 1];
        for (uint32_t j = rs + static_cast<uint32_t>(p); j < re; j += static_cast<uint32_t>(N)) {
          const uint32_t r = M[H.off_rules + j];
          if ((static_cast<uint64_t>(r) & ~en) == 0) on_rule(r, curl);
        }
      }
      wave_sync();
      if (mine) e++;
    }
  };
  auto rule_product = [&](uint32_t r) {
    uint64_t prod = 1;
    for (uint32_t rr = r; rr; rr &= rr - 1) prod *= xch[base + __ffs(rr) - 1].y;
    return prod;
  };
  // does T_{N-1}(cur) contain local state 1 (successor violates safety)
  auto last_hits_one = [&]() {
    const uint2 t = xch[base + N - 1];
    for (uint32_t i = 0; i < t.y; i++)
      if ((M[H.off_syn_tgt + t.x + i] & L.fmask_last) == 1u) return true;
    return false;
  };
  // is successor s new: not in the visited set at a level <= `level`
  auto is_new = [&](const uint64_t (&s)[W]) {
    uint64_t v[WT];
    slot_value<W, WT>(s, fmt, v);
    uint32_t stamp = 0;
    const int f = find_thread<WT>(T, v, fmt, key_hash<W>(s, T.seed), bucket_words, H.state_bits, &stamp);
    return !f || stamp > level;
  };

  // ---------------- pass 1: enabledness, safety, (POR) clusters
  bool any_enabled = false;
  uint32_t cluster0 = in_group ? (1u << p) : 0u;
  uint32_t stub = cluster0;  // stubborn-set relation d(p)
  bool in_en = false;
  uint32_t rnew = 0;  // enabled rules with a new successor (proviso)
  bool local_new = false;
  if (safety && L.last_proc) {
    for (uint32_t t = l0; t < l1; t++)
      if ((M[H.off_loc_tgt + t] & L.fmask) == 1u) viol = true;
  }
  walk(
      [&](uint32_t curl, uint64_t en, bool mine) {
        if (MODE != kModePlain && mine) {
          const uint32_t rs = M[H.off_rule_start + curl], re = M[H.off_rule_start + curl + 1];
          for (uint32_t j = rs; j < re; j++) {
            const uint32_t r = M[H.off_rules + j];
            if ((r >> p) & 1u) {
              cluster0 |= r;
              // the whole rule when every member has the label, else its
              // lowest member without it (the moves that can enable it)
              const uint32_t lack = r & ~static_cast<uint32_t>(en);
              stub |= lack ? (lack & (~lack + 1u)) : r;
              if ((static_cast<uint64_t>(r) & ~en) == 0 && rule_product(r) > 0) in_en = true;
            }
          }
        }
      },
      [&](uint32_t r, uint32_t curl) {
        any_enabled = true;
        if (safety && ((r >> (N - 1)) & 1u) && last_hits_one()) viol = true;
        if (MODE == kModePorProviso) {
          // enumerate until a successor is new (not visited at level <= L)
          const uint64_t prod64 = rule_product(r);
          const uint32_t prod = prod64 > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(prod64);
          bool found_new = false;
          for (uint32_t k = 0; k < prod && !found_new; k++) {
            uint64_t s[W];
#pragma unroll
            for (int j = 0; j < W; j++) s[j] = key[j];
            uint32_t rem = k;
            for (uint32_t rr = r; rr; rr &= rr - 1) {
              const int q = __ffs(rr) - 1;
              const uint2 t = xch[base + q];
              const uint32_t d = rem % t.y;
              rem /= t.y;
              kset<W>(s, M[H.off_floc + q], M[H.off_syn_tgt + t.x + d]);
            }
            if (is_new(s)) found_new = true;
          }
          if (found_new) rnew |= r;
        }
        (void)curl;
      });
  const bool local_entry = (rf & kRowHasLocal) != 0;
  const bool dead = has && ((ballot(local_entry) >> base) & nmask) == 0 &&
                    ((ballot(any_enabled) >> base) & nmask) == 0;
  if (property == GX_PROP_DEADLOCK && dead) viol = true;  // cu:1134-1151

  // ---------------- POR: cluster selection (cu:1633-1672)
  uint32_t csel = static_cast<uint32_t>(nmask);
  if (MODE != kModePlain) {
    bool prov;
    if (MODE == kModePor) {
      prov = (l1 > l0) || in_en;
    } else {
      for (uint32_t t = l0; t < l1 && !local_new; t++) {
        uint64_t s[W];
#pragma unroll
        for (int j = 0; j < W; j++) s[j] = key[j];
        kset<W>(s, L.floc, M[H.off_loc_tgt + t]);
        if (is_new(s)) local_new = true;
      }
      // OR of the new-successor rules over the group
      const uint32_t accr = group_or(rnew, L.rot);
      prov = local_new || ((accr >> p) & 1u);
    }
    if (por_method == GX_POR_AMPLE) {
      // the local moves of one process none of whose current labels has a
      // rule containing it (cluster0 = {p}), with a (new) local
      // successor; the fewest local successors win, the lowest on ties
      const bool cand = in_group && cluster0 == (1u << p) &&
                        (MODE == kModePorProviso ? local_new : (l1 > l0));
      xf[lane] = cand ? (l1 - l0) : 0xFFFFFFFFu;
      wave_sync();
      if (in_group) {
        uint32_t best = 0xFFFFFFFFu;
        for (int q = 0; q < N; q++) {
          const uint32_t nq = xf[base + q];
          if (nq < best) {
            best = nq;
            csel = 1u << q;
          }
        }
      }
      wave_sync();
    } else if (por_method == GX_POR_STUBBORN) {
      // closure of {p} under d; valid when it meets a process with an
      // enabled (under the proviso: new) successor; the smallest valid
      // set below N processes wins, the lowest p on ties
      const uint32_t target = static_cast<uint32_t>((ballot(in_group && prov) >> base) & nmask);
      xc[lane] = stub;
      wave_sync();
      uint32_t K = in_group ? (1u << p) : 0u, done = 0;
      if (in_group) {
        while (K & ~done) {
          const int i = __ffs(K & ~done) - 1;
          done |= 1u << i;
          K |= xc[base + i];
        }
      }
      xf[lane] = (K & target) ? K : 0u;
      wave_sync();
      if (in_group) {
        int best = N;
        for (int q = 0; q < N; q++) {
          const uint32_t fq = xf[base + q];
          const int pc = __popc(fq);
          if (fq && pc < best) {
            best = pc;
            csel = fq;
          }
        }
      }
      wave_sync();
    } else {
      const uint32_t c = prov ? cluster0 : (cluster0 & ~(1u << p));  // cu:1635-1637
      xc[lane] = c;
      wave_sync();
      uint32_t K = c, done = 0;
      bool ok = false;
      if (in_group) {
        while (K & ~done) {  // closure (cu:1641-1649)
          const int i = __ffs(K & ~done) - 1;
          done |= 1u << i;
          const uint32_t ci = xc[base + i];
          if ((ci >> i) & 1u) ok = true;
          K |= ci;
        }
      }
      xf[lane] = ok ? K : 0u;  // cu:1651-1655
      wave_sync();
      if (in_group) {
        int best = N;
        for (int q = 0; q < N; q++) {  // first smallest non-empty cluster (cu:1657-1666)
          const uint32_t fq = xf[base + q];
          const int pc = __popc(fq);
          if (fq && pc < best) {
            best = pc;
            csel = fq;
          }
        }
      }
      wave_sync();
    }
  }

  // ---------------- pass 2: count (ample) successors
  uint64_t cnt = 0;
  if (has && ((csel >> p) & 1u)) cnt += l1 - l0;
  walk([&](uint32_t, uint64_t, bool) {},
       [&](uint32_t r, uint32_t) {
         if (has && (r & ~csel) == 0) cnt += rule_product(r);
       });
  uint64_t total = 0;
  const uint64_t off = wave_excl_scan(cnt, &total);
  uint64_t obase = 0;
  const int ok = reserve(total, &obase);
  if (ok <= 0) return ok;
  acc.viol |= viol;
  acc.trans += total;
  acc.states += __popcll(ballot(has && p == 0));

  // ---------------- pass 3: emit
  uint64_t pos = obase + off;
  if (has && ((csel >> p) & 1u)) {
    for (uint32_t t = l0; t < l1; t++) {
      uint64_t s[W];
#pragma unroll
      for (int j = 0; j < W; j++) s[j] = key[j];
      kset<W>(s, L.floc, M[H.off_loc_tgt + t]);
      emit(pos, s);
      pos++;
    }
  }
  walk([&](uint32_t, uint64_t, bool) {},
       [&](uint32_t r, uint32_t) {
         if (!has || (r & ~csel) != 0) return;
         // emitted products fit the candidate buffer, so < 2^32
         const uint32_t prod = static_cast<uint32_t>(rule_product(r));
         for (uint32_t k = 0; k < prod; k++) {
           uint64_t s[W];
#pragma unroll
           for (int j = 0; j < W; j++) s[j] = key[j];
           uint32_t rem = k;
           for (uint32_t rr = r; rr; rr &= rr - 1) {
             const int q = __ffs(rr) - 1;
             const uint2 t = xch[base + q];
             const uint32_t d = rem % t.y;
             rem /= t.y;
             kset<W>(s, M[H.off_floc + q], M[H.off_syn_tgt + t.x + d]);
           }
           emit(pos, s);
           pos++;
         }
       });
  return 1;
}

}  // namespace gx
