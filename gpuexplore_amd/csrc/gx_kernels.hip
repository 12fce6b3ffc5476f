// gx_kernels.hip -- gather/explore kernels for gfx950.
//
// One BFS level is three launches on one stream:
//   k_expand   frontier -> candidate successors   (cu:870-1151 gather,
//              cu:1306-1696 gather_por)
//   k_insert   candidates -> visited set + next frontier (cu:361-450
//              FINDORPUT_WARP, cu:707-735 publish, cu:804-869 claim)
//   k_advance  level bookkeeping and termination (cu:858-868, 1164-1178,
//              2109-2127)
// All sizes are read from the device control block, so levels can be
// enqueued back to back without host round trips.
#include "../../include/gpuexplore.h"
#include "gx_kernels.h"
#include "gx_group.h"
#include "gx_tps.h"

namespace gx {

// New keys staged per wave before a flush (insert).  Each flush is one
// atomicAdd on the next frontier's shared counter: 256 keys halve those
// same-address atomics against 128 (FIND-OR-PUT 95 vs 101 ms on a ring of
// 20 16-state processes); 384 no longer fits eight workgroups per CU in LDS.
constexpr int kStage = 256;

// adds v to counter `which` through the workgroup's partial-counter slot
__device__ __forceinline__ void add_counter(Ctrl *C, uint32_t which, uint64_t v) {
  if (v)
    atomicAdd(reinterpret_cast<unsigned long long *>(&C->part[(blockIdx.x % kCounterSlots) * 16u + which]),
              static_cast<unsigned long long>(v));
}

// ---------------------------------------------------------------------------
// Successor generation in the process-group layout (gx_group.h): G =
// floor(64/N) states per wave, one lane per process.  POR, the cycle proviso
// and models whose labels do not fit the state-per-lane kernel run here.  A
// wave whose output does not fit in the candidate buffer defers its states
// to a retry list.
template <int W, bool LDS, int MODE, int WT>
__global__ __launch_bounds__(512) void k_expand(ExpandArgs a) {
  extern __shared__ uint32_t smem[];
  const ModelHdr &H = a.hdr;
  Ctrl *C = a.ctrl;
  if (LDS) {
    stage_words(smem, a.model, H.total_words);
  }
  const uint32_t *M = LDS ? smem : a.model;
  const uint32_t xoff = LDS ? ((H.total_words + 3u) & ~3u) : 0u;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  GroupScratch X;
  X.xch = reinterpret_cast<uint2 *>(smem + xoff) + wib * 64;
  X.xc = smem + xoff + wpb * 128 + wib * 64;
  X.xf = smem + xoff + wpb * 192 + wib * 64;
  __syncthreads();
  if (C->stop) return;

  const uint32_t cur = C->cur;
  const uint32_t use_retry = C->use_retry;
  const uint32_t rc = C->retry_cur;
  const uint64_t nwork = use_retry ? C->retry_count[rc] : C->front_count[cur];
  const uint64_t *retry_in = a.retry[rc];
  uint64_t *retry_out = a.retry[rc ^ 1u];
  uint64_t *retry_out_count = &C->retry_count[rc ^ 1u];
  const uint64_t *front = a.front[cur];
  const uint32_t level = C->level;
  const GroupLane L = group_lane(M, H);
  const int G = L.G;
  const int lane = lane_id();

  SlotFmt<WT> fmt;
  make_fmt<WT, MODE == kModePorProviso>(fmt, H.state_bits, 0u);

  const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  GroupAcc acc;
  CandChunk ch(a.cand, a.cand_cap, &C->cand_count);

  for (uint64_t b0 = gw * G; b0 < nwork; b0 += nw * G) {
    const uint64_t si = b0 + L.g;
    const bool has = L.in_group && si < nwork;
    uint64_t fi = 0;
    if (has) fi = use_retry ? retry_in[si] : si;
    uint64_t key[W];
#pragma unroll
    for (int j = 0; j < W; j++) key[j] = has ? front[fi * W + j] : kFlag;
    const uint64_t left = (nwork - b0 + nw * G - 1) / (nw * G);
    const int ok = group_expand<W, MODE, WT>(
        M, H, L, X, key, has, fmt, a.table, a.bucket_words, level, a.property, a.por_method,
        [&](uint64_t total, uint64_t *obase) { return ch.reserve<W>(total, left, obase); },
        [&](uint64_t pos, const uint64_t (&sv)[W]) {
#pragma unroll
          for (int j = 0; j < W; j++) a.cand[pos * W + j] = sv[j];
        },
        acc);
    if (ok <= 0) {
      // defer this wave's states to the retry list
      if (ok < 0 && lane == 0) C->overflow = 2;
      const bool lead = has && L.p == 0;
      const uint64_t m = ballot(lead);
      uint64_t rb = 0;
      if (lane == 0) rb = atomicAdd(reinterpret_cast<unsigned long long *>(retry_out_count),
                                    static_cast<unsigned long long>(__popcll(m)));
      rb = __shfl(rb, 0, 64);
      if (lead) retry_out[rb + __popcll(m & ((1ull << lane) - 1ull))] = fi;
    }
  }
  ch.release_all<W>();
  if (lane == 0) {
    add_counter(C, kCntTransitions, acc.trans);
    add_counter(C, kCntExpanded, acc.states);
  }
  if (ballot(acc.viol) && lane == 0) C->violation = 1;
  if (ballot(acc.bad) && lane == 0) C->bad_state = 1;
}

// ---------------------------------------------------------------------------
// Successor generation, one state per lane (plain mode, labels < 64).
//
// For models whose per-process work is small (a few entries per local
// state) a lane expanding a whole state is cheaper than a group of N lanes:
// 64 states per wave instead of floor(64/N), and no cross-lane traffic
// except the wave's output reservation (gx_tps.h has the per-state code).
template <int W, bool LDS, typename RowT, bool PACKED = false>
__global__ __launch_bounds__(512) void k_expand_tps(ExpandArgs a) {
  extern __shared__ uint32_t smem[];
  // with LDS, only the state-per-lane section of the model is staged
  const ModelHdr H = LDS ? tps_rebased(a.hdr) : a.hdr;
  Ctrl *C = a.ctrl;
  const uint32_t sec = a.hdr.total_words - a.hdr.tps_begin;
  if (LDS) {
    stage_words(smem, a.model + a.hdr.tps_begin, sec);
  }
  const uint32_t *M = LDS ? smem : a.model;
  const uint32_t xoff = LDS ? ((sec + 3u) & ~3u) : 0u;
  const int N = static_cast<int>(H.nprocs);
  RowT *rows = reinterpret_cast<RowT *>(smem + xoff) + static_cast<uint32_t>(threadIdx.x >> 6) * (N * 64u);
  __syncthreads();
  if (C->stop) return;
  const uint32_t cur = C->cur;
  const uint32_t use_retry = C->use_retry;
  const uint32_t rc = C->retry_cur;
  const uint64_t nwork = use_retry ? C->retry_count[rc] : C->front_count[cur];
  const uint64_t *retry_in = a.retry[rc];
  uint64_t *retry_out = a.retry[rc ^ 1u];
  uint64_t *retry_out_count = &C->retry_count[rc ^ 1u];
  const uint64_t *front = a.front[cur];
  const int lane = lane_id();
  const bool safety = a.property == 2;
  const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint64_t acc_trans = 0, acc_states = 0;
  bool viol = false, bad = false;
  CandChunk ch(a.cand, a.cand_cap, &C->cand_count);
  ch.packed = PACKED;

  for (uint64_t b0 = gw * 64; b0 < nwork; b0 += nw * 64) {
    const uint64_t si = b0 + lane;
    const bool has = si < nwork;
    uint64_t fi = 0;
    if (has) fi = use_retry ? retry_in[si] : si;
    uint64_t key[W];
#pragma unroll
    for (int j = 0; j < W; j++) key[j] = has ? front[fi * W + j] : kFlag;
    const TpsInfo inf = tps_scan<W>(M, H, key, has, rows, lane, safety);
    viol |= inf.viol;
    bad |= inf.bad;
    if (a.property == 1 && inf.dead) viol = true;  // cu:1134-1151
    uint64_t total = 0;
    const uint64_t off = wave_excl_scan(inf.cnt, &total);
    const uint64_t left = (nwork - b0 + nw * 64 - 1) / (nw * 64);
    uint64_t obase = 0;
    const int ok = ch.reserve<W>(total, left, &obase);
    if (ok <= 0) {
      if (ok < 0 && lane == 0) C->overflow = 2;
      const uint64_t m = ballot(has);
      uint64_t rb = 0;
      if (lane == 0) rb = atomicAdd(reinterpret_cast<unsigned long long *>(retry_out_count),
                                    static_cast<unsigned long long>(__popcll(m)));
      rb = __shfl(rb, 0, 64);
      if (has) retry_out[rb + __popcll(m & ((1ull << lane) - 1ull))] = fi;
      continue;
    }
    acc_trans += total;
    acc_states += __popcll(ballot(has));
    if (has) {
      // Even lanes fill their range from its end. A line that a lane shares
      // with its right-hand neighbour is then written by both in their first
      // steps, and one shared with its left-hand neighbour in their last
      // steps, so no line stays partly written, and held in the L2, for a
      // whole batch (successor generation on config 2: 31.3 -> 27.7 ms).
      const bool rev = (lane & 1) == 0;
      uint64_t pos = rev ? obase + off + inf.cnt - 1 : obase + off;
      const uint64_t step = rev ? ~0ull : 1ull;  // -1 or +1, modulo 2^64
      tps_emit<W, true>(M, H, key, rows, lane, inf, [&](const uint64_t (&sv)[W]) {
        if constexpr (PACKED) {
          pack12(sv, reinterpret_cast<uint32_t *>(a.cand) + 3 * pos);
        } else {
#pragma unroll
          for (int j = 0; j < W; j++) a.cand[pos * W + j] = sv[j];
        }
        pos += step;
      });
    }
  }
  ch.release_all<W>();
  if (lane == 0) {
    add_counter(C, kCntTransitions, acc_trans);
    add_counter(C, kCntExpanded, acc_states);
  }
  if (ballot(viol) && lane == 0) C->violation = 1;
  if (ballot(bad) && lane == 0) C->bad_state = 1;
}

// Raw buffer loads (the llvm.amdgcn.raw.buffer.load intrinsic): a buffer
// resource is base address (48 bits), stride 0, size in bytes and the
// gfx9 format word 0x00020000; reads past the size return zeros. aux bit 1
// is the non-temporal hint.
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x3 __attribute__((ext_vector_type(3)));
__device__ i32x3 raw_buffer_load_v3(i32x4 rsrc, int voff, int soff, int aux)
    __asm("llvm.amdgcn.raw.buffer.load.v3i32");

// ---------------------------------------------------------------------------
// FIND-OR-PUT over the candidate buffer.  New keys are staged per wave in
// LDS and appended to the next frontier in blocks of up to kStage keys, one
// atomicAdd per block: a single contended counter would otherwise serialise
// all waves of the device.
//
// Eight resident waves per SIMD are asked for where slots allow it
// (kInsertMinWaves).

// OPT: kInsDedup (the duplicate filter below) and kInsSetFlag (keys from a
// caller: bit 63 of every word is set here) are compile-time options.  The
// default instantiation carries neither: as runtime branches they cost the
// plain FIND-OR-PUT registers (five VGPRs spilled to scratch) and 5-10 % of
// its time.
template <int W, int WT, int BW, bool STAMP, int OPT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kInsertMinWaves<WT, STAMP>)))
void k_insert(InsertArgs a) {
  constexpr bool DEDUP = (OPT & kInsDedup) != 0;
  constexpr bool PACKED = (OPT & kInsPacked) != 0;
  static_assert(!PACKED || (W == 2 && WT == 2 && !STAMP), "packed candidates hold 2-word keys");
  constexpr uint64_t fl = (OPT & kInsSetFlag) ? kFlag : 0ull;
  extern __shared__ uint64_t stage_all[];
  Ctrl *C = a.ctrl;
  if (C->stop) return;
  const uint64_t n = C->cand_count < a.cand_cap ? C->cand_count : a.cand_cap;
  const uint32_t nxt = C->cur ^ 1u;
  uint64_t *out = a.front[nxt];
  unsigned long long *out_count = reinterpret_cast<unsigned long long *>(&C->front_count[nxt]);
  SlotFmt<WT> fmt;
  make_fmt<WT, STAMP>(fmt, a.state_bits, C->level + 1u);
  constexpr int LPB = BW / 2;
  constexpr int CPW = 64 / LPB;
  static_assert(CPW <= kStage, "stage smaller than one iteration");
  const int lane = lane_id();
  const int sub = lane % LPB;
  const int cg = lane / LPB;
  const int gbase = cg * LPB;
  uint64_t *stage = stage_all + static_cast<size_t>(threadIdx.x >> 6) * (kStage * W);
  uint32_t *dtag = reinterpret_cast<uint32_t *>(stage_all + static_cast<size_t>(blockDim.x >> 6) * (kStage * W)) +
                   static_cast<size_t>(threadIdx.x >> 6) * kDedupSlots;
  uint32_t staged = 0;  // wave-uniform
  const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint64_t acc_new = 0, acc_look = 0;
  bool full = false, timeout = false, overflow = false;
  auto flush = [&]() {
    if (!staged) return;
    wave_sync();
    uint64_t ob = 0;
    if (lane == 0) ob = atomicAdd(out_count, static_cast<unsigned long long>(staged));
    ob = __shfl(ob, 0, 64);
    for (uint32_t i = lane; i < staged * W; i += 64) {
      if (ob + i / W < a.front_cap)
        out[ob * W + i] = stage[i];
      else
        overflow = true;
    }
    wave_sync();
    staged = 0;
  };
  // One iteration: group cg looks up candidate ci (key, its hash h and
  // first bucket b0).
  auto visit = [&](uint64_t ci, const uint64_t (&key)[W], uint64_t h, uint64_t b0, bool active) {
    uint64_t v[WT];
    slot_value<W, WT>(key, fmt, v);
    const int st = find_or_put_group_b<WT, BW>(a.table, v, fmt, b0, [h]() { return h; }, sub, gbase, active);
    acc_look += __popcll(ballot(active && sub == 0));
    const bool is_new = active && st == kInserted;
    full |= active && st == kFull;
    timeout |= active && st == kTimeout;
    if (a.is_new && active && sub == 0) a.is_new[ci] = is_new ? 1 : 0;
    const uint64_t m = ballot(is_new && sub == 0);
    if (!m) return;
    const uint32_t cnt = __popcll(m);
    acc_new += cnt;
    if (!a.append) return;
    if (staged + cnt > static_cast<uint32_t>(kStage)) flush();
    if (is_new) {
      const uint32_t pos = staged + __popcll(m & ((1ull << gbase) - 1ull));
#pragma unroll
      for (int j = 0; j < W; j++)
        if (j % LPB == sub) stage[pos * W + j] = key[j];
    }
    staged += cnt;
  };
  // Each wave walks runs of consecutive iterations.  Successors of one
  // state, and of states found next to each other, are adjacent in the
  // candidate buffer, so the run keeps them on one wave: its new keys reach
  // the next frontier in candidate order (siblings stay together), and the
  // duplicates that commuting moves produce are looked up close in time.
  const uint64_t iters = insert_run(n, nw * CPW);
  if (iters < kInsertSpread) {
    // a pass spread over every wave: one iteration of CPW candidates at a time
    const uint64_t RUN = CPW * iters;
    for (uint64_t r0 = gw * RUN, c0 = r0; c0 < n;) {
      const uint64_t ci = c0 + cg;
      c0 += CPW;
      if (c0 == r0 + RUN) {
        r0 += nw * RUN;
        c0 = r0;
      }
      uint64_t key[W];
      if constexpr (PACKED) {
        const uint32_t *c32 = reinterpret_cast<const uint32_t *>(a.cand) + 3 * ci;
        const uint32_t u2 = ci < n ? __builtin_nontemporal_load(c32 + 2) : 0u;
        const uint32_t u0 = ci < n ? __builtin_nontemporal_load(c32) : 0u;
        const uint32_t u1 = ci < n ? __builtin_nontemporal_load(c32 + 1) : 0u;
        unpack12(u0, u1, u2, key);
        if (!(u2 >> 31)) key[0] = 0;
      } else {
#pragma unroll
        for (int j = 0; j < W; j++)  // read once: keep the table's lines in the caches
          key[j] = ci < n ? (__builtin_nontemporal_load(a.cand + ci * W + j) | fl) : 0;
      }
      // a zero first word: past the end or a released reservation
      const uint64_t h = key_hash<W>(key, a.table.seed);
      visit(ci, key, h, bucket_of(h, 0, a.table.nbuckets), key[0] != 0);
    }
  } else {
    // Runs of kInsertRun iterations in steps of 64 candidates: each lane
    // reads and hashes one candidate, and the step's LPB iterations hand
    // key, hash and first bucket to the groups in candidate order, so a
    // candidate is hashed once rather than by every lane of its group (the
    // hash is handed over too: recomputed for the later probes, it would be
    // hoisted out of the probe loop and computed for every candidate).
    static_assert(kInsertRun % LPB == 0, "a run is a whole number of steps");
    const uint64_t RUN = CPW * kInsertRun;
    // Duplicate filter (kInsDedup): the successors of commuting moves repeat
    // within a few dozen candidates, so a candidate equal to another of the
    // same step, or of the wave's previous step, is not looked up -- the
    // other lane's lookup (this step) or the finished one (last step) stands
    // for it.  Lanes tag an LDS slot chosen by the key hash with
    // (step << 6 | lane) and compare keys through cross-lane reads, so only
    // equal keys are dropped; slot collisions of different keys just look
    // both up.
    uint32_t sid = 1;  // step of this wave, wave-uniform
    uint64_t key_prev[W];
#pragma unroll
    for (int j = 0; j < W; j++) key_prev[j] = 0;
    if (DEDUP) {
      for (uint32_t i = lane; i < kDedupSlots; i += 64) dtag[i] = 0;
      wave_sync();
    }
    for (uint64_t r0 = gw * RUN, c0 = r0; c0 < n;) {
      const uint64_t cl = c0 + lane;
      const uint64_t cstep = c0;
      c0 += 64;
      if (c0 == r0 + RUN) {
        r0 += nw * RUN;
        c0 = r0;
      }
      uint64_t key_l[W];
      if constexpr (PACKED) {
        // the step's 64 candidates through one buffer resource whose range
        // ends at the buffer's fill: one 12-byte non-temporal load per lane,
        // zeros past the end
        const uint64_t base = reinterpret_cast<uint64_t>(a.cand) + 12ull * cstep;
        const uint64_t left = n > cstep ? n - cstep : 0;
        const uint32_t nrec = static_cast<uint32_t>(12ull * (left < 64 ? left : 64));
        i32x4 rsrc;
        rsrc.x = static_cast<int>(static_cast<uint32_t>(base));
        rsrc.y = static_cast<int>(static_cast<uint32_t>(base >> 32) & 0xFFFFu);
        rsrc.z = static_cast<int>(nrec);
        rsrc.w = 0x00020000;
        const i32x3 q = raw_buffer_load_v3(rsrc, lane * 12, 0, 2);
        const uint32_t u0 = static_cast<uint32_t>(q.x), u1 = static_cast<uint32_t>(q.y),
                       u2 = static_cast<uint32_t>(q.z);
        unpack12(u0, u1, u2, key_l);
        if (!(u2 >> 31)) key_l[0] = 0;
      } else {
#pragma unroll
        for (int j = 0; j < W; j++) key_l[j] = cl < n ? (__builtin_nontemporal_load(a.cand + cl * W + j) | fl) : 0;
      }
      uint64_t act_m = ballot(key_l[0] != 0);
      const uint64_t h_l = key_hash<W>(key_l, a.table.seed);
      const uint64_t b_l = bucket_of(h_l, 0, a.table.nbuckets);
      if (DEDUP) {
        const bool act = (act_m >> lane) & 1ull;
        const uint32_t slot = static_cast<uint32_t>(h_l) & (kDedupSlots - 1u);
        auto equal_to = [&](const uint64_t (&other)[W], int src) {
          bool eq = true;
#pragma unroll
          for (int j = 0; j < W; j++) eq = eq && (__shfl(other[j], src, 64) == key_l[j]);
          return eq;
        };
        // a lane of the previous step with this slot and an equal key
        const uint32_t old = dtag[slot];
        bool dup = equal_to(key_prev, static_cast<int>(old & 63u)) && act && (old >> 6) == sid - 1u;
        wave_sync();
        if (act && !dup) dtag[slot] = (sid << 6) | static_cast<uint32_t>(lane);
        wave_sync();
        // the lane of this step that holds the slot, when its key is equal
        const uint32_t cur = dtag[slot];
        const bool eq_cur = equal_to(key_l, static_cast<int>(cur & 63u));
        dup = dup || (act && (cur >> 6) == sid && static_cast<int>(cur & 63u) != lane && eq_cur);
        act_m = ballot(act && !dup);
#pragma unroll
        for (int j = 0; j < W; j++) key_prev[j] = key_l[j];
        sid++;
      }
#pragma unroll 1
      for (int s = 0; s < LPB; s++) {
        const int src = s * CPW + cg;
        uint64_t key[W];
#pragma unroll
        for (int j = 0; j < W; j++) key[j] = __shfl(key_l[j], src, 64);
        visit(cstep + src, key, __shfl(h_l, src, 64), __shfl(b_l, src, 64), (act_m >> src) & 1ull);
      }
    }
  }
  if (a.append) {
    // The waves' last stages share one reservation per workgroup: one
    // same-address atomic instead of one per wave.  The barriers cost no
    // occupancy, since a workgroup's LDS is held until its last wave ends.
    __shared__ unsigned long long wg_count, wg_base;
    if (threadIdx.x == 0) wg_count = 0;
    __syncthreads();
    unsigned long long off = 0;
    if (lane == 0 && staged) off = atomicAdd(&wg_count, static_cast<unsigned long long>(staged));
    off = __shfl(off, 0, 64);
    __syncthreads();
    if (threadIdx.x == 0) wg_base = wg_count ? atomicAdd(out_count, wg_count) : 0ull;
    __syncthreads();
    const uint64_t ob = wg_base + off;
    for (uint32_t i = lane; i < staged * W; i += 64) {
      if (ob + i / W < a.front_cap)
        out[ob * W + i] = stage[i];
      else
        overflow = true;
    }
  }
  if (lane == 0) {
    add_counter(C, kCntInserts, acc_new);
    add_counter(C, kCntLookups, acc_look);
  }
  if (ballot(full) && lane == 0) C->table_full = 1;
  if (ballot(timeout) && lane == 0) C->spin_timeout = 1;
  if (ballot(overflow) && lane == 0) C->overflow = 1;
}

// ---------------------------------------------------------------------------
// Level bookkeeping (one thread).  A level is complete when no states were
// deferred; then the frontiers swap.  Exploration stops after a level with a
// violation (level-granular early stop), on a full table or a fatal error,
// or when the next frontier is empty.
// Folds the partial-counter slots into the counters and clears them.  One
// wave of kCounterSlots lanes, lane i taking slot i.
__device__ __forceinline__ void fold_counters(Ctrl *c) {
  const uint32_t i = threadIdx.x;
  uint64_t *slot = c->part + i * 16u;
  uint64_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = slot[k];
    slot[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int d = kCounterSlots / 2; d > 0; d >>= 1) v[k] += __shfl_xor(v[k], d, kCounterSlots);
  if (i == 0) {
    c->inserts += v[kCntInserts];
    c->lookups += v[kCntLookups];
    c->transitions += v[kCntTransitions];
    c->expanded += v[kCntExpanded];
  }
}

__global__ void k_fold(Ctrl *c) { fold_counters(c); }

// zeroes the control block (head and partial counters) and sets cand_count
__global__ void k_ctrl_reset(Ctrl *c, uint64_t n) {
  uint64_t *w = reinterpret_cast<uint64_t *>(c);
  constexpr uint32_t nw = sizeof(Ctrl) / 8;
  for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) w[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) c->cand_count = n;
}

__global__ void k_advance(Ctrl *c) {
  fold_counters(c);
  if (threadIdx.x != 0) return;
  if (c->stop) return;
  c->cand_count = 0;
  if (c->overflow || c->bad_state || c->spin_timeout || c->table_full) {
    c->stop = 1;
    return;
  }
  const uint32_t rc = c->retry_cur;
  if (c->retry_count[rc ^ 1u] > 0) {
    c->retry_count[rc] = 0;
    c->retry_cur = rc ^ 1u;
    c->use_retry = 1;
    return;
  }
  c->retry_count[rc] = 0;
  c->use_retry = 0;
  c->front_count[c->cur] = 0;
  c->cur ^= 1u;
  c->level += 1u;
  c->levels_done += 1u;
  if (c->violation) c->stop = 1;
  if (c->max_states && c->inserts >= c->max_states && c->front_count[c->cur] != 0) {
    c->truncated = 1;
    c->stop = 1;
  }
  if (c->front_count[c->cur] == 0) {
    c->done = 1;
    c->stop = 1;
  }
}

// occupied slots (cu:687-704)
__global__ void k_count(Table t, int WT, int BW, unsigned long long *out) {
  __shared__ unsigned long long acc;
  if (threadIdx.x == 0) acc = 0;
  __syncthreads();
  const int SPB = BW / WT;
  unsigned long long mine = 0;
  const uint64_t nslots = t.nbuckets * static_cast<uint64_t>(SPB);
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nslots;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t b = i / SPB, s = i % SPB;
    mine += t.words[b * BW + s * WT] != 0;
  }
  atomicAdd(&acc, mine);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, acc);
}

// ---------------------------------------------------------------------------
size_t expand_scratch_bytes(uint32_t threads) {
  // per wave: 64 x uint2 exchange + 64 x uint32 clusters + 64 x uint32 finals
  return static_cast<size_t>(threads / 64) * kGroupScratchBytesPerWave + 16;
}

template <int W, int WT>
static hipError_t expand_w(const ExpandArgs &a, int mode, bool lds, uint32_t blocks,
                           uint32_t threads, hipStream_t s) {
  const size_t model_b = lds ? ((a.hdr.total_words + 3u) & ~3u) * 4u : 0u;
  if (mode == kModePlain && a.per_state) {
    const size_t sec_b = lds ? ((a.hdr.total_words - a.hdr.tps_begin + 3u) & ~3u) * 4u : 0u;
    const bool r16 = a.hdr.nrows < 0xFFFFu;
    const size_t shm = sec_b + expand_tps_scratch_bytes(threads, a.hdr.nprocs, r16);
    if constexpr (W == 2 && WT == 2) {
      if (a.packed) {
        if (lds && r16)
          hipLaunchKernelGGL((k_expand_tps<W, true, uint16_t, true>), dim3(blocks), dim3(threads), shm, s, a);
        else if (lds)
          hipLaunchKernelGGL((k_expand_tps<W, true, uint32_t, true>), dim3(blocks), dim3(threads), shm, s, a);
        else if (r16)
          hipLaunchKernelGGL((k_expand_tps<W, false, uint16_t, true>), dim3(blocks), dim3(threads), shm, s, a);
        else
          hipLaunchKernelGGL((k_expand_tps<W, false, uint32_t, true>), dim3(blocks), dim3(threads), shm, s, a);
        return hipGetLastError();
      }
    }
    if (a.packed) return hipErrorInvalidValue;
    if (lds && r16)
      hipLaunchKernelGGL((k_expand_tps<W, true, uint16_t>), dim3(blocks), dim3(threads), shm, s, a);
    else if (lds)
      hipLaunchKernelGGL((k_expand_tps<W, true, uint32_t>), dim3(blocks), dim3(threads), shm, s, a);
    else if (r16)
      hipLaunchKernelGGL((k_expand_tps<W, false, uint16_t>), dim3(blocks), dim3(threads), shm, s, a);
    else
      hipLaunchKernelGGL((k_expand_tps<W, false, uint32_t>), dim3(blocks), dim3(threads), shm, s, a);
    return hipGetLastError();
  }
  const size_t shm = model_b + expand_scratch_bytes(threads);
  if (mode == kModePlain) {
    if (lds)
      hipLaunchKernelGGL((k_expand<W, true, kModePlain, W>), dim3(blocks), dim3(threads), shm, s, a);
    else
      hipLaunchKernelGGL((k_expand<W, false, kModePlain, W>), dim3(blocks), dim3(threads), shm, s, a);
  } else if (mode == kModePor) {
    if (lds)
      hipLaunchKernelGGL((k_expand<W, true, kModePor, W>), dim3(blocks), dim3(threads), shm, s, a);
    else
      hipLaunchKernelGGL((k_expand<W, false, kModePor, W>), dim3(blocks), dim3(threads), shm, s, a);
  } else {
    if (lds)
      hipLaunchKernelGGL((k_expand<W, true, kModePorProviso, WT>), dim3(blocks), dim3(threads), shm, s, a);
    else
      hipLaunchKernelGGL((k_expand<W, false, kModePorProviso, WT>), dim3(blocks), dim3(threads), shm, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_expand(const ExpandArgs &a, int mode, int key_words, int slot_words, bool lds,
                         uint32_t blocks, uint32_t threads, hipStream_t s) {
  const bool plus = slot_words == key_words + 1;
  if (slot_words != key_words && !plus) return hipErrorInvalidValue;
  switch (key_words) {
    case 1: return plus ? expand_w<1, 2>(a, mode, lds, blocks, threads, s) : expand_w<1, 1>(a, mode, lds, blocks, threads, s);
    case 2: return plus ? expand_w<2, 3>(a, mode, lds, blocks, threads, s) : expand_w<2, 2>(a, mode, lds, blocks, threads, s);
    case 3: return plus ? expand_w<3, 4>(a, mode, lds, blocks, threads, s) : expand_w<3, 3>(a, mode, lds, blocks, threads, s);
    case 4: return plus ? expand_w<4, 5>(a, mode, lds, blocks, threads, s) : expand_w<4, 4>(a, mode, lds, blocks, threads, s);
    case 5: return plus ? hipErrorInvalidValue : expand_w<5, 5>(a, mode, lds, blocks, threads, s);
    default: return hipErrorInvalidValue;
  }
}

template <int W, int WT, bool STAMP, int OPT>
static hipError_t insert_bw_opt(const InsertArgs &a, int bw, uint32_t blocks, uint32_t threads, hipStream_t s) {
  const size_t shm = static_cast<size_t>(threads / 64) * (kStage * W * 8 + ((OPT & kInsDedup) ? kDedupSlots * 4 : 0));
  switch (bw) {
    case 8:
      if (WT <= 8) { hipLaunchKernelGGL((k_insert<W, WT, 8, STAMP, OPT>), dim3(blocks), dim3(threads), shm, s, a); break; }
      return hipErrorInvalidValue;
    case 16: hipLaunchKernelGGL((k_insert<W, WT, 16, STAMP, OPT>), dim3(blocks), dim3(threads), shm, s, a); break;
    case 32: hipLaunchKernelGGL((k_insert<W, WT, 32, STAMP, OPT>), dim3(blocks), dim3(threads), shm, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// the filter runs in explorations (every mode); caller keys (set_flag) come
// only from the visited-set API, whose slots are the keys themselves
template <int W, int WT, bool STAMP>
static hipError_t insert_bw(const InsertArgs &a, int bw, uint32_t blocks, uint32_t threads, hipStream_t s) {
  if (a.set_flag) {
    if constexpr (W == WT && !STAMP) return insert_bw_opt<W, WT, STAMP, kInsSetFlag>(a, bw, blocks, threads, s);
    return hipErrorInvalidValue;
  }
  if (a.packed) {
    if constexpr (W == 2 && WT == 2 && !STAMP) {
      if (!a.dedup && a.append && !a.is_new) return insert_bw_opt<W, WT, STAMP, kInsPacked>(a, bw, blocks, threads, s);
    }
    return hipErrorInvalidValue;
  }
  if (a.dedup) return insert_bw_opt<W, WT, STAMP, kInsDedup>(a, bw, blocks, threads, s);
  return insert_bw_opt<W, WT, STAMP, 0>(a, bw, blocks, threads, s);
}

hipError_t launch_insert(const InsertArgs &a, int key_words, int slot_words, int bucket_words,
                         bool stamp, uint32_t blocks, uint32_t threads, hipStream_t s) {
  const bool plus = slot_words == key_words + 1;
  if (!stamp && plus) return hipErrorInvalidValue;
  if (slot_words != key_words && !plus) return hipErrorInvalidValue;
#define GX_INS(Wk, WTk)                                                            \
  return stamp ? insert_bw<Wk, WTk, true>(a, bucket_words, blocks, threads, s)     \
               : insert_bw<Wk, WTk, false>(a, bucket_words, blocks, threads, s)
  switch (key_words) {
    case 1: if (plus) { GX_INS(1, 2); } else { GX_INS(1, 1); }
    case 2: if (plus) { GX_INS(2, 3); } else { GX_INS(2, 2); }
    case 3: if (plus) { GX_INS(3, 4); } else { GX_INS(3, 3); }
    case 4: if (plus) { GX_INS(4, 5); } else { GX_INS(4, 4); }
    case 5: if (plus) return hipErrorInvalidValue; else { GX_INS(5, 5); }
    default: return hipErrorInvalidValue;
  }
#undef GX_INS
}

// Measurement kernels: the attainable streaming copy and read bandwidth, and
// the rate of random 64-byte bucket reads (the FIND-OR-PUT's access), each
// with known bytes per launch, so the hardware counters' byte accounting can
// be checked against them.
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

// Four independent 16-byte loads in flight per thread; NT: non-temporal
// (streaming) hints on loads and stores.
template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  const v4u32 *s = reinterpret_cast<const v4u32 *>(src);
  v4u32 *d = reinterpret_cast<v4u32 *>(dst);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    v4u32 x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = NT ? __builtin_nontemporal_load(s + i + k * stride) : s[i + k * stride];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (NT)
        __builtin_nontemporal_store(x[k], d + i + k * stride);
      else
        d[i + k * stride] = x[k];
    }
  }
  for (; i < n; i += stride) {
    if (NT)
      __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
    else
      d[i] = s[i];
  }
}

// Streaming read of n 16-byte elements; the XOR of everything goes to *sink
// only if it equals a value no input produces in practice (keeps the loads).
__global__ __launch_bounds__(256) void k_read(const uint4 *__restrict__ src, uint64_t n, uint32_t *sink) {
  const v4u32 *s = reinterpret_cast<const v4u32 *>(src);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  v4u32 acc = {0, 0, 0, 0};
  for (; i + 3 * stride < n; i += 4 * stride) {
    v4u32 x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = __builtin_nontemporal_load(s + i + k * stride);
#pragma unroll
    for (int k = 0; k < 4; k++) acc ^= x[k];
  }
  for (; i < n; i += stride) acc ^= __builtin_nontemporal_load(s + i);
  const uint32_t r = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (r == 0x9e3779b9u) *sink = r;
}

// n random 64-byte bucket reads from a table of nbuckets buckets, four lanes
// per bucket (16 bytes each), as the FIND-OR-PUT's lane groups read them;
// bucket of lookup i: the FIND-OR-PUT's bucket_of on a hash of i.
__global__ __launch_bounds__(256) void k_random_read(const uint64_t *__restrict__ table, uint64_t nbuckets,
                                                      uint64_t n, uint64_t seed, uint32_t *sink) {
  const int lane = lane_id();
  const int sub = lane & 3;
  const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint64_t acc = 0;
  for (uint64_t j = gw * 16 + (lane >> 2); j < n; j += nw * 16) {
    const uint64_t b = bucket_of(mix64(j ^ seed), 0, nbuckets);
    const ulonglong2 q = *reinterpret_cast<const ulonglong2 *>(table + b * 8 + 2 * sub);
    acc ^= q.x + q.y;
  }
  if (static_cast<uint32_t>(acc) == 0x9e3779b9u) *sink = static_cast<uint32_t>(acc >> 32);
}

hipError_t launch_copy(const uint4 *src, uint4 *dst, uint64_t n, uint32_t blocks, bool nt, hipStream_t s) {
  if (nt)
    hipLaunchKernelGGL(k_copy<true>, dim3(blocks), dim3(256), 0, s, src, dst, n);
  else
    hipLaunchKernelGGL(k_copy<false>, dim3(blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_read(const uint4 *src, uint64_t n, uint32_t *sink, uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_read, dim3(blocks), dim3(256), 0, s, src, n, sink);
  return hipGetLastError();
}

hipError_t launch_random_read(const uint64_t *table, uint64_t nbuckets, uint64_t n, uint64_t seed, uint32_t *sink,
                              uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_random_read, dim3(blocks), dim3(256), 0, s, table, nbuckets, n, seed, sink);
  return hipGetLastError();
}

hipError_t launch_advance(Ctrl *c, hipStream_t s) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(kCounterSlots), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_ctrl_reset(Ctrl *c, uint64_t n, hipStream_t s) {
  static_assert(sizeof(Ctrl) % 8 == 0, "control block of whole words");
  hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(256), 0, s, c, n);
  return hipGetLastError();
}

hipError_t launch_fold(Ctrl *c, hipStream_t s) {
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(kCounterSlots), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_count(const Table &t, int slot_words, int bucket_words, uint64_t *dev_count,
                        uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_count, dim3(blocks), dim3(256), 0, s, t, slot_words, bucket_words,
                     reinterpret_cast<unsigned long long *>(dev_count));
  return hipGetLastError();
}

}  // namespace gx
