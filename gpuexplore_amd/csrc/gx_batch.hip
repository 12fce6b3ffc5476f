// gx_batch.hip -- many independent models explored together.
//
// Every instance owns a partition of one visited set and two frontier
// regions; the instances advance one BFS level per round, in lockstep:
//   k_batch_expand   64-state chunks of one instance per wave (the model is
//                    wave-uniform), state-per-lane successor generation
//   k_batch_insert   FIND-OR-PUT into the candidate's instance partition
//   k_batch_advance  per-instance level bookkeeping and termination, then
//                    the chunk prefix of the next round (one workgroup)
// A round reads all its sizes from device memory, so `levels_per_sync`
// rounds are captured once in a hipGraph and replayed until every instance
// has stopped.  Semantics per instance are those of gx_explore (plain
// method with an optional deadlock or safety check).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <vector>

#include "../../include/gpuexplore.h"
#include "gx_kernels.h"
#include "gx_model.h"
#include "gx_group.h"
#include "gx_tps.h"
#include "gx_batch.h"
#include "gx_devbuf.h"

namespace gx {

// last instance whose chunk prefix (plain or POR) is <= c: instances without
// chunks of that kind share the prefix of the next one, so the search lands on
// the instance that owns chunk c
template <bool POR>
__device__ __forceinline__ uint32_t find_instance(const BatchInst *I, uint32_t n, uint64_t c) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((POR ? I[mid].chunk_prefix_por : I[mid].chunk_prefix) <= c)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

constexpr uint32_t kNoInst = 0xFFFFFFFFu;

// Transitions and expanded states of the wave's chunks, added to their
// instance's counters when the wave moves on to another instance (one
// atomic per switch, not per chunk: on a single large instance every wave
// of the device would otherwise add into the same two words per chunk).
struct InstAcc {
  uint32_t i = kNoInst;
  uint64_t trans = 0, expanded = 0;
  __device__ void flush(BatchInst *inst) {
    if (i != kNoInst && lane_id() == 0 && (trans || expanded)) {
      atomicAdd(reinterpret_cast<unsigned long long *>(&inst[i].transitions), static_cast<unsigned long long>(trans));
      atomicAdd(reinterpret_cast<unsigned long long *>(&inst[i].expanded), static_cast<unsigned long long>(expanded));
    }
    trans = expanded = 0;
  }
  __device__ void add(BatchInst *inst, uint32_t j, uint64_t t, uint64_t e) {
    if (j != i) {
      flush(inst);
      i = j;
    }
    trans += t;
    expanded += e;
  }
};

// A wave's share of the candidate buffer in the batch kernels: the
// single-model reservation (CandChunk, whole 64-entry blocks from the shared
// counter) where a wave also moves to a fresh block whenever the instance it
// emits for changes, so every block holds successors of one instance and
// cand_blk[block] names it: one id per 64 candidates instead of one per
// candidate, and the FIND-OR-PUT's 64-candidate steps read one id.
struct BlockChunk {
  CandChunk ch;
  uint32_t inst = 0xFFFFFFFFu;
  __device__ BlockChunk(uint64_t *c, uint64_t cap, uint64_t *cnt) : ch(c, cap, cnt) {}
  template <int W>
  __device__ int reserve(uint32_t *cand_blk, uint32_t i, uint64_t total, uint64_t batches, uint64_t *at) {
    if (!total) return 1;
    if (i != inst && ch.left) {  // skip to the next block boundary, releasing the gap
      const uint64_t gap0 = ((ch.base + 63) & ~63ull) - ch.base;
      const uint64_t gap = gap0 < ch.left ? gap0 : ch.left;
      ch.release<W>(ch.base, gap);
      ch.base += gap;
      ch.left -= gap;
    }
    inst = i;
    const int ok = ch.reserve<W>(total, batches, at);
    if (ok == 1) {
      const uint64_t b1 = (*at + total - 1) >> 6;
      for (uint64_t b = (*at >> 6) + lane_id(); b <= b1; b += 64) cand_blk[b] = i;
    }
    return ok;
  }
  template <int W>
  __device__ void release_all() {
    ch.release_all<W>();
  }
};

// One 64-state chunk of a plain instance by one wave, one state per lane
// (the state-per-lane code of k_expand_tps).  M/H: the model (staged in LDS
// or global), rows: the wave's row table.  Returns false if the chunk is
// deferred (the candidate buffer is full) or dropped (it never fits).
template <int W, typename RowT>
__device__ __forceinline__ void batch_chunk(const BatchArgs &a, BatchCtrl *C, BlockChunk &ch, InstAcc &acc,
                                            uint32_t i, uint64_t start, const uint32_t *M, const ModelHdr &H,
                                            RowT *rows, uint64_t left, uint64_t *def_out, uint64_t *def_out_count) {
  const int lane = lane_id();
  BatchInst *I = &a.inst[i];
  const uint32_t cur = I->cur;
  const uint64_t si = start + lane;
  const bool has = si < I->front_count[cur];
  if (!ballot(has)) return;
  uint64_t key[W];
#pragma unroll
  for (int j = 0; j < W; j++) key[j] = has ? a.front[cur][(I->fo + si) * W + j] : kFlag;
  const TpsInfo inf = tps_scan<W>(M, H, key, has, rows, lane, I->property == 2);
  const bool viol = inf.viol || (I->property == 1 && inf.dead);
  if (ballot(viol) && lane == 0) I->violation = 1;
  if (ballot(inf.bad) && lane == 0) I->bad_state = 1;
  uint64_t total = 0;
  const uint64_t off = wave_excl_scan(inf.cnt, &total);
  uint64_t obase = 0;
  const int ok = ch.reserve<W>(a.cand_blk, i, total, left, &obase);
  if (ok < 0) {
    // this chunk can never fit: stop its instance (at the level's end),
    // not the batch, and do not defer the chunk
    if (lane == 0) {
      I->cand_overflow = 1;
      C->cand_overflow = 1;
    }
    return;
  }
  if (ok == 0) {
    if (lane == 0) {
      const uint64_t at = atomicAdd(reinterpret_cast<unsigned long long *>(def_out_count), 1ull);
      def_out[at] = (static_cast<uint64_t>(i) << 40) | start;
    }
    return;
  }
  acc.add(a.inst, i, total, static_cast<uint64_t>(__popcll(ballot(has))));
  if (has) {
    // even lanes fill their range from its end, as in k_expand_tps: lines
    // shared by neighbouring lanes complete within a few steps
    const bool rev = (lane & 1) == 0;
    uint64_t pos = rev ? obase + off + inf.cnt - 1 : obase + off;
    const uint64_t step = rev ? ~0ull : 1ull;  // -1 or +1, modulo 2^64
    tps_emit<W>(M, H, key, rows, lane, inf, [&](const uint64_t (&sv)[W]) {
#pragma unroll
      for (int j = 0; j < W; j++) a.cand[pos * W + j] = sv[j];
      pos += step;
    });
  }
}

// Plain instances.  A round's plain work is cut into workgroup chunks of
// waves x 64 states of one instance (a.chunk_states), so all waves of a
// workgroup expand the same model: the workgroup stages that model's
// state-per-lane section into LDS when its instance changes (when it fits
// beside the row tables), as the single-model kernel does, instead of
// gathering it through the caches lane by lane.  Deferred chunks (64
// states, from a round whose candidate buffer filled) are expanded one per
// wave with the model read from global memory.
template <int W, typename RowT>
__global__ __launch_bounds__(512) void k_batch_expand(BatchArgs a) {
  extern __shared__ uint32_t smem[];
  BatchCtrl *C = a.ctrl;
  if (C->active == 0) return;
  const uint32_t n = C->n;
  const uint32_t use_def = C->use_deferred;
  const uint32_t dc = C->deferred_cur;
  const uint64_t *def_in = a.deferred[dc];
  uint64_t *def_out = a.deferred[dc ^ 1u];
  uint64_t *def_out_count = &C->deferred_count[dc ^ 1u];
  const uint32_t wib = threadIdx.x >> 6;
  const uint32_t wpb = blockDim.x >> 6;
  BlockChunk ch(a.cand, a.cand_cap, &C->cand_count);
  InstAcc acc;
  if (use_def) {
    const uint64_t nchunks = C->deferred_count[dc];
    const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
    RowT *rows = reinterpret_cast<RowT *>(smem) + wib * (a.max_procs * 64u);
    for (uint64_t c = gw; c < nchunks; c += nw) {
      const uint64_t d = def_in[c];
      const uint32_t i = static_cast<uint32_t>(d >> 40);
      if (a.inst[i].mode != kModePlain) continue;  // the POR kernel's
      const ModelHdr H = a.inst[i].hdr;
      batch_chunk<W, RowT>(a, C, ch, acc, i, d & ((1ull << 40) - 1), a.models + a.inst[i].model_off, H, rows,
                           (nchunks - c + nw - 1) / nw, def_out, def_out_count);
    }
  } else {
    const uint64_t nchunks = C->total_chunks;
    uint32_t staged = kNoInst;
    uint32_t i = kNoInst;
    uint64_t lo = 0, hi = 0;  // chunks [lo, hi) belong to instance i
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {  // workgroup-uniform
      if (c < lo || c >= hi) {  // the chunk left the current instance: search
        i = find_instance<false>(a.inst, n, c);
        lo = a.inst[i].chunk_prefix;
        hi = lo + (a.inst[i].front_count[a.inst[i].cur] + a.chunk_states - 1) / a.chunk_states;
      }
      const BatchInst *I = &a.inst[i];
      const ModelHdr Hg = I->hdr;
      const uint32_t *Mg = a.models + I->model_off;
      const uint32_t sec = Hg.total_words - Hg.tps_begin;
      const uint32_t sec_w = (sec + 3u) & ~3u;
      const size_t rows_b = static_cast<size_t>(wpb) * Hg.nprocs * 64u * sizeof(RowT);
      const bool lds = static_cast<size_t>(sec_w) * 4u + rows_b <= a.expand_lds;
      // LDS layout: the staged section and N-process row tables behind it,
      // or (model not staged) row tables of max_procs processes per wave.
      // Every wave must be done with the old layout before it changes.
      const uint32_t layout = lds ? i : kNoInst;
      if (layout != staged) {
        __syncthreads();
        if (lds) stage_words(smem, Mg + Hg.tps_begin, sec);
        __syncthreads();
        staged = layout;
      }
      // two call sites, so that each knows its model's address space (a
      // pointer that may be either would make every model read a FLAT access)
      const uint64_t start = (c - I->chunk_prefix) * a.chunk_states + wib * 64u;
      const uint64_t left = (nchunks - c + gridDim.x - 1) / gridDim.x;
      if (lds)
        batch_chunk<W, RowT>(a, C, ch, acc, i, start, smem, tps_rebased(Hg),
                             reinterpret_cast<RowT *>(smem + sec_w) + wib * (Hg.nprocs * 64u), left, def_out,
                             def_out_count);
      else
        batch_chunk<W, RowT>(a, C, ch, acc, i, start, Mg, Hg,
                             reinterpret_cast<RowT *>(smem) + wib * (a.max_procs * 64u), left, def_out,
                             def_out_count);
    }
  }
  acc.flush(a.inst);
  ch.release_all<W>();
}

// The slot format of an instance: with the cycle proviso a 32-bit level
// stamp (the level its states are found at) sits after the state's K bits
// and is left out of comparisons; otherwise every bit is compared (a stamp
// position past the key's words places nothing).
template <int W>
__device__ __forceinline__ void batch_fmt(SlotFmt<W> &f, const BatchInst *I, uint32_t stamp) {
  const uint32_t K = I->mode == kModePorProviso ? I->hdr.state_bits : 63u * W;
  make_fmt<W, true>(f, K, stamp);
}

// POR instances: 64-state chunks in the process-group layout of the
// single-model POR kernel (same cluster, ample or stubborn selection, cycle
// proviso against the instance's own partition), G = floor(64/N) states per
// pass.  A chunk is walked twice: the first walk only counts its successors,
// so the chunk reserves its output once (or is deferred whole); the second
// emits them.  Both walks select the same sets: the visited set does not
// change while successors are generated.
template <int W>
__global__ __launch_bounds__(256) void k_batch_expand_por(BatchArgs a) {
  extern __shared__ uint32_t smem[];
  BatchCtrl *C = a.ctrl;
  if (C->active == 0) return;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  GroupScratch X;
  X.xch = reinterpret_cast<uint2 *>(smem) + wib * 64;
  X.xc = smem + wpb * 128 + wib * 64;
  X.xf = smem + wpb * 192 + wib * 64;
  const uint32_t n = C->n;
  const uint32_t use_def = C->use_deferred;
  const uint32_t dc = C->deferred_cur;
  const uint64_t nchunks = use_def ? C->deferred_count[dc] : C->total_chunks_por;
  const uint64_t *def_in = a.deferred[dc];
  uint64_t *def_out = a.deferred[dc ^ 1u];
  uint64_t *def_out_count = &C->deferred_count[dc ^ 1u];
  const int lane = lane_id();
  const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  BlockChunk ch(a.cand, a.cand_cap, &C->cand_count);
  for (uint64_t c = gw; c < nchunks; c += nw) {
    uint32_t i;
    uint64_t start;
    if (use_def) {
      const uint64_t d = def_in[c];
      i = static_cast<uint32_t>(d >> 40);
      start = d & ((1ull << 40) - 1);
      if (a.inst[i].mode == kModePlain) continue;  // the plain kernel's
    } else {
      i = find_instance<true>(a.inst, n, c);
      start = (c - a.inst[i].chunk_prefix_por) * 64;
    }
    BatchInst *I = &a.inst[i];
    const ModelHdr H = I->hdr;
    const uint32_t *M = a.models + I->model_off;
    const GroupLane L = group_lane(M, H);
    const uint32_t cur = I->cur;
    const uint64_t f = I->front_count[cur];
    const uint64_t end = start + 64 < f ? start + 64 : f;
    const Table T{a.table + I->tb * static_cast<uint64_t>(a.bucket_words), I->nb, a.seed};
    SlotFmt<W> fmt;
    batch_fmt<W>(fmt, I, 0u);
    const bool proviso = I->mode == kModePorProviso;
    const int method = static_cast<int>(I->por_method);
    GroupAcc acc;
    bool counting = true;
    uint64_t total = 0, pos = 0;
    auto reserve = [&](uint64_t t, uint64_t *ob) {
      if (counting) {
        total += t;
        return 0;
      }
      *ob = pos;
      pos += t;
      return 1;
    };
    auto emit = [&](uint64_t q, const uint64_t (&sv)[W]) {
#pragma unroll
      for (int j = 0; j < W; j++) a.cand[q * W + j] = sv[j];
    };
    auto walk = [&]() {
      for (uint64_t b = start; b < end; b += static_cast<uint64_t>(L.G)) {
        const uint64_t si = b + static_cast<uint64_t>(L.g);
        const bool has = L.in_group && si < end;
        uint64_t key[W];
#pragma unroll
        for (int j = 0; j < W; j++) key[j] = has ? a.front[cur][(I->fo + si) * W + j] : kFlag;
        if (proviso)
          group_expand<W, kModePorProviso, W>(M, H, L, X, key, has, fmt, T, static_cast<int>(a.bucket_words),
                                              I->level, I->property, method, reserve, emit, acc);
        else
          group_expand<W, kModePor, W>(M, H, L, X, key, has, fmt, T, static_cast<int>(a.bucket_words), I->level,
                                       I->property, method, reserve, emit, acc);
      }
    };
    walk();  // count
    if (ballot(acc.bad) && lane == 0) I->bad_state = 1;
    uint64_t obase = 0;
    const int ok = ch.reserve<W>(a.cand_blk, i, total, (nchunks - c + nw - 1) / nw, &obase);
    if (ok < 0) {
      if (lane == 0) {
        I->cand_overflow = 1;
        C->cand_overflow = 1;
      }
      continue;
    }
    if (ok == 0) {
      if (lane == 0) {
        const uint64_t at = atomicAdd(reinterpret_cast<unsigned long long *>(def_out_count), 1ull);
        def_out[at] = (static_cast<uint64_t>(i) << 40) | start;
      }
      continue;
    }
    counting = false;
    pos = obase;
    walk();  // emit
    if (ballot(acc.viol) && lane == 0) I->violation = 1;
    if (lane == 0 && (acc.trans || acc.states)) {
      atomicAdd(reinterpret_cast<unsigned long long *>(&I->transitions), static_cast<unsigned long long>(acc.trans));
      atomicAdd(reinterpret_cast<unsigned long long *>(&I->expanded), static_cast<unsigned long long>(acc.states));
    }
  }
  ch.release_all<W>();
}

// New keys are staged per wave in LDS together with their instance and
// appended to that instance's next frontier in blocks, one atomicAdd per
// block: candidates of one instance come in long runs, and a per-key atomic
// on one instance's counter would serialise the device.
constexpr int kBatchStage = 256;

// FIND-OR-PUT of every candidate into its instance's partition.  As in
// k_insert, a wave walks runs of kInsertRun consecutive iterations in steps
// of 64 candidates (a pass too small for that spreads one iteration at a time
// over every wave): each lane reads one candidate, resolves its instance's
// partition (first bucket, size, and with stamps the slot format's state
// bits and level) and hashes it once, and the step's iterations hand all of
// that to the lane groups in candidate order.
template <int W, int BW, bool STAMPS>
__global__ __launch_bounds__(512)
void k_batch_insert(BatchArgs a) {
  extern __shared__ uint64_t bstage_all[];
  BatchCtrl *C = a.ctrl;
  if (C->active == 0) return;
  const uint64_t n = C->cand_count < a.cand_cap ? C->cand_count : a.cand_cap;
  constexpr int LPB = BW / 2;
  constexpr int CPW = 64 / LPB;
  static_assert(CPW <= kBatchStage, "stage smaller than one iteration");
  const int lane = lane_id();
  const int sub = lane % LPB;
  const int cg = lane / LPB;
  const int gbase = cg * LPB;
  uint64_t *stage = bstage_all + static_cast<size_t>(threadIdx.x >> 6) * (kBatchStage * W);
  uint32_t staged = 0;      // wave-uniform
  uint32_t stage_inst = 0;  // instance of the staged keys
  const uint64_t gw = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint64_t acc_look = 0;
  auto flush = [&]() {
    if (!staged) return;
    wave_sync();
    BatchInst *I = &a.inst[stage_inst];
    const uint32_t nxt = I->cur ^ 1u;
    uint64_t ob = 0;
    if (lane == 0)
      ob = atomicAdd(reinterpret_cast<unsigned long long *>(&I->front_count[nxt]),
                     static_cast<unsigned long long>(staged));
    ob = __shfl(ob, 0, 64);
    bool over = false;
    for (uint32_t k = lane; k < staged * W; k += 64) {
      const uint64_t idx = ob + k / W;
      if (idx < I->fcap)
        a.front[nxt][(I->fo + ob) * W + k] = stage[k];
      else
        over = true;
    }
    if (ballot(over) && lane == 0) I->overflow = 1;
    wave_sync();
    staged = 0;
  };
  // A step of 64 candidates (or an iteration of CPW) lies in one 64-entry
  // block of the candidate buffer, so its instance is wave-uniform: one
  // scalar read of the block's id, and the partition (first bucket, size,
  // and with stamps the slot format's state bits and level) is read again
  // only when the instance changes.
  uint32_t pi = kNoInst;
  uint64_t ptb = 0, pnb = 1;
  uint32_t pK = 63u * W, pstamp = 0;
  const uint32_t ninst = C->n;
  // the id of the block holding candidate c (a wave-uniform read)
  auto block_id = [&](uint64_t c) {
    const uint64_t blk = (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(c >> 38))) << 32) |
                         __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(c >> 6));
    return a.cand_blk[blk];
  };
  auto partition = [&](uint32_t id) {
    uint32_t i = __builtin_amdgcn_readfirstlane(id);
    i = i < ninst ? i : 0u;  // a block with an active entry holds a valid id; stay in bounds anyway
    if (i != pi) {
      const BatchInst *I = &a.inst[i];
      pi = i;
      ptb = I->tb;
      pnb = I->nb;
      if (STAMPS) {
        pK = I->mode == kModePorProviso ? I->hdr.state_bits : 63u * W;
        pstamp = I->level + 1u;  // new states are stamped with the level they are found at
      }
    }
  };
  auto visit = [&](const uint64_t (&key)[W], uint64_t h, uint64_t b0, bool active) {
    const Table T{a.table + ptb * BW, pnb, a.seed};
    SlotFmt<W> f;
    if (STAMPS)
      make_fmt<W, true>(f, pK, pstamp);
    else
      make_fmt<W, false>(f, 0, 0);
    uint64_t v[W];
    slot_value<W, W>(key, f, v);
    const int st = find_or_put_group_b<W, BW>(T, v, f, b0, [h]() { return h; }, sub, gbase, active);
    acc_look += __popcll(ballot(active && sub == 0));
    if (active && sub == 0) {
      if (st == kFull) a.inst[pi].table_full = 1;
      if (st == kTimeout) a.inst[pi].spin_timeout = 1;
    }
    const bool is_new = active && st == kInserted;
    const uint64_t m = ballot(is_new && sub == 0);
    if (!m) return;
    const uint32_t cnt = __popcll(m);
    if (staged && (stage_inst != pi || staged + cnt > static_cast<uint32_t>(kBatchStage))) flush();
    stage_inst = pi;
    if (is_new) {
      const uint32_t pos = staged + __popcll(m & ((1ull << gbase) - 1ull));
#pragma unroll
      for (int j = 0; j < W; j++)
        if (j % LPB == sub) stage[pos * W + j] = key[j];
    }
    staged += cnt;
  };
  const uint64_t iters = insert_run(n, nw * CPW);
  if (iters < kInsertSpread) {
    const uint64_t RUN = CPW * iters;
    for (uint64_t r0 = gw * RUN, c0 = r0; c0 < n;) {
      const uint64_t cstep = c0;
      const uint64_t ci = c0 + cg;
      c0 += CPW;
      if (c0 == r0 + RUN) {
        r0 += nw * RUN;
        c0 = r0;
      }
      uint64_t key[W];
#pragma unroll
      for (int j = 0; j < W; j++) key[j] = ci < n ? __builtin_nontemporal_load(a.cand + ci * W + j) : 0;
      const bool active = key[0] != 0;  // past the end or a released reservation: zero
      if (!ballot(active)) continue;
      partition(block_id(cstep));
      const uint64_t h = key_hash<W>(key, a.seed);
      visit(key, h, bucket_of(h, 0, pnb), active);
    }
  } else {
    static_assert(kInsertRun % LPB == 0, "a run is a whole number of steps");
    const uint64_t RUN = CPW * kInsertRun;
    // Software-pipelined by one step: the next step's keys and block id are
    // read while this step's lookups run (the block id decides the partition,
    // hence the first bucket, so reading it in the step itself would put one
    // more memory round trip before every step's lookups).
    uint64_t r0 = gw * RUN, c0 = r0;
    uint64_t key_n[W];
    uint32_t id_n = 0;
    auto fetch = [&](uint64_t cs) {
      const uint64_t cl = cs + lane;
#pragma unroll
      for (int j = 0; j < W; j++) key_n[j] = cl < n ? __builtin_nontemporal_load(a.cand + cl * W + j) : 0;
      id_n = block_id(cs);
    };
    if (c0 < n) fetch(c0);
    while (c0 < n) {
      uint64_t key_l[W];
#pragma unroll
      for (int j = 0; j < W; j++) key_l[j] = key_n[j];
      const uint32_t id = id_n;
      c0 += 64;
      if (c0 == r0 + RUN) {
        r0 += nw * RUN;
        c0 = r0;
      }
      if (c0 < n) fetch(c0);
      const uint64_t act_m = ballot(key_l[0] != 0);
      if (!act_m) continue;
      partition(id);
      const uint64_t h_l = key_hash<W>(key_l, a.seed);
      const uint64_t b_l = bucket_of(h_l, 0, pnb);
#pragma unroll 1
      for (int s2 = 0; s2 < LPB; s2++) {
        const int src = s2 * CPW + cg;
        uint64_t key[W];
#pragma unroll
        for (int j = 0; j < W; j++) key[j] = __shfl(key_l[j], src, 64);
        visit(key, __shfl(h_l, src, 64), __shfl(b_l, src, 64), (act_m >> src) & 1ull);
      }
    }
  }
  flush();
  if (lane == 0 && acc_look)
    atomicAdd(reinterpret_cast<unsigned long long *>(&C->lookups), static_cast<unsigned long long>(acc_look));
}

// One workgroup of 1024 threads; instances are handled in tiles of 1024.
// Per instance: the level step and termination; then inclusive scans of the
// next round's plain and POR chunk counts.
__global__ __launch_bounds__(1024) void k_batch_advance(BatchArgs a) {
  __shared__ uint64_t scan[2][1024];
  __shared__ uint32_t nact;
  BatchCtrl *C = a.ctrl;
  if (C->active == 0 && C->prefix_only == 0) return;
  const uint32_t tid = threadIdx.x;
  const bool deferred = C->deferred_count[C->deferred_cur ^ 1u] > 0;
  const bool prefix_only = C->prefix_only != 0;  // recompute chunks after admissions
  __syncthreads();
  if (tid == 0) {
    C->cand_entries += C->cand_count < a.cand_cap ? C->cand_count : a.cand_cap;
    C->cand_count = 0;
    if (!prefix_only) C->rounds += 1;
  }
  if (deferred && !prefix_only) {  // the level is not complete: run the deferred chunks
    if (tid == 0) {
      const uint32_t dc = C->deferred_cur;
      C->deferred_count[dc] = 0;
      C->deferred_cur = dc ^ 1u;
      C->use_deferred = 1;
    }
    return;
  }
  if (tid == 0) {
    C->deferred_count[C->deferred_cur] = 0;
    C->use_deferred = 0;
    nact = 0;
  }
  __syncthreads();
  const uint32_t n = C->n;
  uint64_t carry[2] = {0, 0};
  for (uint32_t t0 = 0; t0 < n; t0 += 1024) {
    const uint32_t i = t0 + tid;
    uint64_t chunks[2] = {0, 0};  // plain, POR
    if (i < n) {
      BatchInst *I = &a.inst[i];
      if (!I->stop && !prefix_only) {
        const uint32_t cur = I->cur, nxt = cur ^ 1u;
        const uint64_t fresh = I->front_count[nxt] < I->fcap ? I->front_count[nxt] : I->fcap;
        I->inserts += fresh;
        I->front_count[cur] = 0;
        I->cur = nxt;
        I->level += 1u;
        if (I->violation || I->table_full || I->overflow || I->bad_state || I->spin_timeout || I->cand_overflow)
          I->stop = 1;
        if (I->front_count[nxt] == 0) I->stop = 1;
      }
      if (!I->stop) {
        const uint64_t f = I->front_count[I->cur];
        if (I->mode == kModePlain)
          chunks[0] = (f + a.chunk_states - 1) / a.chunk_states;
        else
          chunks[1] = (f + 63) / 64;
        atomicAdd(&nact, 1u);
      }
    }
    for (int k = 0; k < 2; k++) scan[k][tid] = chunks[k];
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {  // inclusive scans
      const uint64_t v0 = tid >= d ? scan[0][tid - d] : 0;
      const uint64_t v1 = tid >= d ? scan[1][tid - d] : 0;
      __syncthreads();
      scan[0][tid] += v0;
      scan[1][tid] += v1;
      __syncthreads();
    }
    if (i < n) {
      a.inst[i].chunk_prefix = carry[0] + scan[0][tid] - chunks[0];
      a.inst[i].chunk_prefix_por = carry[1] + scan[1][tid] - chunks[1];
    }
    carry[0] += scan[0][1023];
    carry[1] += scan[1][1023];
    __syncthreads();
  }
  if (tid == 0) {
    C->total_chunks = carry[0];
    C->total_chunks_por = carry[1];
    C->active = nact;
    C->prefix_only = 0;
  }
}

template <int W, int BW>
static void launch_insert(const BatchArgs &a, uint32_t blocks, hipStream_t s) {
  const uint32_t it = kInsertThreads;
  const size_t ishm = static_cast<size_t>(it / 64) * kBatchStage * W * 8;
  if (a.stamps)
    hipLaunchKernelGGL((k_batch_insert<W, BW, true>), dim3(blocks), dim3(it), ishm, s, a);
  else
    hipLaunchKernelGGL((k_batch_insert<W, BW, false>), dim3(blocks), dim3(it), ishm, s, a);
}

template <int W>
hipError_t launch_round(const BatchArgs &a, int BW, uint32_t blocks, uint32_t threads, uint32_t ins_blocks,
                        hipStream_t s, const RoundEvents *ev) {
  const size_t shm = a.expand_lds;
  if (ev) {
    const hipError_t e = hipEventRecord(ev->e0, s);
    if (e != hipSuccess) return e;
  }
  if (a.rows16)
    hipLaunchKernelGGL((k_batch_expand<W, uint16_t>), dim3(blocks), dim3(threads), shm, s, a);
  else
    hipLaunchKernelGGL((k_batch_expand<W, uint32_t>), dim3(blocks), dim3(threads), shm, s, a);
  if (a.por) {
    const size_t pshm = static_cast<size_t>(a.por_threads / 64) * kGroupScratchBytesPerWave;
    hipLaunchKernelGGL((k_batch_expand_por<W>), dim3(a.por_blocks), dim3(a.por_threads), pshm, s, a);
  }
  if (ev) {
    const hipError_t e = hipEventRecord(ev->e1, s);
    if (e != hipSuccess) return e;
  }
  switch (BW) {
    case 8: launch_insert<W, 8>(a, ins_blocks, s); break;
    case 16: launch_insert<W, 16>(a, ins_blocks, s); break;
    case 32: launch_insert<W, 32>(a, ins_blocks, s); break;
    default: return hipErrorInvalidValue;
  }
  if (ev) {
    const hipError_t e = hipEventRecord(ev->e2, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(1024), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_round_w(const BatchArgs &a, int W, int BW, uint32_t blocks, uint32_t threads,
                          uint32_t ins_blocks, hipStream_t s, const RoundEvents *ev) {
  switch (W) {
    case 1: return launch_round<1>(a, BW, blocks, threads, ins_blocks, s, ev);
    case 2: return launch_round<2>(a, BW, blocks, threads, ins_blocks, s, ev);
    case 3: return launch_round<3>(a, BW, blocks, threads, ins_blocks, s, ev);
    case 4: return launch_round<4>(a, BW, blocks, threads, ins_blocks, s, ev);
    case 5: return launch_round<5>(a, BW, blocks, threads, ins_blocks, s, ev);
    default: return hipErrorInvalidValue;
  }
}

// The initial state (all processes in local state 0, cu:668-682) goes to
// slot 0 of its first probe bucket: the partition was zeroed just before in
// stream order, so that is where FIND-OR-PUT would have put it.
template <int W>
__global__ void k_place_initial(BatchArgs a, uint64_t tb, uint64_t nb, uint64_t fo, uint32_t BW) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t key[W];
#pragma unroll
  for (int j = 0; j < W; j++) key[j] = kFlag;
  const uint64_t b = bucket_of(key_hash<W>(key, a.seed), 0, nb);
  uint64_t *slot0 = a.table + (tb + b) * BW;
#pragma unroll
  for (int j = 0; j < W; j++) {
    slot0[j] = key[j];
    a.front[0][fo * W + j] = key[j];
  }
}

// every instance's initial state, one thread per instance (placed as
// k_place_initial places one)
template <int W>
__global__ void k_place_initial_all(BatchArgs a, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BatchInst &I = a.inst[i];
  uint64_t key[W];
#pragma unroll
  for (int j = 0; j < W; j++) key[j] = kFlag;
  const uint64_t b = bucket_of(key_hash<W>(key, a.seed), 0, I.nb);
  uint64_t *slot0 = a.table + (I.tb + b) * a.bucket_words;
#pragma unroll
  for (int j = 0; j < W; j++) {
    slot0[j] = key[j];
    a.front[0][I.fo * W + j] = key[j];
  }
}

hipError_t launch_place_initial_all(const BatchArgs &a, int W, uint32_t n, hipStream_t s) {
  const dim3 g((n + 255) / 256), b(256);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_place_initial_all<1>, g, b, 0, s, a, n); break;
    case 2: hipLaunchKernelGGL(k_place_initial_all<2>, g, b, 0, s, a, n); break;
    case 3: hipLaunchKernelGGL(k_place_initial_all<3>, g, b, 0, s, a, n); break;
    case 4: hipLaunchKernelGGL(k_place_initial_all<4>, g, b, 0, s, a, n); break;
    case 5: hipLaunchKernelGGL(k_place_initial_all<5>, g, b, 0, s, a, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_place_initial(const BatchArgs &a, int W, int BW, uint64_t tb, uint64_t nb, uint64_t fo,
                                hipStream_t s) {
  if (nb == 0) return hipErrorInvalidValue;
  const uint32_t bw = static_cast<uint32_t>(BW);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_place_initial<1>, dim3(1), dim3(64), 0, s, a, tb, nb, fo, bw); break;
    case 2: hipLaunchKernelGGL(k_place_initial<2>, dim3(1), dim3(64), 0, s, a, tb, nb, fo, bw); break;
    case 3: hipLaunchKernelGGL(k_place_initial<3>, dim3(1), dim3(64), 0, s, a, tb, nb, fo, bw); break;
    case 4: hipLaunchKernelGGL(k_place_initial<4>, dim3(1), dim3(64), 0, s, a, tb, nb, fo, bw); break;
    case 5: hipLaunchKernelGGL(k_place_initial<5>, dim3(1), dim3(64), 0, s, a, tb, nb, fo, bw); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_advance(const BatchArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(1024), 0, s, a);
  return hipGetLastError();
}

}  // namespace gx

using namespace gx;

namespace {


int map_err(hipError_t e) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return GX_E_NOMEM;
  return GX_E_HIP;
}

#define BX_TRY(expr)                          \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return map_err(e_); \
  } while (0)

}  // namespace

extern "C" void gx_batch_options_default(gx_batch_options *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->threads = 256;
  o->levels_per_sync = 8;
  o->use_graph = 1;
  o->hash_seed = 0x5eed5eed5eedull;
}

extern "C" int gx_explore_batch(const gx_model *const *models, const int32_t *properties,
                                const uint64_t *weights, int n, const gx_batch_options *o,
                                gx_result *results, gx_batch_stats *stats) {
  if (!models || n <= 0 || n > (1 << 20)) return GX_E_ARG;
  std::vector<gx_job_options> jobs(n);
  for (int i = 0; i < n; i++) {
    if (models[i] && models[i]->por) return GX_E_ARG;  // plain method only (gx_explore_batch_jobs has POR)
    jobs[i] = gx_job_options{properties ? properties[i] : GX_PROP_NONE, 0, 0, GX_POR_CAMPLE};
  }
  return gx_explore_batch_jobs(models, jobs.data(), weights, n, o, results, stats);
}

extern "C" int gx_explore_batch_jobs(const gx_model *const *models, const gx_job_options *jobs,
                                     const uint64_t *weights, int n, const gx_batch_options *o,
                                     gx_result *results, gx_batch_stats *stats) {
  if (!models || !jobs || !o || !results || n <= 0 || n > (1 << 20)) return GX_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GX_E_NODEVICE;
  memset(results, 0, sizeof(gx_result) * n);
  if (stats) memset(stats, 0, sizeof *stats);
  // layout: common key width, concatenated models
  int W = 1;
  uint32_t max_procs = 1, max_rows = 0;
  bool any_por = false, any_stamp = false;
  std::vector<uint32_t> mbuf;
  std::vector<BatchInst> inst(n);
  for (int i = 0; i < n; i++) {
    const gx_model *m = models[i];
    const gx_job_options &J = jobs[i];
    if (!m) return GX_E_ARG;
    if (J.property < GX_PROP_NONE || J.property > GX_PROP_LIVENESS || J.por_method > GX_POR_STUBBORN)
      return GX_E_ARG;
    const ModelHdr &H = m->comp.hdr;
    if (!J.por && !H.lab64) return GX_E_LIMIT;  // the state-per-lane kernel's label masks
    W = std::max<int>(W, static_cast<int>(H.key_words));
    // a level stamp after the state bits needs K + 32 bits in the slot
    if (J.por && J.cycle_proviso) W = std::max<int>(W, static_cast<int>((H.state_bits + 32 + 62) / 63));
    max_procs = std::max(max_procs, H.nprocs);
    max_rows = std::max(max_rows, H.nrows);
    any_por = any_por || J.por;
    any_stamp = any_stamp || (J.por && J.cycle_proviso);
    memset(&inst[i], 0, sizeof(BatchInst));
    inst[i].hdr = H;
    inst[i].model_off = static_cast<uint32_t>(mbuf.size());
    inst[i].property = J.property;
    inst[i].mode = batch_mode(J.por, J.cycle_proviso);
    inst[i].por_method = J.por_method;
    inst[i].cur = 0;
    inst[i].level = 0;
    inst[i].front_count[0] = 1;  // the initial state, placed at level 0 below
    inst[i].inserts = 1;
    mbuf.insert(mbuf.end(), m->comp.words.begin(), m->comp.words.end());
    while (mbuf.size() % 4) mbuf.push_back(0);
  }
  if (W > kMaxKeyWords) return GX_E_LIMIT;
  const uint32_t bb = o->bucket_bytes ? o->bucket_bytes : (W <= 2 ? 64u : 128u);
  if (bb != 64 && bb != 128 && bb != 256) return GX_E_ARG;
  const int BW = static_cast<int>(bb / 8);
  if (BW < W) return GX_E_ARG;
  const uint64_t spb = static_cast<uint64_t>(BW / W);
  uint32_t threads = o->threads ? o->threads : 256;
  threads = std::min<uint32_t>(512, std::max<uint32_t>(64, (threads + 63) / 64 * 64));
  const bool rows16 = max_rows < 0xFFFFu;
  if (batch_rows_lds(threads, max_procs, rows16) > 64 * 1024) return GX_E_LIMIT;
  int dev = 0;
  BX_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  BX_TRY(hipGetDeviceProperties(&prop, dev));
  const uint32_t blocks = o->blocks ? o->blocks : prop.multiProcessorCount * (2048 / threads);
  // rounds of a batch are small and captured once: a modest insert grid
  const uint32_t ins_blocks = o->insert_blocks ? o->insert_blocks : prop.multiProcessorCount * 8;
  // table partitions proportional to the weights (equal without weights)
  // default: 2^22 ints (16 MiB) per instance
  const uint64_t q_ints = o->q_ints ? o->q_ints : (static_cast<uint64_t>(n) << 22);
  const uint64_t total_buckets = std::max<uint64_t>(q_ints * 4 / bb, static_cast<uint64_t>(n));
  double wsum = 0;
  for (int i = 0; i < n; i++) wsum += weights ? std::max<double>(1.0, static_cast<double>(weights[i])) : 1.0;
  uint64_t tb = 0, fo = 0;
  for (int i = 0; i < n; i++) {
    const double wi = weights ? std::max<double>(1.0, static_cast<double>(weights[i])) : 1.0;
    uint64_t nb = static_cast<uint64_t>(static_cast<double>(total_buckets) * wi / wsum);
    nb = std::max<uint64_t>(nb, 1);
    if (tb + nb > total_buckets) nb = total_buckets > tb ? total_buckets - tb : 0;
    if (nb == 0) return GX_E_ARG;
    inst[i].tb = tb;
    inst[i].nb = nb;
    inst[i].fo = fo;
    inst[i].fcap = nb * spb;  // states fit the partition's slots
    tb += nb;
    fo += inst[i].fcap;
  }
  const uint64_t front_total = fo;
  uint64_t cand_cap = o->cand_capacity ? o->cand_capacity : std::max<uint64_t>(1u << 22, front_total);
  cand_cap = std::max<uint64_t>(std::min<uint64_t>(cand_cap, 0xFFFFFFFFull), static_cast<uint64_t>(n));
  const uint64_t max_chunks = front_total / 64 + static_cast<uint64_t>(n) + 1;

  const auto t_init0 = std::chrono::steady_clock::now();
  DevBuf d_inst, d_ctrl, d_models, d_f0, d_f1, d_cand, d_cinst, d_def0, d_def1, d_table;
  BX_TRY(d_inst.alloc(sizeof(BatchInst) * n));
  BX_TRY(d_ctrl.alloc(sizeof(BatchCtrl)));
  BX_TRY(d_models.alloc(mbuf.size() * 4));
  BX_TRY(d_f0.alloc(front_total * W * 8));
  BX_TRY(d_f1.alloc(front_total * W * 8));
  BX_TRY(d_cand.alloc(cand_cap * W * 8));
  BX_TRY(d_cinst.alloc((cand_cap / 64 + 1) * 4));
  BX_TRY(d_def0.alloc(max_chunks * 8));
  BX_TRY(d_def1.alloc(max_chunks * 8));
  BX_TRY(d_table.alloc(total_buckets * bb));
  hipStream_t st;
  BX_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  // on every return, the stream's work ends before the buffers are freed
  std::unique_ptr<void, void (*)(void *)> sg(st, [](void *s) {
    (void)hipStreamSynchronize(static_cast<hipStream_t>(s));
    (void)hipStreamDestroy(static_cast<hipStream_t>(s));
  });
  BX_TRY(hipMemsetAsync(d_table.p, 0, total_buckets * bb, st));
  BX_TRY(hipMemcpyAsync(d_models.p, mbuf.data(), mbuf.size() * 4, hipMemcpyHostToDevice, st));
  BX_TRY(hipMemcpyAsync(d_inst.p, inst.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice, st));
  // every instance starts at level 0 with its initial state (all processes
  // in local state 0, cu:668-682) in its partition and frontier; the first
  // advance only computes the chunk prefixes
  BatchCtrl c0{};
  c0.n = static_cast<uint32_t>(n);
  c0.active = static_cast<uint32_t>(n);
  c0.prefix_only = 1;
  BX_TRY(hipMemcpyAsync(d_ctrl.p, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  BatchArgs a{};
  a.inst = static_cast<BatchInst *>(d_inst.p);
  a.ctrl = static_cast<BatchCtrl *>(d_ctrl.p);
  a.models = static_cast<const uint32_t *>(d_models.p);
  a.front[0] = static_cast<uint64_t *>(d_f0.p);
  a.front[1] = static_cast<uint64_t *>(d_f1.p);
  a.cand = static_cast<uint64_t *>(d_cand.p);
  a.cand_blk = static_cast<uint32_t *>(d_cinst.p);
  a.cand_cap = cand_cap;
  a.deferred[0] = static_cast<uint64_t *>(d_def0.p);
  a.deferred[1] = static_cast<uint64_t *>(d_def1.p);
  a.table = static_cast<uint64_t *>(d_table.p);
  a.seed = o->hash_seed;
  a.bucket_words = static_cast<uint32_t>(BW);
  a.max_procs = max_procs;
  a.rows16 = rows16 ? 1u : 0u;
  a.por = any_por ? 1u : 0u;
  a.stamps = any_stamp ? 1u : 0u;
  a.por_threads = 256;
  a.por_blocks = prop.multiProcessorCount * 4;
  a.chunk_states = threads;  // 64 states per wave
  a.expand_lds = static_cast<uint32_t>(batch_expand_lds(threads, max_procs, rows16));
  BX_TRY(launch_place_initial_all(a, W, static_cast<uint32_t>(n), st));
  BX_TRY(launch_batch_advance(a, st));
  BX_TRY(hipStreamSynchronize(st));
  const double t_init = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_init0).count();

  // kernel timing: one round per host synchronisation, launched directly
  const bool timed = o->time_kernels != 0;
  const uint32_t k = timed ? 1u : std::max<uint32_t>(1, o->levels_per_sync);
  BatchCtrl ch{};
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  struct EventGuard {
    RoundEvents ev{nullptr, nullptr, nullptr};
    ~EventGuard() {
      for (hipEvent_t e : {ev.e0, ev.e1, ev.e2})
        if (e) (void)hipEventDestroy(e);
    }
  } eg;
  if (timed) {
    BX_TRY(hipEventCreate(&eg.ev.e0));
    BX_TRY(hipEventCreate(&eg.ev.e1));
    BX_TRY(hipEventCreate(&eg.ev.e2));
  }
  double t_exp_ms = 0, t_ins_ms = 0;
  uint64_t ins_launches = 0;
  if (o->use_graph && !timed) {
    BX_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipSuccess;
    for (uint32_t r = 0; r < k && ce == hipSuccess; r++)
      ce = launch_round_w(a, W, BW, blocks, threads, ins_blocks, st);
    hipError_t ee = hipStreamEndCapture(st, &graph);
    BX_TRY(ce);
    BX_TRY(ee);
    BX_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  struct GraphGuard {
    hipGraph_t g;
    hipGraphExec_t e;
    ~GraphGuard() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } gg{graph, exec};
  uint64_t launches = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    if (exec) {
      BX_TRY(hipGraphLaunch(exec, st));
    } else {
      for (uint32_t r = 0; r < k; r++)
        BX_TRY(launch_round_w(a, W, BW, blocks, threads, ins_blocks, st, timed ? &eg.ev : nullptr));
    }
    launches++;
    BX_TRY(hipMemcpyAsync(&ch, d_ctrl.p, sizeof ch, hipMemcpyDeviceToHost, st));
    BX_TRY(hipStreamSynchronize(st));
    if (timed) {
      float ms = 0;
      BX_TRY(hipEventElapsedTime(&ms, eg.ev.e0, eg.ev.e1));
      t_exp_ms += ms;
      BX_TRY(hipEventElapsedTime(&ms, eg.ev.e1, eg.ev.e2));
      t_ins_ms += ms;
      ins_launches++;
    }
    if (ch.active == 0) break;
  }
  const double t_explore = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  BX_TRY(hipMemcpy(inst.data(), d_inst.p, sizeof(BatchInst) * n, hipMemcpyDeviceToHost));
  int rc = GX_OK;
  for (int i = 0; i < n; i++) {
    const BatchInst &I = inst[i];
    gx_result &r = results[i];
    r.states = I.inserts;
    r.transitions = I.transitions;
    r.levels = I.level;
    r.inserts = I.inserts;
    r.violation = I.violation ? 1 : 0;
    r.table_full = I.table_full ? 1 : 0;
    r.table_bytes = I.nb * bb;
    r.t_explore_s = t_explore;
    r.t_init_s = t_init;
    if (I.bad_state) rc = GX_E_STATE;
    if ((I.overflow || I.spin_timeout || I.cand_overflow) && rc == GX_OK) rc = GX_E_CAPACITY;
  }
  if (stats) {
    stats->t_explore_s = t_explore;
    stats->t_init_s = t_init;
    stats->rounds = ch.rounds;
    stats->graph_launches = launches;
    stats->lookups = ch.lookups;
    stats->cand_entries = ch.cand_entries;
    stats->t_expand_s = t_exp_ms * 1e-3;
    stats->t_insert_s = t_ins_ms * 1e-3;
    stats->insert_launches = ins_launches;
    for (int i = 0; i < n; i++) {
      stats->states += inst[i].inserts;
      stats->transitions += inst[i].transitions;
    }
  }
  return rc;
}
