// gx_device.h -- device primitives for gfx950 (wave64).
//
//  * key codec: process fields inside the 63-bit-chunk key (the reference's
//    GETSTATEVECTORSTATE / SETSTATEVECTORSTATE, cu:162-178, on a different
//    word layout);
//  * hashing and bucket selection (replaces HASHALL, cu:307-327, with a
//    seeded 64-bit mix and a multiply-shift range reduction, so bucket
//    indices are 64-bit and the table can span all of HBM);
//  * wave and process-group primitives (64-lane ballot / bpermute);
//  * the cooperative FIND-OR-PUT bucket probe (replaces FINDORPUT_WARP,
//    cu:361-450) and the thread-level FIND used by the cycle proviso
//    (replaces FIND_WARP, cu:454-493).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gx_common.h"

namespace gx {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

template <int W>
__host__ __device__ __forceinline__ uint64_t key_hash(const uint64_t (&k)[W], uint64_t seed) {
  uint64_t h = seed ^ 0x9e3779b97f4a7c15ull;
#pragma unroll
  for (int j = 0; j < W; j++) h = mix64(h ^ k[j]) + static_cast<uint64_t>(j);
  return h;
}

// bucket of probe i: multiply-shift reduction of an independent 64-bit hash
__host__ __device__ __forceinline__ uint64_t bucket_of(uint64_t h, int i, uint64_t nbuckets) {
  uint64_t hi = mix64(h + static_cast<uint64_t>(i + 1) * 0x9e3779b97f4a7c15ull);
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(hi, nbuckets);
#else
  return static_cast<uint64_t>((static_cast<unsigned __int128>(hi) * nbuckets) >> 64);
#endif
}

// ---------------------------------------------------------------------------
// key codec.  floc = chunk | offset << 8 | width << 16.
template <int W>
__device__ __forceinline__ uint32_t kget(const uint64_t (&k)[W], uint32_t floc) {
  const uint32_t c = floc & 0xFFu, o = (floc >> 8) & 0xFFu, w = floc >> 16;
  uint64_t lo = 0, hi = 0;
#pragma unroll
  for (int j = 0; j < W; j++) {
    if (static_cast<uint32_t>(j) == c) lo = k[j];
    if (static_cast<uint32_t>(j) == c + 1) hi = k[j];
  }
  uint64_t x = (lo & kPayload) >> o;
  if (o + w > 63) x |= (hi & kPayload) << (63 - o);
  const uint64_t m = (w >= 32) ? 0xFFFFFFFFull : ((1ull << w) - 1ull);
  return static_cast<uint32_t>(x & m);
}

template <int W>
__device__ __forceinline__ void kset(uint64_t (&k)[W], uint32_t floc, uint32_t v) {
  const uint32_t c = floc & 0xFFu, o = (floc >> 8) & 0xFFu, w = floc >> 16;
  const uint64_t m = (w >= 32) ? 0xFFFFFFFFull : ((1ull << w) - 1ull);
  const uint64_t x = static_cast<uint64_t>(v) & m;
  const uint64_t mlo = (m << o) & kPayload;
  const uint64_t xlo = (x << o) & kPayload;
  const bool straddle = (o + w > 63);
  const uint64_t mhi = straddle ? ((1ull << (o + w - 63)) - 1ull) : 0ull;
  const uint64_t xhi = straddle ? (x >> (63 - o)) : 0ull;
#pragma unroll
  for (int j = 0; j < W; j++) {
    if (static_cast<uint32_t>(j) == c) k[j] = (k[j] & ~mlo) | xlo;
    if (static_cast<uint32_t>(j) == c + 1) k[j] = (k[j] & ~mhi) | xhi;
  }
}

// ---------------------------------------------------------------------------
// wave primitives
__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// wave-wide exclusive prefix sum; *total gets the sum over all 64 lanes
__device__ __forceinline__ uint64_t wave_excl_scan(uint64_t v, uint64_t *total) {
  const int lane = lane_id();
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// Copies n words of the model into LDS with the whole workgroup, eight
// loads in flight per thread: a loop with one load per iteration waits a
// memory round trip per 4 B x blockDim, which made a level of one state
// take 30-58 us in successor generation.
__device__ __forceinline__ void stage_words(uint32_t *dst, const uint32_t *__restrict__ src, uint32_t n) {
  constexpr uint32_t K = 8;
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += K * blockDim.x) {
    uint32_t v[K];
#pragma unroll
    for (uint32_t k = 0; k < K; k++) {
      const uint32_t i = i0 + k * blockDim.x;
      v[k] = i < n ? src[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < K; k++) {
      const uint32_t i = i0 + k * blockDim.x;
      if (i < n) dst[i] = v[k];
    }
  }
}

// ordering of LDS traffic between lanes of one wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Reductions over the N lanes of one process group: rotate-and-combine
// with distances 1, 2, 4, ... (covers any N, not only powers of two).  The
// source lanes of the rotations are computed once per kernel.
struct GroupRot {
  int src[5];
  int steps;
};

__device__ __forceinline__ GroupRot make_rot(int N, int base, int p, bool in_group) {
  GroupRot r;
  r.steps = 0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    const int j = 1 << i;
    r.src[i] = in_group ? base + ((p + j) % N) : lane_id();
    if (j < N) r.steps = i + 1;
  }
  return r;
}

__device__ __forceinline__ uint32_t group_min(uint32_t v, const GroupRot &r) {
#pragma unroll
  for (int i = 0; i < 5; i++) {
    if (i < r.steps) {
      const uint32_t o = __shfl(v, r.src[i], 64);
      v = o < v ? o : v;
    }
  }
  return v;
}

__device__ __forceinline__ uint32_t group_or(uint32_t v, const GroupRot &r) {
#pragma unroll
  for (int i = 0; i < 5; i++)
    if (i < r.steps) v |= __shfl(v, r.src[i], 64);
  return v;
}

// ---------------------------------------------------------------------------
// Visited set.  nbuckets buckets of BW 64-bit words; slot s of a bucket is
// words [s*WT, s*WT+WT).  A slot is empty while its first word is 0.
struct Table {
  uint64_t *words;
  uint64_t nbuckets;
  uint64_t seed;
};

// Per-slot-word comparison masks and stamp bits.  Without stamps every bit
// is compared; with stamps the 32-bit level stamp at bits [K, K+32) of the
// payload string is excluded from comparisons.
template <int WT>
struct SlotFmt {
  uint64_t cmp[WT];
  uint64_t stamp[WT];
};

template <int WT, bool STAMP>
__device__ __forceinline__ void make_fmt(SlotFmt<WT> &f, uint32_t K, uint32_t stamp) {
#pragma unroll
  for (int j = 0; j < WT; j++) {
    f.cmp[j] = ~0ull;
    f.stamp[j] = 0;
    if (STAMP) {
      // bits of word j cover payload bits [63j, 63j+63)
      const int64_t lo = 63 * j, hi = lo + 63;
      uint64_t keep = kFlag;
      // key bits [0, K)
      if (static_cast<int64_t>(K) > lo) {
        int64_t n = (static_cast<int64_t>(K) < hi ? static_cast<int64_t>(K) : hi) - lo;
        keep |= (n >= 63) ? kPayload : ((1ull << n) - 1ull);
      }
      f.cmp[j] = keep;
      // stamp bits [K, K+32)
      const int64_t s0 = K, s1 = static_cast<int64_t>(K) + 32;
      const int64_t a = s0 > lo ? s0 : lo, b = s1 < hi ? s1 : hi;
      if (a < b) {
        uint64_t part = static_cast<uint64_t>(stamp) >> (a - s0);
        uint64_t m = (1ull << (b - a)) - 1ull;
        f.stamp[j] = (part & m) << (a - lo);
      }
    }
  }
}

// slot value of a key (W key words, WT slot words)
template <int W, int WT>
__device__ __forceinline__ void slot_value(const uint64_t (&k)[W], const SlotFmt<WT> &f,
                                           uint64_t (&v)[WT]) {
#pragma unroll
  for (int j = 0; j < WT; j++) v[j] = (j < W ? k[j] : kFlag) | f.stamp[j];
}

template <int WT>
__device__ __forceinline__ uint32_t read_stamp(const uint64_t (&s)[WT], uint32_t K) {
  uint64_t r = 0;
#pragma unroll
  for (int j = 0; j < WT; j++) {
    const int64_t lo = 63 * j, hi = lo + 63;
    const int64_t s0 = K, s1 = static_cast<int64_t>(K) + 32;
    const int64_t a = s0 > lo ? s0 : lo, b = s1 < hi ? s1 : hi;
    if (a < b) {
      uint64_t m = (1ull << (b - a)) - 1ull;
      r |= (((s[j] & kPayload) >> (a - lo)) & m) << (a - s0);
    }
  }
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ uint64_t load_coherent(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_coherent(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

enum ProbeStatus : int { kPending = 0, kFound = 1, kInserted = 2, kFull = 3, kTimeout = 4 };

constexpr int kSpinLimit = 1 << 22;

// Cooperative FIND-OR-PUT of one key by a group of BW/2 lanes; lane `sub`
// of the group holds bucket words 2*sub and 2*sub+1.  Every lane of the
// group calls it with the same key; all lanes of the group return the same
// status.  Groups of one wave may call it with different keys (different
// candidates) and run the probe loop independently under the exec mask.
//
// Protocol: the first word of a slot is claimed by a 64-bit CAS from 0; the
// claimer then publishes the remaining WT-1 words with agent-scope stores.
// A reader whose first word matches but sees unpublished (zero) words
// re-reads them with agent-scope loads until they appear.  First words never
// change once written, so a stale (cached) view of a bucket can only show a
// slot as emptier than it is; a CAS on such a slot fails and returns the
// real first word, which replaces the stale one.
//
// `b0` is the key's first bucket, bucket_of(h, 0, T.nbuckets); `rehash()`
// returns h, for the later probes only (callers that computed the first
// bucket elsewhere need not keep h live).
template <int WT, int BW, typename Rehash>
__device__ __forceinline__ int find_or_put_group_b(const Table &T, const uint64_t (&v)[WT],
                                                   const SlotFmt<WT> &f, uint64_t b0, Rehash rehash,
                                                   int sub, int group_base, bool active) {
  constexpr int LPB = BW / 2;
  constexpr int SPB = BW / WT;
  static_assert(SPB >= 1, "slot wider than bucket");
  int status = active ? kPending : kFound;
  int probe = 0;
  int load = 1;  // 0: keep registers, 1: plain load, 2: coherent load
  int spins = 0;
  uint64_t w0 = 0, w1 = 0;
  uint64_t *bucket = nullptr;
  const int x0 = 2 * sub, x1 = 2 * sub + 1;
  const int s_of0 = x0 / WT, j_of0 = x0 % WT;
  const int s_of1 = x1 / WT, j_of1 = x1 % WT;
  uint64_t e0 = 0, c0 = 0, e1 = 0, c1 = 0;  // expected word / compare mask for my two words
#pragma unroll
  for (int j = 0; j < WT; j++) {
    if (j == j_of0) { e0 = v[j]; c0 = f.cmp[j]; }
    if (j == j_of1) { e1 = v[j]; c1 = f.cmp[j]; }
  }
  const bool in0 = s_of0 < SPB, in1 = s_of1 < SPB;
  if constexpr (WT == 2) {
    // 16-byte slots: lane `sub` holds slot `sub` whole, so the group's match,
    // half-published and empty masks are three ballots (same protocol as the
    // general path below)
    constexpr uint64_t gm = (LPB >= 64) ? ~0ull : ((1ull << LPB) - 1ull);
    while (ballot(status == kPending)) {
      if (status == kPending) {
        if (load) {
          uint64_t bi = b0;
          if (probe) bi = bucket_of(rehash(), probe, T.nbuckets);
          bucket = T.words + bi * BW;
          if (load == 1) {
            const ulonglong2 q = *reinterpret_cast<const ulonglong2 *>(bucket + x0);
            w0 = q.x;
            w1 = q.y;
          } else {
            w0 = load_coherent(bucket + x0);
            w1 = load_coherent(bucket + x1);
          }
        }
        const bool first = ((w0 ^ e0) & c0) == 0;
        const uint64_t mm = (ballot(first && ((w1 ^ e1) & c1) == 0) >> group_base) & gm;
        const uint64_t pm = (ballot(first && w1 == 0) >> group_base) & gm;
        const uint64_t em = (ballot(w0 == 0) >> group_base) & gm;
        if (mm) {
          status = kFound;
        } else if (pm) {
          load = 2;
          if (++spins > kSpinLimit) status = kTimeout;
        } else if (!em) {
          load = 1;
          if (++probe == kHashFunctions) status = kFull;
        } else {
          const int owner = __ffsll(static_cast<unsigned long long>(em)) - 1;
          uint64_t old = 0;
          if (sub == owner) {
            old = atomicCAS(reinterpret_cast<unsigned long long *>(bucket + 2 * owner), 0ull,
                            static_cast<unsigned long long>(v[0]));
            if (old == 0)
              store_coherent(bucket + 2 * owner + 1, v[1]);
            else
              w0 = old;
          }
          old = __shfl(old, group_base + owner, 64);
          if (old == 0)
            status = kInserted;
          else
            load = 0;
        }
      }
    }
    return status;
  }
  while (ballot(status == kPending)) {
    if (status == kPending) {
      if (load) {
        uint64_t bi = b0;
        if (probe) bi = bucket_of(rehash(), probe, T.nbuckets);
        bucket = T.words + bi * BW;
        if (load == 1) {
          const ulonglong2 q = *reinterpret_cast<const ulonglong2 *>(bucket + x0);
          w0 = q.x;
          w1 = q.y;
        } else {
          w0 = load_coherent(bucket + x0);
          w1 = load_coherent(bucket + x1);
        }
      }
      uint32_t eq = 0, em = 0;
      if (in0) {
        eq |= static_cast<uint32_t>((w0 & c0) == (e0 & c0)) << x0;
        em |= static_cast<uint32_t>(w0 == 0) << x0;
      }
      if (in1) {
        eq |= static_cast<uint32_t>((w1 & c1) == (e1 & c1)) << x1;
        em |= static_cast<uint32_t>(w1 == 0) << x1;
      }
#pragma unroll
      for (int d = 1; d < LPB; d <<= 1) {
        eq |= __shfl_xor(eq, d, LPB);
        em |= __shfl_xor(em, d, LPB);
      }
      uint32_t full = 0, partial = 0, empty = 0;
#pragma unroll
      for (int s = 0; s < SPB; s++) {
        const uint32_t sm = ((1u << WT) - 1u) << (s * WT);
        const uint32_t first = 1u << (s * WT);
        const uint32_t later = sm & ~first;
        if (eq & first) {
          if ((eq & sm) == sm)
            full |= 1u << s;
          else if (((eq | em) & later) == later)
            partial |= 1u << s;
        }
        if (em & first) empty |= 1u << s;
      }
      if (full) {
        status = kFound;
      } else if (partial) {
        load = 2;
        if (++spins > kSpinLimit) status = kTimeout;
      } else if (!empty) {
        load = 1;
        if (++probe == kHashFunctions) status = kFull;
      } else {
        const int s = __ffs(empty) - 1;
        const int xw = s * WT;          // first word of the slot
        const int owner = xw >> 1;      // lane (sub) holding it
        uint64_t old = 0;
        if (sub == owner) {
          old = atomicCAS(reinterpret_cast<unsigned long long *>(bucket + xw), 0ull,
                          static_cast<unsigned long long>(v[0]));
          if (old == 0) {
#pragma unroll
            for (int j = 1; j < WT; j++) store_coherent(bucket + xw + j, v[j]);
          }
        }
        old = __shfl(old, group_base + owner, 64);
        if (old == 0) {
          status = kInserted;
        } else {
          if (sub == owner) {
            if ((xw & 1) == 0)
              w0 = old;
            else
              w1 = old;
          }
          load = 0;
        }
      }
    }
  }
  return status;
}

template <int WT, int BW>
__device__ __forceinline__ int find_or_put_group(const Table &T, const uint64_t (&v)[WT],
                                                 const SlotFmt<WT> &f, uint64_t h, int sub,
                                                 int group_base, bool active) {
  return find_or_put_group_b<WT, BW>(T, v, f, bucket_of(h, 0, T.nbuckets), [h]() { return h; }, sub,
                                     group_base, active);
}

// Thread-level FIND: returns 1 and the slot's stamp if the key is present.
// Used while no FIND-OR-PUT runs concurrently, so no slot is half written.
template <int WT>
__device__ __forceinline__ int find_thread(const Table &T, const uint64_t (&v)[WT],
                                           const SlotFmt<WT> &f, uint64_t h, int BW,
                                           uint32_t K, uint32_t *stamp) {
  const int SPB = BW / WT;
  for (int probe = 0; probe < kHashFunctions; probe++) {
    const uint64_t *b = T.words + bucket_of(h, probe, T.nbuckets) * static_cast<uint64_t>(BW);
    bool any_empty = false;
    for (int s = 0; s < SPB; s++) {
      uint64_t w[WT];
#pragma unroll
      for (int j = 0; j < WT; j++) w[j] = b[s * WT + j];
      if (w[0] == 0) {
        any_empty = true;
        continue;
      }
      bool eq = true;
#pragma unroll
      for (int j = 0; j < WT; j++) eq = eq && ((w[j] & f.cmp[j]) == (v[j] & f.cmp[j]));
      if (eq) {
        *stamp = read_stamp<WT>(w, K);
        return 1;
      }
    }
    // keys are placed in the first bucket of their sequence that had room
    if (any_empty) return 0;
  }
  return 0;
}

}  // namespace gx
