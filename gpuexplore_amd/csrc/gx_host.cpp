// gx_host.cpp -- C ABI implementation: model handles, single-model
// exploration driver, visited-set API and device helpers.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/gpuexplore.h"
#include "gx_devbuf.h"
#include "gx_kernels.h"
#include "gx_model.h"

using namespace gx;

namespace {

#define GX_TRY(expr)                      \
  do {                                    \
    hipError_t e_ = (expr);               \
    if (e_ != hipSuccess) return map_hip(e_); \
  } while (0)

int map_hip(hipError_t e) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return GX_E_NOMEM;
  if (e == hipErrorNoDevice) return GX_E_NODEVICE;
  return GX_E_HIP;
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// RAII device buffer

struct Events {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  ~Events() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  hipEvent_t next() {
    if (used == ev.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      ev.push_back(e);
    }
    return ev[used++];
  }
};

// FIND-OR-PUT grid for a level of about front * degree candidates: enough
// waves for ~8 iterations of 16 candidates each, between one and
// kInsertBlocksPerCU workgroups per CU (tiny levels must not pay for
// dispatching thousands of empty workgroups)
uint32_t insert_grid(uint32_t cus, double front, double degree) {
  const double cand = front * std::max(1.0, degree);
  const double waves = cand / (16.0 * 8.0) + 1.0;
  const double blocks = std::ceil(waves / (kInsertThreads / 64));
  return static_cast<uint32_t>(std::min<double>(cus * kInsertBlocksPerCU, std::max<double>(cus, blocks)));
}

// successor grid for a frontier of `front` states: about two 64-state
// batches per wave, between one workgroup per CU and `max_blocks`
uint32_t expand_grid(uint32_t cus, uint32_t threads, uint32_t max_blocks, double front) {
  const double waves = front / (64.0 * 2.0) + 1.0;
  const double blocks = std::ceil(waves / (threads / 64));
  return static_cast<uint32_t>(std::min<double>(max_blocks, std::max<double>(cus, blocks)));
}

uint32_t round_threads(uint32_t t) {
  if (t == 0) t = 256;
  t = (t + 63) / 64 * 64;
  return std::min<uint32_t>(512, std::max<uint32_t>(64, t));
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int gx_model_load_gpf_text(const char *text, size_t len, int por, gx_model **out) {
  if (!text || !out) return GX_E_ARG;
  gx_model *m = new (std::nothrow) gx_model();
  if (!m) return GX_E_NOMEM;
  m->por = por != 0;
  int rc = parse_gpf(std::string(text, len), m->por, &m->gpf);
  if (rc == GX_OK) rc = compile_model(m->gpf, &m->comp);
  if (rc != GX_OK) {
    delete m;
    return rc;
  }
  *out = m;
  return GX_OK;
}

extern "C" int gx_model_load_gpf(const char *path_noext, int por, gx_model **out) {
  if (!path_noext || !out) return GX_E_ARG;
  std::string fn = std::string(path_noext) + ".gpf";  // cu:1814-1815
  std::ifstream f(fn, std::ios::binary);
  if (!f) return GX_E_NOFILE;
  std::stringstream ss;
  ss << f.rdbuf();
  std::string t = ss.str();
  return gx_model_load_gpf_text(t.data(), t.size(), por, out);
}

extern "C" void gx_model_free(gx_model *m) { delete m; }

extern "C" int gx_model_info_get(const gx_model *m, gx_model_info *o) {
  if (!m || !o) return GX_E_ARG;
  const Gpf &g = m->gpf;
  o->nr_procs = g.nr_procs;
  o->bits_act = g.bits_act;
  o->proc_nrstates = g.proc_nrstates;
  o->bits_statevector = g.bits_statevector;
  o->sv_nints = g.sv_nints();
  o->liveness = g.liveness;
  o->max_buf_ints = g.max_buf_ints;
  o->nr_trans = static_cast<uint32_t>(g.proc_trans.size());
  o->state_bits = m->comp.hdr.state_bits;
  o->key_words = m->comp.hdr.key_words;
  return GX_OK;
}

extern "C" void gx_options_default(gx_options *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->kernel_iters = 1;
  o->bucket_bytes = 0;
  o->hash_seed = 0x5eed5eed5eedull;
  o->dedup = 0;  // on a 20-process ring it drops 15 % of the lookups, but not the kernel's time
}

extern "C" const char *gx_strerror(int s) {
  switch (s) {
    case GX_OK: return "ok";
    case GX_E_NOFILE: return "input network does not exist";
    case GX_E_FORMAT: return "malformed gpf input";
    case GX_E_RANGE: return "gpf field out of range";
    case GX_E_NOMEM: return "out of memory";
    case GX_E_HIP: return "HIP runtime error";
    case GX_E_NODEVICE: return "no HIP device detected";
    case GX_E_ARG: return "invalid argument";
    case GX_E_LIMIT: return "model exceeds a supported limit";
    case GX_E_STATE: return "reached a local state without transition offsets";
    case GX_E_CAPACITY: return "frontier or successor buffer exhausted";
    case GX_E_SCRATCH: return "input larger than the no-allocation scratch buffer";
    default: return "unknown error";
  }
}

// ---------------------------------------------------------------------------
// Single-model exploration (cu:1989-2142, 2168-2172).  An explorer owns the
// device buffers of one model and configuration; every run resets the
// visited set and explores from the initial state.
struct gx_explorer {
  const gx_model *m = nullptr;
  gx_options opt{};
  int W = 1, WT = 1, BW = 16, mode = kModePlain;
  bool stamp = false, lds = false, per_state = false;
  uint32_t threads = 256, blocks = 1, cus = 1;
  uint32_t ins_blocks = 1;  // FIND-OR-PUT launch: ins_blocks x kInsertThreads
  uint64_t nbuckets = 0, table_bytes = 0, front_cap = 0, cand_cap = 0;
  DevBuf d_table, d_front0, d_front1, d_cand, d_retry0, d_retry1, d_ctrl, d_model, d_count;
  hipStream_t st = nullptr;
  Events evs[2];        // per batch of levels in flight
  Ctrl *h_ctrl = nullptr;  // pinned: the control block after each batch, two slots
  Ctrl last{};       // control block at the end of the last run
  bool ran = false;
  ~gx_explorer() {
    if (st) (void)hipStreamSynchronize(st);  // no copy may still target h_ctrl
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Lists the visited set in the reference's format (cu:544-568, 594-612):
// per stored state an optional "new: " (not yet expanded), the local states
// "p0,p1,...", a space, the sv_nints 32-bit words of the packed vector with
// the flag bits stripped ("%d " each), a newline; then the summary line.
// States come in table order.  Unexplored = the frontier that the next
// level would expand.  The table is streamed to the host in slices.
int print_states(gx_explorer *e, FILE *f, const Ctrl &ch) {
  const Compiled &c = e->m->comp;
  const Gpf &g = e->m->gpf;
  const uint32_t N = g.nr_procs, nw32 = g.sv_nints();
  const int W = e->W, WT = e->WT, BW = e->BW;
  const int SPB = BW / WT;
  std::vector<uint32_t> vec(nw32 ? nw32 : 1);
  auto words32 = [&](const uint64_t *key) {
    key_to_words32(c, key, vec.data(), nw32);
    return std::string(reinterpret_cast<const char *>(vec.data()), nw32 * 4);
  };
  // unexplored states: the current frontier
  std::unordered_set<std::string> fresh;
  const uint64_t nf = ch.front_count[ch.cur];
  if (nf) {
    std::vector<uint64_t> fk(nf * W);
    const DevBuf &fb = ch.cur ? e->d_front1 : e->d_front0;
    GX_TRY(hipMemcpyAsync(fk.data(), fb.p, nf * W * 8, hipMemcpyDeviceToHost, e->st));
    GX_TRY(hipStreamSynchronize(e->st));
    for (uint64_t i = 0; i < nf; i++) fresh.insert(words32(&fk[i * W]));
  }
  const uint64_t slice = 1u << 20;  // buckets per copy
  std::vector<uint64_t> buf(slice * BW);
  uint64_t count = 0, newcount = 0;
  for (uint64_t b0 = 0; b0 < e->nbuckets; b0 += slice) {
    const uint64_t nb = std::min<uint64_t>(slice, e->nbuckets - b0);
    GX_TRY(hipMemcpyAsync(buf.data(), e->d_table.as<uint64_t>() + b0 * BW, nb * BW * 8, hipMemcpyDeviceToHost, e->st));
    GX_TRY(hipStreamSynchronize(e->st));
    for (uint64_t b = 0; b < nb; b++)
      for (int s = 0; s < SPB; s++) {
        const uint64_t *slot = &buf[b * BW + s * WT];
        if (slot[0] == 0) continue;
        count++;
        const std::string w = words32(slot);
        if (fresh.count(w)) {
          newcount++;
          fputs("new: ", f);
        }
        for (uint32_t p = 0; p < N; p++) fprintf(f, p + 1 < N ? "%u," : "%u", key_get(c, slot, p));
        fputc(' ', f);
        for (uint32_t j = 0; j < nw32; j++) fprintf(f, "%d ", static_cast<int32_t>(vec[j]));
        fputc('\n', f);
      }
  }
  fprintf(f, "nr. of states in hash table: %llu (%llu unexplored states)\n", static_cast<unsigned long long>(count),
          static_cast<unsigned long long>(newcount));
  return GX_OK;
}

}  // namespace

extern "C" int gx_explorer_create(const gx_model *m, const gx_options *opt, gx_explorer **out) {
  if (!m || !opt || !out) return GX_E_ARG;
  if (device_count() == 0) return GX_E_NODEVICE;
  const ModelHdr &H = m->comp.hdr;
  const int W = static_cast<int>(H.key_words);
  const bool proviso = opt->por && opt->cycle_proviso;
  if (opt->por_method > GX_POR_STUBBORN) return GX_E_ARG;
  if (opt->property < GX_PROP_NONE || opt->property > GX_PROP_LIVENESS) return GX_E_ARG;
  const uint32_t K = H.state_bits;
  const int WT = proviso ? std::max<int>(1, static_cast<int>((K + 32 + 62) / 63)) : W;
  if (WT > kMaxKeyWords || W > kMaxKeyWords) return GX_E_LIMIT;
  // bucket width: one 64-byte half line holds 4-8 slots of up to 2 words,
  // wider slots get a full 128-byte line
  const uint32_t bb = opt->bucket_bytes ? opt->bucket_bytes : (WT <= 2 ? 64u : 128u);
  if (bb != 64 && bb != 128 && bb != 256) return GX_E_ARG;
  const int BW = static_cast<int>(bb / 8);
  if (BW / WT < 1) return GX_E_ARG;
  gx_explorer *e = new (std::nothrow) gx_explorer();
  if (!e) return GX_E_NOMEM;
  std::unique_ptr<gx_explorer> guard(e);
  e->m = m;
  e->opt = *opt;
  e->W = W;
  e->WT = WT;
  e->BW = BW;
  e->stamp = proviso;
  e->mode = opt->por ? (proviso ? kModePorProviso : kModePor) : kModePlain;
  e->threads = round_threads(opt->threads);
  int dev = 0;
  GX_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GX_TRY(hipGetDeviceProperties(&prop, dev));
  e->cus = static_cast<uint32_t>(prop.multiProcessorCount);
  // default grid: 8192 threads per CU (measured best for the state-per-lane
  // successor kernel on a 256-CU part; grid-stride loops take any size)
  e->blocks = opt->blocks ? std::max<uint32_t>(1, opt->blocks) : e->cus * (8192 / e->threads);
  // random bucket reads are latency bound: a deep grid keeps every SIMD at
  // its register-limited wave count (sized per level in gx_explorer_run)
  e->ins_blocks = opt->insert_blocks ? opt->insert_blocks : e->cus * kInsertBlocksPerCU;
  size_t free_b = 0, total_b = 0;
  GX_TRY(hipMemGetInfo(&free_b, &total_b));
  // -q in 32-bit ints (cu:1838-1842); the default is a quarter of free
  // memory, at most 8 GiB (the reference takes all of it, cu:2054-2066)
  uint64_t table_bytes = opt->q_ints ? opt->q_ints * 4ull : std::min<uint64_t>(free_b / 4, 8ull << 30);
  e->nbuckets = table_bytes / bb;
  if (e->nbuckets == 0) return GX_E_ARG;
  e->table_bytes = e->nbuckets * bb;
  const uint64_t slots = e->nbuckets * static_cast<uint64_t>(BW / WT);
  // buffers from what the visited set leaves: frontiers are bounded by the
  // slot count.  The visited set itself is allocated last: random bucket
  // reads ran 5 % slower in the first device memory a process allocates (a
  // 2^34-int set on config 2: FIND-OR-PUT 79.8 ms as the first allocation,
  // 75.5 ms behind a 64 GiB one, the same box and process)
  free_b = free_b > e->table_bytes ? free_b - e->table_bytes : 0;
  const uint64_t per_front = static_cast<uint64_t>(W) * 8ull + 8ull;  // key + retry index
  uint64_t front_cap = opt->frontier_capacity;
  if (!front_cap) front_cap = std::min<uint64_t>(slots, (free_b / 4) / (2 * per_front));
  e->front_cap = std::max<uint64_t>(front_cap, 64);
  uint64_t cand_cap = opt->cand_capacity;
  if (!cand_cap) cand_cap = std::min<uint64_t>(std::max<uint64_t>(4 * slots, 1u << 20), (free_b / 4) / (W * 8ull));
  e->cand_cap = std::max<uint64_t>(std::min<uint64_t>(cand_cap, 0xFFFFFFFFull), 64);
  GX_TRY(e->d_front0.alloc(e->front_cap * W * 8));
  GX_TRY(e->d_front1.alloc(e->front_cap * W * 8));
  GX_TRY(e->d_retry0.alloc(e->front_cap * 8));
  GX_TRY(e->d_retry1.alloc(e->front_cap * 8));
  GX_TRY(e->d_cand.alloc(e->cand_cap * W * 8));
  GX_TRY(e->d_ctrl.alloc(sizeof(Ctrl)));
  GX_TRY(e->d_model.alloc(m->comp.words.size() * 4));
  GX_TRY(e->d_count.alloc(8));
  GX_TRY(e->d_table.alloc(e->table_bytes));
  GX_TRY(hipHostMalloc(reinterpret_cast<void **>(&e->h_ctrl), 2 * sizeof(Ctrl), hipHostMallocDefault));
  GX_TRY(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
  GX_TRY(hipMemcpyAsync(e->d_model.p, m->comp.words.data(), m->comp.words.size() * 4, hipMemcpyHostToDevice,
                        e->st));
  GX_TRY(hipStreamSynchronize(e->st));
  // successor kernel: one state per lane for plain exploration of models
  // whose labels fit the 64-bit row masks, else the process-group kernel
  const size_t lds_limit = 64 * 1024;
  const size_t model_b = ((H.total_words + 3u) & ~3u) * 4u;
  const size_t sec_b = ((H.total_words - H.tps_begin + 3u) & ~3u) * 4u;  // state-per-lane section
  const size_t tps_b = expand_tps_scratch_bytes(e->threads, H.nprocs, H.nrows < 0xFFFFu);
  if (opt->expand_kernel > 2) return GX_E_ARG;
  e->per_state = e->mode == kModePlain && H.lab64 && opt->expand_kernel != 1 && tps_b <= lds_limit;
  if (opt->expand_kernel == 2 && !e->per_state) return GX_E_ARG;
  if (e->per_state)
    e->lds = sec_b + tps_b <= lds_limit;
  else
    e->lds = model_b + expand_scratch_bytes(e->threads) <= 48 * 1024;
  *out = guard.release();
  return GX_OK;
}

extern "C" void gx_explorer_free(gx_explorer *e) { delete e; }

extern "C" int gx_explorer_dump(gx_explorer *e, const char *path) {
  if (!e) return GX_E_ARG;
  if (!e->ran) return GX_E_ARG;
  FILE *f = path ? fopen(path, "w") : stdout;
  if (!f) return GX_E_NOFILE;
  const int rc = print_states(e, f, e->last);
  if (path) fclose(f);
  else fflush(stdout);
  return rc;
}

extern "C" int gx_explorer_run(gx_explorer *e, gx_result *res) {
  if (!e || !res) return GX_E_ARG;
  memset(res, 0, sizeof *res);
  const auto t_init0 = std::chrono::steady_clock::now();
  const gx_options *opt = &e->opt;
  const ModelHdr &H = e->m->comp.hdr;
  const int W = e->W, WT = e->WT, BW = e->BW;
  hipStream_t st = e->st;
  GX_TRY(hipMemsetAsync(e->d_table.p, 0, e->table_bytes, st));  // cu:656-663 init_queue
  Ctrl c0{};
  c0.cur = 1;
  c0.level = 0xFFFFFFFFu;
  c0.levels_done = 0xFFFFFFFFu;
  c0.cand_count = 1;
  c0.max_states = opt->max_states;
  GX_TRY(hipMemcpyAsync(e->d_ctrl.p, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  std::vector<uint64_t> s0(W, kFlag);  // all processes in local state 0 (cu:668-682)
  GX_TRY(hipMemcpyAsync(e->d_cand.p, s0.data(), W * 8, hipMemcpyHostToDevice, st));

  Table T{e->d_table.as<uint64_t>(), e->nbuckets, opt->hash_seed};
  ExpandArgs ea{};
  ea.model = e->d_model.as<uint32_t>();
  ea.hdr = H;
  ea.ctrl = e->d_ctrl.as<Ctrl>();
  ea.front[0] = e->d_front0.as<uint64_t>();
  ea.front[1] = e->d_front1.as<uint64_t>();
  ea.cand = e->d_cand.as<uint64_t>();
  ea.cand_cap = e->cand_cap;
  ea.retry[0] = e->d_retry0.as<uint64_t>();
  ea.retry[1] = e->d_retry1.as<uint64_t>();
  ea.property = opt->property;
  ea.bucket_words = BW;
  ea.table = T;
  ea.per_state = e->per_state ? 1 : 0;
  ea.por_method = static_cast<int>(opt->por_method);
  InsertArgs ia{};
  ia.ctrl = e->d_ctrl.as<Ctrl>();
  ia.cand = e->d_cand.as<uint64_t>();
  ia.cand_cap = e->cand_cap;
  ia.front[0] = e->d_front0.as<uint64_t>();
  ia.front[1] = e->d_front1.as<uint64_t>();
  ia.front_cap = e->front_cap;
  ia.table = T;
  ia.state_bits = H.state_bits;
  ia.is_new = nullptr;
  ia.append = 1;
  ia.dedup = opt->dedup ? 1 : 0;

  // initial state: FIND-OR-PUT of s0 into frontier 0 with stamp 0
  GX_TRY(launch_insert(ia, W, WT, BW, e->stamp, 1, 64, st));
  GX_TRY(launch_advance(ia.ctrl, st));
  // level passes hand successors over as 12-byte candidates where keys have
  // two words of at most kPackedMaxBits state bits (the initial state above
  // went in the 16-byte form)
  {
    const int pk = (e->mode == kModePlain && e->per_state && W == 2 && WT == 2 && !e->stamp && !opt->dedup &&
                    H.state_bits <= kPackedMaxBits) ? 1 : 0;
    ea.packed = pk;
    ia.packed = pk;
  }
  GX_TRY(hipStreamSynchronize(st));
  const auto t_init1 = std::chrono::steady_clock::now();
  res->t_init_s = std::chrono::duration<double>(t_init1 - t_init0).count();

  const uint32_t kiters = std::max<uint32_t>(1, opt->kernel_iters);
  // Without per-level output the host keeps one more batch of levels
  // enqueued while it reads the control block of the previous one, so the
  // GPU does not idle through the read-back.  A batch enqueued after the
  // last level finds `stop` set: every level kernel then returns at once.
  const bool ahead = opt->verbosity == 0;
  struct Batch {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ins, exp;
    hipEvent_t done = nullptr;
  } batch[2];
  double hash_ms = 0, exp_ms = 0;
  Ctrl ch{};
  // frontier and out-degree estimates for the launch geometry
  double front = 1.0, prev_front = 1.0, degree = 16.0;
  auto enqueue = [&](int k, double front_est) -> int {
    Batch &b = batch[k];
    e->evs[k].used = 0;
    b.ins.clear();
    b.exp.clear();
    const uint32_t ins_blocks = opt->insert_blocks ? opt->insert_blocks
                                : kiters == 1 ? insert_grid(e->cus, front_est, degree)
                                              : e->cus * kInsertBlocksPerCU;
    const uint32_t exp_blocks = (opt->blocks || kiters != 1) ? e->blocks
                                : expand_grid(e->cus, e->threads, e->blocks, front_est);
    for (uint32_t it = 0; it < kiters; it++) {
      hipEvent_t a0 = e->evs[k].next(), a1 = e->evs[k].next(), b0 = e->evs[k].next(), b1 = e->evs[k].next();
      if (!a0 || !a1 || !b0 || !b1) return GX_E_HIP;
      GX_TRY(hipEventRecord(a0, st));
      GX_TRY(launch_expand(ea, e->mode, W, WT, e->lds, exp_blocks, e->threads, st));
      GX_TRY(hipEventRecord(a1, st));
      GX_TRY(hipEventRecord(b0, st));
      GX_TRY(launch_insert(ia, W, WT, BW, e->stamp, ins_blocks, kInsertThreads, st));
      GX_TRY(hipEventRecord(b1, st));
      GX_TRY(launch_advance(ia.ctrl, st));
      b.exp.push_back({a0, a1});
      b.ins.push_back({b0, b1});
    }
    GX_TRY(hipMemcpyAsync(e->h_ctrl + k, e->d_ctrl.p, kCtrlHead, hipMemcpyDeviceToHost, st));
    b.done = e->evs[k].next();
    if (!b.done) return GX_E_HIP;
    GX_TRY(hipEventRecord(b.done, st));
    return GX_OK;
  };
  auto collect = [&](int k) -> int {
    Batch &b = batch[k];
    GX_TRY(hipEventSynchronize(b.done));
    for (auto &p : b.exp) {
      float ms = 0;
      GX_TRY(hipEventElapsedTime(&ms, p.first, p.second));
      exp_ms += ms;
    }
    for (auto &p : b.ins) {
      float ms = 0;
      GX_TRY(hipEventElapsedTime(&ms, p.first, p.second));
      hash_ms += ms;
    }
    res->hash_launches += b.ins.size();
    return GX_OK;
  };
  // the frontier one batch further on, extrapolated from the last growth
  auto ahead_front = [&]() {
    const double g = std::min(8.0, std::max(0.125, front / std::max(1.0, prev_front)));
    return std::max(1.0, front * std::pow(g, static_cast<double>(kiters)));
  };
  const auto t0 = std::chrono::steady_clock::now();
  int cur = 0;
  // an error leaves no batch in flight behind it
  auto fail = [&](int rc) {
    (void)hipStreamSynchronize(st);
    return rc;
  };
  {
    const int rc = enqueue(0, front);
    if (rc != GX_OK) return fail(rc);
  }
  if (ahead) {
    const int rc = enqueue(1, front);
    if (rc != GX_OK) return fail(rc);
  }
  for (;;) {
    int rc = collect(cur);
    if (rc != GX_OK) return fail(rc);
    ch = e->h_ctrl[cur];
    prev_front = front;
    front = static_cast<double>(ch.front_count[ch.cur]);
    if (ch.expanded) degree = static_cast<double>(ch.transitions) / static_cast<double>(ch.expanded);
    if (opt->verbosity == 1) printf("%u\n", ch.levels_done);  // cu:2129-2131
    if (opt->verbosity == 2) {  // cu:2132-2135, 631-651
      GX_TRY(hipMemsetAsync(e->d_count.p, 0, 8, st));
      GX_TRY(launch_count(T, WT, BW, e->d_count.as<uint64_t>(), e->cus * 8, st));
      uint64_t n = 0;
      GX_TRY(hipMemcpyAsync(&n, e->d_count.p, 8, hipMemcpyDeviceToHost, st));
      GX_TRY(hipStreamSynchronize(st));
      printf("nr. of states in hash table: %llu (%llu unexplored states)\n", static_cast<unsigned long long>(n),
             static_cast<unsigned long long>(ch.front_count[ch.cur]));
    }
    if (opt->verbosity == 3) {  // cu:2136-2139
      const int prc = print_states(e, stdout, ch);
      if (prc != GX_OK) return fail(prc);
    }
    if (ch.stop) {
      if (ahead) {  // the batch behind it found stop set and did nothing
        rc = collect(cur ^ 1);
        if (rc != GX_OK) return fail(rc);
      }
      break;
    }
    if (ahead) {
      rc = enqueue(cur, ahead_front());
      if (rc != GX_OK) return fail(rc);
      cur ^= 1;
    } else {
      rc = enqueue(cur, front);
      if (rc != GX_OK) return fail(rc);
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  res->t_explore_s = std::chrono::duration<double>(t1 - t0).count();
  res->t_hash_s = hash_ms / 1e3;
  res->t_expand_s = exp_ms / 1e3;

  GX_TRY(hipMemsetAsync(e->d_count.p, 0, 8, st));
  GX_TRY(launch_count(T, WT, BW, e->d_count.as<uint64_t>(), e->cus * 8, st));
  uint64_t counted = 0;
  GX_TRY(hipMemcpyAsync(&counted, e->d_count.p, 8, hipMemcpyDeviceToHost, st));
  GX_TRY(hipStreamSynchronize(st));

  e->last = ch;
  e->ran = true;
  res->states = counted;
  res->transitions = ch.transitions;
  res->levels = ch.level;  // levels completed
  res->lookups = ch.lookups;
  res->inserts = ch.inserts;
  res->violation = ch.violation ? 1 : 0;
  res->table_full = ch.table_full ? 1 : 0;
  res->truncated = ch.truncated ? 1 : 0;
  res->table_bytes = e->table_bytes;
  if (ch.bad_state) return GX_E_STATE;
  if (ch.overflow) return GX_E_CAPACITY;
  if (ch.spin_timeout) return GX_E_HIP;
  return GX_OK;
}

extern "C" int gx_explore(const gx_model *m, const gx_options *opt, gx_result *res) {
  if (!m || !opt || !res) return GX_E_ARG;
  memset(res, 0, sizeof *res);
  const auto t0 = std::chrono::steady_clock::now();
  gx_explorer *e = nullptr;
  int rc = gx_explorer_create(m, opt, &e);
  if (rc != GX_OK) return rc;
  const double t_alloc = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  rc = gx_explorer_run(e, res);
  res->t_init_s += t_alloc;
  gx_explorer_free(e);
  return rc;
}

// ---------------------------------------------------------------------------
// Visited-set API
struct gx_table {
  DevBuf words, ctrl, cand, status;
  uint64_t nbuckets = 0, cand_cap = 0;
  uint32_t W = 1, BW = 16;
  uint32_t cus = 1;       // compute units of the table's device
  int dev = 0;            // the device the table lives on
  void *scratch = nullptr;  // that device's no-allocation scratch (device address)
  uint64_t seed = 0;
  hipStream_t st = nullptr;
  ~gx_table() {
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
};

namespace {

// Scratch of gx_table_insert_host_noalloc: the control block, the staged keys
// and their is-new bytes, in one file-scope device array.  Every device that
// loads the library holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in, never the device memory: the device
// address comes from hipGetSymbolAddress, once per device (scratch_of).
constexpr size_t kScratchCtrl = (sizeof(Ctrl) + 255) / 256 * 256;
constexpr size_t kScratchKeys = static_cast<size_t>(GX_NOALLOC_MAX_KEYS) * kMaxKeyWords * 8;
constexpr size_t kScratchBytes = kScratchCtrl + kScratchKeys + GX_NOALLOC_MAX_KEYS;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];  // serialises the calls that use a device's scratch
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_init_mu;

// device address of `scratch` on device `dev` (current), resolved once
int scratch_of(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices) return GX_E_LIMIT;
  std::lock_guard<std::mutex> g(g_scratch_init_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    GX_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return GX_E_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GX_OK;
}

}  // namespace

extern "C" int gx_table_create(uint64_t q_ints, uint32_t key_words, uint32_t bucket_bytes,
                               uint64_t hash_seed, gx_table **out) {
  if (!out || key_words < 1 || key_words > static_cast<uint32_t>(kMaxKeyWords)) return GX_E_ARG;
  if (bucket_bytes == 0) bucket_bytes = 128;
  if (bucket_bytes != 64 && bucket_bytes != 128 && bucket_bytes != 256) return GX_E_ARG;
  if (bucket_bytes / 8 < key_words) return GX_E_ARG;
  if (device_count() == 0) return GX_E_NODEVICE;
  gx_table *t = new (std::nothrow) gx_table();
  if (!t) return GX_E_NOMEM;
  std::unique_ptr<gx_table> guard(t);
  t->W = key_words;
  t->BW = bucket_bytes / 8;
  t->seed = hash_seed;
  t->nbuckets = q_ints * 4ull / bucket_bytes;
  if (t->nbuckets == 0) return GX_E_ARG;
  GX_TRY(hipGetDevice(&t->dev));
  hipDeviceProp_t prop;
  GX_TRY(hipGetDeviceProperties(&prop, t->dev));
  t->cus = static_cast<uint32_t>(prop.multiProcessorCount);
  // set-up of the no-allocation entry point on this device
  const int src = scratch_of(t->dev, &t->scratch);
  if (src != GX_OK) return src;
  if (hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking) != hipSuccess) return GX_E_HIP;
  if (t->words.alloc(t->nbuckets * bucket_bytes) != hipSuccess) return GX_E_NOMEM;
  if (t->ctrl.alloc(sizeof(Ctrl)) != hipSuccess) return GX_E_NOMEM;
  GX_TRY(hipMemsetAsync(t->words.p, 0, t->nbuckets * bucket_bytes, t->st));
  GX_TRY(hipStreamSynchronize(t->st));
  *out = guard.release();
  return GX_OK;
}

extern "C" void gx_table_free(gx_table *t) { delete t; }

// One FIND-OR-PUT pass of n keys (device memory) with control block `ctrl`,
// is-new bytes to dev_status (optional).  set_flag: the keys come from the
// caller, so bit 63 of every word is set on the device.  kernel_ms: timed
// with HIP events when non-NULL (events are created only then).
static int table_insert(gx_table *t, Ctrl *ctrl, const uint64_t *dev_keys, uint64_t n, uint8_t *dev_status,
                        bool set_flag, uint64_t *inserted, float *kernel_ms) {
  InsertArgs ia{};
  ia.ctrl = ctrl;
  ia.cand = dev_keys;
  ia.cand_cap = n;
  ia.table = Table{t->words.as<uint64_t>(), t->nbuckets, t->seed};
  ia.state_bits = 63 * t->W;
  ia.is_new = dev_status;
  ia.append = 0;
  ia.set_flag = set_flag ? 1 : 0;
  GX_TRY(launch_ctrl_reset(ctrl, n, t->st));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (kernel_ms) {
    GX_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return GX_E_HIP;
    }
    (void)hipEventRecord(e0, t->st);
  }
  hipError_t le = launch_insert(ia, t->W, t->W, t->BW, false, t->cus * kInsertBlocksPerCU, kInsertThreads, t->st);
  if (kernel_ms) (void)hipEventRecord(e1, t->st);
  if (le == hipSuccess) le = launch_fold(ctrl, t->st);
  Ctrl c{};
  hipError_t ce = hipMemcpyAsync(&c, ctrl, kCtrlHead, hipMemcpyDeviceToHost, t->st);
  hipError_t se = hipStreamSynchronize(t->st);
  if (kernel_ms) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *kernel_ms = ms;
  }
  GX_TRY(le);
  GX_TRY(ce);
  GX_TRY(se);
  if (inserted) *inserted = c.inserts;
  if (c.spin_timeout) return GX_E_HIP;
  if (c.table_full) return GX_E_CAPACITY;
  return GX_OK;
}

extern "C" int gx_table_insert_device(gx_table *t, const uint64_t *dev_keys, uint64_t n,
                                      uint64_t *inserted, float *kernel_ms) {
  if (!t || (!dev_keys && n)) return GX_E_ARG;
  return table_insert(t, t->ctrl.as<Ctrl>(), dev_keys, n, nullptr, true, inserted, kernel_ms);
}

extern "C" int gx_table_insert_host(gx_table *t, const uint64_t *keys, uint64_t n, uint8_t *is_new,
                                    uint64_t *inserted) {
  if (!t || (!keys && n)) return GX_E_ARG;
  DevBuf k, s;
  GX_TRY(k.alloc(n * t->W * 8));
  GX_TRY(s.alloc(n));
  std::vector<uint64_t> tmp(keys, keys + n * t->W);
  for (auto &w : tmp) w |= kFlag;
  GX_TRY(hipMemcpyAsync(k.p, tmp.data(), n * t->W * 8, hipMemcpyHostToDevice, t->st));
  int rc = table_insert(t, t->ctrl.as<Ctrl>(), k.as<uint64_t>(), n, s.as<uint8_t>(), false, inserted, nullptr);
  if (is_new && n) {
    GX_TRY(hipMemcpyAsync(is_new, s.p, n, hipMemcpyDeviceToHost, t->st));
    GX_TRY(hipStreamSynchronize(t->st));
  }
  return rc;
}

extern "C" int gx_table_insert_host_noalloc(gx_table *t, const uint64_t *keys, uint64_t n, uint8_t *is_new,
                                            uint64_t *inserted) {
  if (!t || (!keys && n)) return GX_E_ARG;
  if (n > GX_NOALLOC_MAX_KEYS) return GX_E_SCRATCH;
  int dev = 0;
  GX_TRY(hipGetDevice(&dev));
  if (dev != t->dev || !t->scratch) return GX_E_ARG;  // the table's scratch is on its own device
  std::lock_guard<std::mutex> g(g_scratch_mu[dev]);
  unsigned char *base = static_cast<unsigned char *>(t->scratch);
  Ctrl *ctrl = reinterpret_cast<Ctrl *>(base);
  uint64_t *dkeys = reinterpret_cast<uint64_t *>(base + kScratchCtrl);
  uint8_t *dstat = base + kScratchCtrl + kScratchKeys;
  if (n) GX_TRY(hipMemcpyAsync(dkeys, keys, n * t->W * 8, hipMemcpyHostToDevice, t->st));
  const int rc = table_insert(t, ctrl, dkeys, n, dstat, true, inserted, nullptr);
  if (is_new && n) {
    GX_TRY(hipMemcpyAsync(is_new, dstat, n, hipMemcpyDeviceToHost, t->st));
    GX_TRY(hipStreamSynchronize(t->st));
  }
  return rc;
}

extern "C" int gx_table_count(gx_table *t, uint64_t *count) {
  if (!t || !count) return GX_E_ARG;
  DevBuf d;
  GX_TRY(d.alloc(8));
  GX_TRY(hipMemsetAsync(d.p, 0, 8, t->st));
  GX_TRY(launch_count(Table{t->words.as<uint64_t>(), t->nbuckets, t->seed}, t->W, t->BW,
                      d.as<uint64_t>(), 1024, t->st));
  GX_TRY(hipMemcpyAsync(count, d.p, 8, hipMemcpyDeviceToHost, t->st));
  GX_TRY(hipStreamSynchronize(t->st));
  return GX_OK;
}

// ---------------------------------------------------------------------------
extern "C" int gx_device_count(int *count) {
  if (!count) return GX_E_ARG;
  *count = device_count();
  return GX_OK;
}

extern "C" int gx_device_alloc(uint64_t bytes, void **p) {
  if (!p) return GX_E_ARG;
  if (device_count() == 0) return GX_E_NODEVICE;
  GX_TRY(hipMalloc(p, bytes ? bytes : 8));
  return GX_OK;
}

extern "C" int gx_device_free(void *p) {
  GX_TRY(hipFree(p));
  return GX_OK;
}

extern "C" int gx_copy_to_device(void *dst, const void *src, uint64_t bytes) {
  GX_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GX_OK;
}

extern "C" int gx_copy_to_host(void *dst, const void *src, uint64_t bytes) {
  GX_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GX_OK;
}

namespace {

// Average time of `reps` launches of `launch` on the null stream after one
// warm-up, in milliseconds.
template <typename Launch>
int time_launches(int reps, Launch &&launch, double *ms_out) {
  GX_TRY(launch());  // warm-up
  hipEvent_t e0, e1;
  GX_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return GX_E_HIP;
  }
  hipError_t err = hipEventRecord(e0, nullptr);
  for (int i = 0; i < reps && err == hipSuccess; i++) err = launch();
  if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  GX_TRY(err);
  *ms_out = static_cast<double>(ms) / reps;
  return GX_OK;
}

uint32_t device_cus() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
  return static_cast<uint32_t>(prop.multiProcessorCount);
}

}  // namespace

extern "C" int gx_bench_copy(uint64_t bytes, int reps, double *gbps) {
  if (!gbps || bytes < 16 || reps < 1) return GX_E_ARG;
  if (device_count() == 0) return GX_E_NODEVICE;
  const uint64_t n = bytes / 16;
  DevBuf a, b;
  GX_TRY(a.alloc(n * 16));
  GX_TRY(b.alloc(n * 16));
  GX_TRY(hipMemset(a.p, 1, n * 16));
  const uint32_t cus = device_cus();
  // the best of a few launch shapes: workgroups per CU x non-temporal hints
  double best = 0;
  for (uint32_t per_cu : {8u, 16u, 32u, 64u}) {
    for (bool nt : {true, false}) {
      double ms = 0;
      const int rc = time_launches(
          reps, [&]() { return launch_copy(a.as<uint4>(), b.as<uint4>(), n, cus * per_cu, nt, nullptr); }, &ms);
      if (rc != GX_OK) return rc;
      best = std::max(best, 2.0 * static_cast<double>(n * 16) / (ms / 1e3) / 1e9);
    }
  }
  *gbps = best;
  return GX_OK;
}

extern "C" int gx_bench_read(uint64_t bytes, int reps, double *gbps) {
  if (!gbps || bytes < 16 || reps < 1) return GX_E_ARG;
  if (device_count() == 0) return GX_E_NODEVICE;
  const uint64_t n = bytes / 16;
  DevBuf a, sink;
  GX_TRY(a.alloc(n * 16));
  GX_TRY(sink.alloc(4));
  GX_TRY(hipMemset(a.p, 1, n * 16));
  const uint32_t cus = device_cus();
  double best = 0;
  for (uint32_t per_cu : {8u, 16u, 32u, 64u}) {
    double ms = 0;
    const int rc = time_launches(
        reps, [&]() { return launch_read(a.as<uint4>(), n, sink.as<uint32_t>(), cus * per_cu, nullptr); }, &ms);
    if (rc != GX_OK) return rc;
    best = std::max(best, static_cast<double>(n * 16) / (ms / 1e3) / 1e9);
  }
  *gbps = best;
  return GX_OK;
}

extern "C" int gx_bench_random(uint64_t table_bytes, uint64_t lookups, int reps, double *lookups_per_s) {
  if (!lookups_per_s || table_bytes < 64 || lookups < 1 || reps < 1) return GX_E_ARG;
  if (device_count() == 0) return GX_E_NODEVICE;
  const uint64_t nb = table_bytes / 64;
  DevBuf t, sink;
  GX_TRY(t.alloc(nb * 64));
  GX_TRY(sink.alloc(4));
  GX_TRY(hipMemset(t.p, 1, nb * 64));
  const uint32_t cus = device_cus();
  double best_ms = 0;
  for (uint32_t per_cu : {16u, 32u, 64u}) {
    double ms = 0;
    const int rc = time_launches(
        reps,
        [&]() {
          return launch_random_read(t.as<uint64_t>(), nb, lookups, 0x5eedull, sink.as<uint32_t>(), cus * per_cu,
                                    nullptr);
        },
        &ms);
    if (rc != GX_OK) return rc;
    if (best_ms == 0 || ms < best_ms) best_ms = ms;
  }
  *lookups_per_s = static_cast<double>(lookups) / (best_ms / 1e3);
  return GX_OK;
}
