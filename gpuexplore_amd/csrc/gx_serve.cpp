// gx_serve.cpp -- continuous-batching model-checking server.
//
// A server owns fixed device pools: a visited set of `table_bytes` cut into
// pages, frontier pages paired with the table pages (half a page's slots
// each), and a model region per instance slot.  Jobs are served in two
// lanes by estimated footprint: small jobs and large ones (an estimate of at
// least `large_states` states).  Each lane has its own instance slots,
// control block, successor buffer, captured rounds and stream, so the two
// lanes' rounds run beside each other on the device: small jobs advance in
// rounds of their own instead of waiting, level after level, for the levels of
// a large job.  Every step, for each lane whose last replay has finished:
//   1. retire finished jobs; a job whose partition (or frontier) filled up is
//      queued again with twice the pages, so the result never depends on the
//      estimate;
//   2. admit waiting jobs of the lane (FIFO) while a slot and a contiguous run
//      of pages for 2x the job's estimated states are free (the partition is
//      zeroed and the initial state placed in it and in the job's frontier);
//   3. replay the lane's captured rounds once (each round expands one BFS
//      level of every running job of the lane: their work overlaps on the
//      device), without waiting for it.
// Visited sets and frontiers of running jobs stay resident between steps.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/gpuexplore.h"
#include "gx_batch.h"
#include "gx_devbuf.h"
#include "gx_device.h"
#include "gx_kernels.h"
#include "gx_model.h"

using namespace gx;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int map_err(hipError_t e) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return GX_E_NOMEM;
  return GX_E_HIP;
}

#define SV_TRY(expr)                          \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return map_err(e_); \
  } while (0)

struct Job {
  uint64_t id;
  std::vector<uint32_t> words;  // compiled model
  ModelHdr hdr;
  int32_t property;
  uint32_t mode, por_method;    // ExpandMode and GX_POR_* (gx_job_options)
  uint64_t pages;               // pages requested
  double t_submit;
  int32_t restarts;
};

// One lane: instance slots [slot0, slot0 + nslots) of the server, and the
// device state and captured rounds that expand them.
struct Lane {
  uint32_t slot0 = 0, nslots = 0;
  DevBuf ctrl, cand, cblk, def0, def1;
  hipStream_t st = nullptr;
  // rounds_per_step rounds, captured twice: with level stamps (replayed while
  // a job with the cycle proviso runs) and without (the leaner FIND-OR-PUT)
  hipGraph_t graph = nullptr, graph_nostamp = nullptr;
  hipGraphExec_t exec = nullptr, exec_nostamp = nullptr;
  hipEvent_t done = nullptr;  // end of the last replay and of its read-back
  bool in_flight = false;
  BatchArgs a{};
  BatchInst *h_inst = nullptr;  // pinned: the lane's instances after the last replay
  BatchCtrl *h_ctrl = nullptr;  // pinned: its control block
  std::deque<Job> waiting;
  uint64_t rounds = 0;
  ~Lane() {
    if (st) (void)hipStreamSynchronize(st);  // nothing may still run on the pools
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (exec_nostamp) (void)hipGraphExecDestroy(exec_nostamp);
    if (graph_nostamp) (void)hipGraphDestroy(graph_nostamp);
    if (done) (void)hipEventDestroy(done);
    if (h_inst) (void)hipHostFree(h_inst);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (st) (void)hipStreamDestroy(st);
  }
};

constexpr int kLanes = 2;  // 0: small jobs, 1: large jobs

}  // namespace

struct gx_server {
  gx_server_options o{};
  int W = 2, BW = 8;
  uint64_t spb = 4, page_buckets = 0, page_entries = 0, page_front = 0, npages = 0, model_words = 0;
  // pools first: members are destroyed in reverse order, so the lanes (whose
  // destructors wait for their streams) go before the memory they use
  DevBuf inst, models, f0, f1, table;
  Lane lane[kLanes];
  uint64_t next_id = 1;
  std::vector<Job> running;  // per slot
  std::vector<char> slot_used;
  std::vector<uint64_t> slot_page0, slot_npages;
  std::vector<double> slot_admit;
  std::vector<char> page_used;
  std::vector<gx_job_result> done;
  uint64_t steps = 0;
  uint32_t one = 1;  // host source of the prefix_only flag copy
};

extern "C" void gx_server_options_default(gx_server_options *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->slots = 256;
  o->key_words = 2;
  o->max_procs = 32;
  o->threads = 256;
  o->rounds_per_step = 16;
  o->table_bytes = 16ull << 30;
  o->page_bytes = 4ull << 20;
  o->model_words = 1u << 15;
  o->hash_seed = 0x5eed5eed5eedull;
  o->por_jobs = 1;
  o->large_states = 1ull << 24;
  o->large_slots = 8;
}

namespace {

int create_lane(gx_server *s, int k, uint32_t slot0, uint32_t nslots, uint64_t cand_cap, uint32_t cus,
                bool rows16) {
  Lane &L = s->lane[k];
  L.slot0 = slot0;
  L.nslots = nslots;
  const uint64_t W = s->W;
  const uint64_t front_total = s->npages * s->page_front;
  const uint64_t max_chunks = front_total / 64 + nslots + 1;
  SV_TRY(L.ctrl.alloc(sizeof(BatchCtrl)));
  SV_TRY(L.cand.alloc(cand_cap * W * 8));
  SV_TRY(L.cblk.alloc((cand_cap / 64 + 1) * 4));
  SV_TRY(L.def0.alloc(max_chunks * 8));
  SV_TRY(L.def1.alloc(max_chunks * 8));
  if (k == 1 && s->o.large_cus < cus) {
    // the large-job lane's kernels run on the first large_cus compute units
    // only, so the small-job lane's rounds always find free ones
    std::vector<uint32_t> mask((cus + 31) / 32, 0u);
    for (uint32_t c = 0; c < s->o.large_cus; c++) mask[c / 32] |= 1u << (c % 32);
    SV_TRY(hipExtStreamCreateWithCUMask(&L.st, static_cast<uint32_t>(mask.size()), mask.data()));
  } else {
    SV_TRY(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
  }
  SV_TRY(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
  SV_TRY(hipHostMalloc(reinterpret_cast<void **>(&L.h_inst), sizeof(BatchInst) * nslots, hipHostMallocDefault));
  SV_TRY(hipHostMalloc(reinterpret_cast<void **>(&L.h_ctrl), sizeof(BatchCtrl), hipHostMallocDefault));
  memset(L.h_ctrl, 0, sizeof(BatchCtrl));
  L.h_ctrl->n = nslots;
  for (uint32_t i = 0; i < nslots; i++) {
    memset(&L.h_inst[i], 0, sizeof(BatchInst));
    L.h_inst[i].stop = 1;
  }
  SV_TRY(hipMemcpyAsync(s->inst.as<BatchInst>() + slot0, L.h_inst, sizeof(BatchInst) * nslots,
                        hipMemcpyHostToDevice, L.st));
  SV_TRY(hipMemcpyAsync(L.ctrl.p, L.h_ctrl, sizeof(BatchCtrl), hipMemcpyHostToDevice, L.st));
  BatchArgs &a = L.a;
  a.inst = s->inst.as<BatchInst>() + slot0;
  a.ctrl = L.ctrl.as<BatchCtrl>();
  a.models = s->models.as<uint32_t>();
  a.front[0] = s->f0.as<uint64_t>();
  a.front[1] = s->f1.as<uint64_t>();
  a.cand = L.cand.as<uint64_t>();
  a.cand_blk = L.cblk.as<uint32_t>();
  a.cand_cap = cand_cap;
  a.deferred[0] = L.def0.as<uint64_t>();
  a.deferred[1] = L.def1.as<uint64_t>();
  a.table = s->table.as<uint64_t>();
  a.seed = s->o.hash_seed;
  a.bucket_words = static_cast<uint32_t>(s->BW);
  a.max_procs = s->o.max_procs;
  a.rows16 = rows16 ? 1u : 0u;
  a.por = s->o.por_jobs ? 1u : 0u;
  a.stamps = s->o.por_jobs ? 1u : 0u;
  a.por_threads = 256;
  a.por_blocks = cus * 4;
  a.chunk_states = s->o.threads;  // 64 states per wave
  a.expand_lds = static_cast<uint32_t>(batch_expand_lds(s->o.threads, s->o.max_procs, rows16));
  // capture rounds_per_step rounds once per slot format; replays reuse the
  // fixed pools
  SV_TRY(hipStreamSynchronize(L.st));
  for (int stamped = 0; stamped < (s->o.por_jobs ? 2 : 1); stamped++) {
    BatchArgs ca = a;
    ca.stamps = static_cast<uint32_t>(stamped);
    SV_TRY(hipStreamBeginCapture(L.st, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipSuccess;
    for (uint32_t r = 0; r < s->o.rounds_per_step && ce == hipSuccess; r++)
      ce = launch_round_w(ca, s->W, s->BW, s->o.blocks, s->o.threads, s->o.insert_blocks, L.st);
    hipGraph_t &g = stamped ? L.graph : L.graph_nostamp;
    hipGraphExec_t &x = stamped ? L.exec : L.exec_nostamp;
    hipError_t ee = hipStreamEndCapture(L.st, &g);
    SV_TRY(ce);
    SV_TRY(ee);
    SV_TRY(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
  }
  SV_TRY(hipStreamSynchronize(L.st));
  return GX_OK;
}

}  // namespace

extern "C" int gx_server_create(const gx_server_options *opt, gx_server **out) {
  if (!opt || !out) return GX_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GX_E_NODEVICE;
  std::unique_ptr<gx_server> s(new (std::nothrow) gx_server());
  if (!s) return GX_E_NOMEM;
  gx_server_options o = *opt;
  if (!o.slots) o.slots = 256;
  if (o.slots < 2 || o.slots > (1u << 16)) return GX_E_ARG;
  if (o.key_words < 1 || o.key_words > static_cast<uint32_t>(kMaxKeyWords)) return GX_E_ARG;
  if (o.max_procs < 1 || o.max_procs > static_cast<uint32_t>(kMaxProcs)) return GX_E_ARG;
  o.threads = std::min<uint32_t>(512, std::max<uint32_t>(64, (o.threads ? o.threads : 256) / 64 * 64));
  if (!o.rounds_per_step) o.rounds_per_step = 1;
  if (!o.model_words) o.model_words = 1u << 15;
  o.model_words = (o.model_words + 3u) & ~3u;  // 16-byte aligned model regions
  if (!o.large_states) o.large_states = ~0ull;  // every job in the small-job lane
  o.large_slots = std::min<uint32_t>(std::max<uint32_t>(o.large_slots, 1), std::max<uint32_t>(1, o.slots / 4));
  // a model has fewer rows than words: 16-bit row indices when models are small
  const bool rows16 = o.model_words < 0xFFFFu;
  if (batch_rows_lds(o.threads, o.max_procs, rows16) > 64 * 1024) return GX_E_LIMIT;
  s->o = o;
  s->W = static_cast<int>(o.key_words);
  const uint32_t bb = o.key_words <= 2 ? 64u : 128u;
  s->BW = static_cast<int>(bb / 8);
  s->spb = static_cast<uint64_t>(s->BW / s->W);
  s->page_buckets = std::max<uint64_t>(1, o.page_bytes / bb);
  s->npages = std::max<uint64_t>(1, o.table_bytes / (s->page_buckets * bb));
  s->page_entries = s->page_buckets * s->spb;
  // frontier entries per page: half the page's slots, so a partition sized
  // for twice its estimate holds a level of every estimated state (a 24-process
  // POR job keeps up to two thirds of its states in its last level); a job
  // whose level overflows restarts with more pages, like a full partition
  s->page_front = std::max<uint64_t>(64, s->page_entries / 2);
  s->model_words = o.model_words;
  int dev = 0;
  SV_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SV_TRY(hipGetDeviceProperties(&prop, dev));
  const uint32_t cus = static_cast<uint32_t>(prop.multiProcessorCount);
  if (!s->o.blocks) s->o.blocks = cus * (2048 / s->o.threads);
  if (!s->o.insert_blocks) s->o.insert_blocks = cus * 16;
  // the large-job lane leaves an eighth of the CUs to the small-job lane by
  // default: a 20 jobs/s stream of 24-process POR jobs beside 20-process
  // 3e8-state jobs had a POR-job p50 of 236 ms sharing every CU, 29 ms with
  // 224 of 256 CUs for the large jobs (192: 24 ms, but the large jobs slow)
  if (!s->o.large_cus) s->o.large_cus = cus - cus / 8;
  const uint64_t S = o.slots, W = s->W;
  const uint64_t front_total = s->npages * s->page_front;
  SV_TRY(s->inst.alloc(sizeof(BatchInst) * S));
  SV_TRY(s->models.alloc(S * s->model_words * 4));
  SV_TRY(s->f0.alloc(front_total * W * 8));
  SV_TRY(s->f1.alloc(front_total * W * 8));
  SV_TRY(s->table.alloc(s->npages * s->page_buckets * bb));
  // successor buffers: the large-job lane's holds a whole level of a
  // 3e8-state job (about 7e8 successors) in one round, up to 2^30 entries;
  // the small-job lane's 2^24 (a level that does not fit is expanded over
  // several rounds)
  const uint64_t large_cap = std::min<uint64_t>(
      0xFFFFFFFFull,
      o.cand_capacity ? o.cand_capacity
                      : std::min<uint64_t>(1ull << 30, std::max<uint64_t>(1u << 24, front_total)));
  const uint64_t small_cap = std::min<uint64_t>(large_cap, 1ull << 24);
  const uint32_t large_slots = s->o.large_slots;
  int rc = create_lane(s.get(), 0, 0, o.slots - large_slots, small_cap, cus, rows16);
  if (rc == GX_OK) rc = create_lane(s.get(), 1, o.slots - large_slots, large_slots, large_cap, cus, rows16);
  if (rc != GX_OK) return rc;
  s->running.resize(S);
  s->slot_used.assign(S, 0);
  s->slot_page0.assign(S, 0);
  s->slot_npages.assign(S, 0);
  s->slot_admit.assign(S, 0.0);
  s->page_used.assign(s->npages, 0);
  *out = s.release();
  return GX_OK;
}

extern "C" void gx_server_free(gx_server *s) { delete s; }

extern "C" int gx_server_submit(gx_server *s, const gx_model *m, int32_t property, uint64_t est_states,
                                uint64_t *job_id) {
  if (!m || m->por) return GX_E_ARG;  // plain method only (gx_server_submit_job has POR)
  const gx_job_options j{property, 0, 0, GX_POR_CAMPLE};
  return gx_server_submit_job(s, m, &j, est_states, job_id);
}

extern "C" int gx_server_submit_job(gx_server *s, const gx_model *m, const gx_job_options *job,
                                    uint64_t est_states, uint64_t *job_id) {
  if (!s || !m || !job) return GX_E_ARG;
  if (job->property < GX_PROP_NONE || job->property > GX_PROP_LIVENESS || job->por_method > GX_POR_STUBBORN)
    return GX_E_ARG;
  if (job->por && !s->o.por_jobs) return GX_E_ARG;
  const ModelHdr &H = m->comp.hdr;
  if ((!job->por && !H.lab64) || H.key_words > static_cast<uint32_t>(s->W) || H.nprocs > s->o.max_procs ||
      m->comp.words.size() > s->model_words)
    return GX_E_LIMIT;
  // the level stamp of the cycle proviso follows the state bits in the slot
  if (job->por && job->cycle_proviso && H.state_bits + 32 > 63u * static_cast<uint32_t>(s->W)) return GX_E_LIMIT;
  Job j;
  j.id = s->next_id++;
  j.words = m->comp.words;
  j.hdr = H;
  j.property = job->property;
  j.mode = batch_mode(job->por, job->cycle_proviso);
  j.por_method = job->por_method;
  // footprint: slots for 2x the estimate, at least one page
  const uint64_t slots = std::max<uint64_t>(2 * std::max<uint64_t>(est_states, 1), 1);
  j.pages = std::max<uint64_t>(1, (slots + s->page_entries - 1) / s->page_entries);
  j.t_submit = now_s();
  j.restarts = 0;
  if (j.pages > s->npages) return GX_E_CAPACITY;
  Lane &L = s->lane[est_states >= s->o.large_states ? 1 : 0];
  L.waiting.push_back(std::move(j));
  if (job_id) *job_id = L.waiting.back().id;
  return GX_OK;
}

// A run of `n` free pages: small partitions by first fit from the bottom of
// the pool, large ones (above 1/16 of it) by first fit from the top, so the
// many small jobs do not cut the pool into runs too short for a large one.
static bool alloc_pages(gx_server *s, uint64_t n, uint64_t *first) {
  uint64_t run = 0;
  if (n * 16 > s->npages) {
    for (uint64_t q = s->npages; q-- > 0;) {
      run = s->page_used[q] ? 0 : run + 1;
      if (run == n) {
        *first = q;
        for (uint64_t r = q; r < q + n; r++) s->page_used[r] = 1;
        return true;
      }
    }
    return false;
  }
  for (uint64_t p = 0; p < s->npages; p++) {
    run = s->page_used[p] ? 0 : run + 1;
    if (run == n) {
      *first = p + 1 - n;
      for (uint64_t q = *first; q <= p; q++) s->page_used[q] = 1;
      return true;
    }
  }
  return false;
}

// Admits job j into global slot `slot` of lane L (on L's stream).
static int admit(gx_server *s, Lane &L, uint32_t slot, Job &&j, uint64_t page0) {
  const uint64_t bb = static_cast<uint64_t>(s->BW) * 8;
  BatchInst &I = L.h_inst[slot - L.slot0];
  I = BatchInst{};
  I.hdr = j.hdr;
  I.model_off = static_cast<uint32_t>(slot * s->model_words);
  I.property = j.property;
  I.mode = j.mode;
  I.por_method = j.por_method;
  I.tb = page0 * s->page_buckets;
  I.nb = j.pages * s->page_buckets;
  I.fo = page0 * s->page_front;
  I.fcap = j.pages * s->page_front;
  I.front_count[0] = 1;  // the initial state, level 0
  I.inserts = 1;
  I.cur = 0;
  I.level = 0;
  hipStream_t st = L.st;
  SV_TRY(hipMemcpyAsync(s->models.as<uint32_t>() + I.model_off, j.words.data(), j.words.size() * 4,
                        hipMemcpyHostToDevice, st));
  SV_TRY(hipMemsetAsync(s->table.as<uint64_t>() + I.tb * s->BW, 0, I.nb * bb, st));
  SV_TRY(hipMemcpyAsync(L.a.inst + (slot - L.slot0), &I, sizeof I, hipMemcpyHostToDevice, st));
  // the initial state is written by a kernel behind the zeroing, in stream
  // order (a host copy into the table right after the memset was seen to be
  // missing when the first level probed for it)
  SV_TRY(launch_place_initial(L.a, s->W, s->BW, I.tb, I.nb, I.fo, st));
  s->slot_used[slot] = 1;
  s->slot_page0[slot] = page0;
  s->slot_npages[slot] = j.pages;
  s->slot_admit[slot] = now_s();
  s->running[slot] = std::move(j);
  SV_TRY(hipStreamSynchronize(st));  // host sources of the copies above are reused
  return GX_OK;
}

// Retires the finished jobs of lane L (its state read back after the last
// replay).
static void retire(gx_server *s, Lane &L) {
  const double t = now_s();
  for (uint32_t k = 0; k < L.nslots; k++) {
    const uint32_t i = L.slot0 + k;
    if (!s->slot_used[i]) continue;
    const BatchInst &I = L.h_inst[k];
    if (!I.stop) continue;
    Job &j = s->running[i];
    for (uint64_t q = 0; q < s->slot_npages[i]; q++) s->page_used[s->slot_page0[i] + q] = 0;
    s->slot_used[i] = 0;
    if ((I.table_full || I.overflow) && !I.cand_overflow && j.pages * 2 <= s->npages) {
      // restart with twice the partition (a larger step would let one
      // restarted job hold most of the pool)
      j.pages = j.pages * 2;
      j.restarts++;
      L.waiting.push_front(std::move(j));
      continue;
    }
    gx_job_result r{};
    r.job_id = j.id;
    r.result.states = I.inserts;
    r.result.transitions = I.transitions;
    r.result.levels = I.level;
    r.result.inserts = I.inserts;
    r.result.violation = I.violation ? 1 : 0;
    r.result.table_full = (I.table_full || I.overflow) ? 1 : 0;
    r.result.table_bytes = s->slot_npages[i] * s->page_buckets * s->BW * 8;
    r.result.t_explore_s = t - s->slot_admit[i];
    r.t_submit = j.t_submit;
    r.t_admit = s->slot_admit[i];
    r.t_done = t;
    r.restarts = j.restarts;
    r.status = I.bad_state ? GX_E_STATE
               : I.spin_timeout ? GX_E_HIP
               : I.cand_overflow ? GX_E_CAPACITY  // more pages cannot fix it: no restart
                                 : GX_OK;
    s->done.push_back(r);
  }
}

// Works lane L if its last replay has finished (or it has none): retire,
// admit, replay.  *progress is set when anything was retired or launched.
static int work_lane(gx_server *s, Lane &L, bool *progress) {
  if (L.in_flight) {
    const hipError_t q = hipEventQuery(L.done);
    if (q == hipErrorNotReady) return GX_OK;
    SV_TRY(q);
    L.in_flight = false;
    L.rounds = L.h_ctrl->rounds;
    retire(s, L);
    *progress = true;
  }
  // admissions, only between levels (no deferred chunks pending)
  const BatchCtrl &ch = *L.h_ctrl;
  bool admitted = false;
  if (!ch.use_deferred && ch.deferred_count[0] == 0 && ch.deferred_count[1] == 0) {
    uint32_t slot = L.slot0;
    const uint32_t end = L.slot0 + L.nslots;
    while (!L.waiting.empty()) {
      while (slot < end && s->slot_used[slot]) slot++;
      if (slot == end) break;
      uint64_t p0 = 0;
      if (!alloc_pages(s, L.waiting.front().pages, &p0)) break;  // FIFO: wait for memory
      Job j = std::move(L.waiting.front());
      L.waiting.pop_front();
      const int rc = admit(s, L, slot, std::move(j), p0);
      if (rc != GX_OK) return rc;
      admitted = true;
    }
  }
  if (admitted) {
    SV_TRY(hipMemcpyAsync(&L.a.ctrl->prefix_only, &s->one, 4, hipMemcpyHostToDevice, L.st));
    SV_TRY(launch_batch_advance(L.a, L.st));
  }
  // one replay of the lane's rounds (with level stamps while a job with the
  // cycle proviso runs), then its state back into pinned memory
  bool any = false, stamps = false;
  for (uint32_t i = L.slot0; i < L.slot0 + L.nslots; i++) {
    any = any || s->slot_used[i];
    stamps = stamps || (s->slot_used[i] && s->running[i].mode == kModePorProviso);
  }
  if (any) {
    SV_TRY(hipGraphLaunch(stamps ? L.exec : L.exec_nostamp, L.st));
    SV_TRY(hipMemcpyAsync(L.h_inst, L.a.inst, sizeof(BatchInst) * L.nslots, hipMemcpyDeviceToHost, L.st));
    SV_TRY(hipMemcpyAsync(L.h_ctrl, L.a.ctrl, sizeof(BatchCtrl), hipMemcpyDeviceToHost, L.st));
    SV_TRY(hipEventRecord(L.done, L.st));
    L.in_flight = true;
    s->steps++;
    *progress = true;
  }
  return GX_OK;
}

extern "C" int gx_server_step(gx_server *s, gx_server_status *status) {
  if (!s) return GX_E_ARG;
  // Work every lane that is ready; wait (polling, up to about a millisecond)
  // only while lanes are in flight and none has made progress yet.
  const double t0 = now_s();
  for (;;) {
    bool progress = false;
    for (auto &L : s->lane) {
      const int rc = work_lane(s, L, &progress);
      if (rc != GX_OK) return rc;
    }
    bool in_flight = false;
    for (auto &L : s->lane) in_flight = in_flight || L.in_flight;
    if (progress || !in_flight || now_s() - t0 > 1e-3) break;
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  if (status) {
    uint32_t running = 0, waiting = 0;
    uint64_t rounds = 0;
    for (auto &L : s->lane) {
      waiting += static_cast<uint32_t>(L.waiting.size());
      rounds += L.rounds;
    }
    for (char c : s->slot_used) running += c ? 1 : 0;
    status->running = running;
    status->waiting = waiting;
    status->finished_pending = static_cast<uint32_t>(s->done.size());
    status->steps = s->steps;
    status->rounds = rounds;
    uint64_t used = 0;
    for (char c : s->page_used) used += c ? 1 : 0;
    status->pages_used = used;
    status->pages_total = s->npages;
  }
  return GX_OK;
}

extern "C" int gx_server_poll(gx_server *s, gx_job_result *out, int max, int *n) {
  if (!s || !n || (max > 0 && !out)) return GX_E_ARG;
  const int k = std::min<int>(max, static_cast<int>(s->done.size()));
  for (int i = 0; i < k; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return GX_OK;
}
