// gx_serve.cpp -- continuous-batching model-checking server.
//
// A server owns fixed device pools: `slots` instance slots, a visited set of
// `table_bytes` cut into pages, frontier pages paired one to one with the
// table pages, a model region per slot and one successor buffer.  Jobs wait
// in a FIFO queue.  Every step
//   1. admits waiting jobs while a slot and a contiguous run of pages for the
//      job's estimated footprint are free (the partition is zeroed and the
//      initial state placed in it and in the job's frontier),
//   2. replays the captured rounds once (each round expands one BFS level of
//      every running job: the jobs' work overlaps on the device),
//   3. retires finished jobs; a job whose partition filled up is queued
//      again with four times the pages, so the result never depends on the
//      estimate.
// Visited sets and frontiers of running jobs stay resident between steps.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
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

}  // namespace

struct gx_server {
  gx_server_options o{};
  int W = 2, BW = 8;
  uint64_t spb = 4, page_buckets = 0, page_entries = 0, npages = 0, model_words = 0;
  DevBuf inst, ctrl, models, f0, f1, cand, cinst, def0, def1, table;
  hipStream_t st = nullptr;
  // rounds_per_step rounds, captured twice: with level stamps (replayed while
  // a job with the cycle proviso runs) and without (the leaner FIND-OR-PUT)
  hipGraph_t graph = nullptr, graph_nostamp = nullptr;
  hipGraphExec_t exec = nullptr, exec_nostamp = nullptr;
  BatchArgs a{};
  uint64_t next_id = 1;
  std::deque<Job> waiting;
  std::vector<Job> running;            // per slot
  std::vector<char> slot_used;
  std::vector<uint64_t> slot_page0, slot_npages;
  std::vector<double> slot_admit;
  std::vector<char> page_used;
  std::vector<BatchInst> hinst;
  std::vector<gx_job_result> done;
  uint64_t steps = 0, rounds = 0;
  uint32_t one = 1;  // host source of the prefix_only flag copy
  ~gx_server() {
    if (st) (void)hipStreamSynchronize(st);  // nothing may still run on the pools
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (exec_nostamp) (void)hipGraphExecDestroy(exec_nostamp);
    if (graph_nostamp) (void)hipGraphDestroy(graph_nostamp);
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" void gx_server_options_default(gx_server_options *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->slots = 256;
  o->key_words = 2;
  o->max_procs = 32;
  o->threads = 256;
  o->rounds_per_step = 16;  // a 10 s stream at 20 jobs/s: p50 0.74 ms (4: 0.98 ms, 1: 1.55 ms)
  o->table_bytes = 16ull << 30;
  o->page_bytes = 4ull << 20;
  o->model_words = 1u << 15;
  o->hash_seed = 0x5eed5eed5eedull;
  o->por_jobs = 1;
}

extern "C" int gx_server_create(const gx_server_options *opt, gx_server **out) {
  if (!opt || !out) return GX_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GX_E_NODEVICE;
  std::unique_ptr<gx_server> s(new (std::nothrow) gx_server());
  if (!s) return GX_E_NOMEM;
  gx_server_options o = *opt;
  if (!o.slots) o.slots = 256;
  if (o.slots > (1u << 16)) return GX_E_ARG;
  if (o.key_words < 1 || o.key_words > static_cast<uint32_t>(kMaxKeyWords)) return GX_E_ARG;
  if (o.max_procs < 1 || o.max_procs > static_cast<uint32_t>(kMaxProcs)) return GX_E_ARG;
  o.threads = std::min<uint32_t>(512, std::max<uint32_t>(64, (o.threads ? o.threads : 256) / 64 * 64));
  if (!o.rounds_per_step) o.rounds_per_step = 1;
  if (!o.model_words) o.model_words = 1u << 15;
  o.model_words = (o.model_words + 3u) & ~3u;  // 16-byte aligned model regions
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
  s->model_words = o.model_words;
  int dev = 0;
  SV_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SV_TRY(hipGetDeviceProperties(&prop, dev));
  if (!s->o.blocks) s->o.blocks = prop.multiProcessorCount * (2048 / s->o.threads);
  if (!s->o.insert_blocks) s->o.insert_blocks = prop.multiProcessorCount * 16;
  const uint64_t S = o.slots, W = s->W;
  const uint64_t front_total = s->npages * s->page_entries;
  // successor buffer: as many entries as the frontier pool, 2^24..2^30 by
  // default.  A level that does not fit is expanded over several rounds; the
  // largest levels of a 3e8-state job emit about 7e8 successors, so 2^30
  // keeps them to one round.
  const uint64_t cand_cap = std::min<uint64_t>(
      0xFFFFFFFFull,
      o.cand_capacity ? o.cand_capacity : std::min<uint64_t>(1ull << 30, std::max<uint64_t>(1u << 24, front_total)));
  const uint64_t max_chunks = front_total / 64 + S + 1;
  SV_TRY(s->inst.alloc(sizeof(BatchInst) * S));
  SV_TRY(s->ctrl.alloc(sizeof(BatchCtrl)));
  SV_TRY(s->models.alloc(S * s->model_words * 4));
  SV_TRY(s->f0.alloc(front_total * W * 8));
  SV_TRY(s->f1.alloc(front_total * W * 8));
  SV_TRY(s->cand.alloc(cand_cap * W * 8));
  SV_TRY(s->cinst.alloc((cand_cap / 64 + 1) * 4));
  SV_TRY(s->def0.alloc(max_chunks * 8));
  SV_TRY(s->def1.alloc(max_chunks * 8));
  SV_TRY(s->table.alloc(s->npages * s->page_buckets * bb));
  SV_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
  s->hinst.assign(S, BatchInst{});
  for (auto &I : s->hinst) I.stop = 1;
  SV_TRY(hipMemcpyAsync(s->inst.p, s->hinst.data(), sizeof(BatchInst) * S, hipMemcpyHostToDevice, s->st));
  BatchCtrl c0{};
  c0.n = static_cast<uint32_t>(S);
  SV_TRY(hipMemcpyAsync(s->ctrl.p, &c0, sizeof c0, hipMemcpyHostToDevice, s->st));
  BatchArgs &a = s->a;
  a.inst = s->inst.as<BatchInst>();
  a.ctrl = s->ctrl.as<BatchCtrl>();
  a.models = s->models.as<uint32_t>();
  a.front[0] = s->f0.as<uint64_t>();
  a.front[1] = s->f1.as<uint64_t>();
  a.cand = s->cand.as<uint64_t>();
  a.cand_blk = s->cinst.as<uint32_t>();
  a.cand_cap = cand_cap;
  a.deferred[0] = s->def0.as<uint64_t>();
  a.deferred[1] = s->def1.as<uint64_t>();
  a.table = s->table.as<uint64_t>();
  a.seed = o.hash_seed;
  a.bucket_words = static_cast<uint32_t>(s->BW);
  a.max_procs = o.max_procs;
  a.rows16 = rows16 ? 1u : 0u;
  a.por = o.por_jobs ? 1u : 0u;
  a.stamps = o.por_jobs ? 1u : 0u;
  a.por_threads = 256;
  a.por_blocks = static_cast<uint32_t>(prop.multiProcessorCount) * 4;
  a.chunk_states = o.threads;  // 64 states per wave
  a.expand_lds = static_cast<uint32_t>(batch_expand_lds(o.threads, o.max_procs, rows16));
  // capture rounds_per_step rounds once per slot format; replays reuse the
  // fixed pools
  SV_TRY(hipStreamSynchronize(s->st));
  for (int stamped = 0; stamped < (o.por_jobs ? 2 : 1); stamped++) {
    BatchArgs ca = a;
    ca.stamps = static_cast<uint32_t>(stamped);
    SV_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipSuccess;
    for (uint32_t r = 0; r < o.rounds_per_step && ce == hipSuccess; r++)
      ce = launch_round_w(ca, s->W, s->BW, s->o.blocks, s->o.threads, s->o.insert_blocks, s->st);
    hipGraph_t &g = stamped ? s->graph : s->graph_nostamp;
    hipGraphExec_t &x = stamped ? s->exec : s->exec_nostamp;
    hipError_t ee = hipStreamEndCapture(s->st, &g);
    SV_TRY(ce);
    SV_TRY(ee);
    SV_TRY(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
  }
  SV_TRY(hipStreamSynchronize(s->st));
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
  s->waiting.push_back(std::move(j));
  if (job_id) *job_id = s->waiting.back().id;
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

static int admit(gx_server *s, uint32_t slot, Job &&j, uint64_t page0) {
  const uint64_t bb = static_cast<uint64_t>(s->BW) * 8;
  BatchInst I{};
  I.hdr = j.hdr;
  I.model_off = static_cast<uint32_t>(slot * s->model_words);
  I.property = j.property;
  I.mode = j.mode;
  I.por_method = j.por_method;
  I.tb = page0 * s->page_buckets;
  I.nb = j.pages * s->page_buckets;
  I.fo = page0 * s->page_entries;
  I.fcap = j.pages * s->page_entries;
  I.front_count[0] = 1;  // the initial state, level 0
  I.inserts = 1;
  I.cur = 0;
  I.level = 0;
  hipStream_t st = s->st;
  SV_TRY(hipMemcpyAsync(s->models.as<uint32_t>() + I.model_off, j.words.data(), j.words.size() * 4,
                        hipMemcpyHostToDevice, st));
  SV_TRY(hipMemsetAsync(s->table.as<uint64_t>() + I.tb * s->BW, 0, I.nb * bb, st));
  s->hinst[slot] = I;
  SV_TRY(hipMemcpyAsync(s->inst.as<BatchInst>() + slot, &s->hinst[slot], sizeof I, hipMemcpyHostToDevice, st));
  // the initial state is written by a kernel behind the zeroing, in stream
  // order (a host copy into the table right after the memset was seen to be
  // missing when the first level probed for it)
  SV_TRY(launch_place_initial(s->a, s->W, s->BW, I.tb, I.nb, I.fo, st));
  s->slot_used[slot] = 1;
  s->slot_page0[slot] = page0;
  s->slot_npages[slot] = j.pages;
  s->slot_admit[slot] = now_s();
  s->running[slot] = std::move(j);
  SV_TRY(hipStreamSynchronize(st));  // host sources of the copies above are reused
  return GX_OK;
}

extern "C" int gx_server_step(gx_server *s, gx_server_status *status) {
  if (!s) return GX_E_ARG;
  BatchCtrl ch{};
  SV_TRY(hipMemcpyAsync(&ch, s->ctrl.p, sizeof ch, hipMemcpyDeviceToHost, s->st));
  SV_TRY(hipStreamSynchronize(s->st));
  // 1. admissions (only between levels: no deferred chunks pending)
  bool admitted = false;
  if (!ch.use_deferred && ch.deferred_count[0] == 0 && ch.deferred_count[1] == 0) {
    const uint32_t S = s->o.slots;
    uint32_t slot = 0;
    while (!s->waiting.empty()) {
      while (slot < S && s->slot_used[slot]) slot++;
      if (slot == S) break;
      uint64_t p0 = 0;
      if (!alloc_pages(s, s->waiting.front().pages, &p0)) break;  // FIFO: wait for memory
      Job j = std::move(s->waiting.front());
      s->waiting.pop_front();
      int rc = admit(s, slot, std::move(j), p0);
      if (rc != GX_OK) return rc;
      admitted = true;
    }
  }
  if (admitted) {
    SV_TRY(hipMemcpyAsync(&s->a.ctrl->prefix_only, &s->one, 4, hipMemcpyHostToDevice, s->st));
    SV_TRY(launch_batch_advance(s->a, s->st));
  }
  // 2. one replay of the captured rounds (with level stamps while a job with
  // the cycle proviso runs)
  bool any = false, stamps = false;
  for (uint32_t i = 0; i < s->o.slots; i++) {
    any = any || s->slot_used[i];
    stamps = stamps || (s->slot_used[i] && s->running[i].mode == kModePorProviso);
  }
  if (any) {
    SV_TRY(hipGraphLaunch(stamps ? s->exec : s->exec_nostamp, s->st));
    s->steps++;
  }
  // 3. retire finished jobs
  SV_TRY(hipMemcpyAsync(s->hinst.data(), s->inst.p, sizeof(BatchInst) * s->o.slots, hipMemcpyDeviceToHost, s->st));
  SV_TRY(hipMemcpyAsync(&ch, s->ctrl.p, sizeof ch, hipMemcpyDeviceToHost, s->st));
  SV_TRY(hipStreamSynchronize(s->st));
  const double t = now_s();
  uint32_t running = 0;
  for (uint32_t i = 0; i < s->o.slots; i++) {
    if (!s->slot_used[i]) continue;
    const BatchInst &I = s->hinst[i];
    if (!I.stop) {
      running++;
      continue;
    }
    Job &j = s->running[i];
    for (uint64_t q = 0; q < s->slot_npages[i]; q++) s->page_used[s->slot_page0[i] + q] = 0;
    s->slot_used[i] = 0;
    if ((I.table_full || I.overflow) && !I.cand_overflow && j.pages * 2 <= s->npages) {
      // restart with a larger partition: 4x (2x at the pool's limit); a
      // full partition usually means the estimate was off by a large factor
      j.pages = j.pages * 4 <= s->npages ? j.pages * 4 : j.pages * 2;
      j.restarts++;
      s->waiting.push_front(std::move(j));
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
  s->rounds = ch.rounds;
  if (status) {
    status->running = running;
    status->waiting = static_cast<uint32_t>(s->waiting.size());
    status->finished_pending = static_cast<uint32_t>(s->done.size());
    status->steps = s->steps;
    status->rounds = s->rounds;
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
