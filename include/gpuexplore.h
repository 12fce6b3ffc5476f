/*
 * gpuexplore.h -- C ABI of the MI355X GPUexplore library (libgpuexplore.so).
 *
 * The reference (GPUexplore/src/GPUexplore.cu, cited as cu:N) has no library
 * interface: its boundary is the `GPUexplore <model> [options]` process
 * (argv, <model>.gpf, stdout/stderr, exit code) and, inside it, the loader,
 * device setup, BFS driver and count steps of main().  Each entry point
 * below replaces one of those steps, so the CLI (gpuexplore_amd/csrc/
 * gx_cli.cpp) and any other host can be rebuilt on top of it.
 *
 * Conventions: every function returns 0 or a negative GX_E* status and never
 * prints or exits; the caller owns option and result structs; a model handle
 * owns its host arrays; explorations allocate and free their device buffers
 * themselves.  Distinct handles may be used from distinct host threads.
 */
#ifndef GPUEXPLORE_H
#define GPUEXPLORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GX_OK 0
#define GX_E_NOFILE (-1)    /* cu:1984-1986 "input network does not exist" */
#define GX_E_FORMAT (-2)    /* truncated / malformed gpf */
#define GX_E_RANGE (-3)     /* field widths, offsets or labels out of range */
#define GX_E_NOMEM (-4)     /* host or device allocation failed */
#define GX_E_HIP (-5)       /* HIP runtime error */
#define GX_E_NODEVICE (-6)  /* cu:2000-2003 "No CUDA compatible GPU detected" */
#define GX_E_ARG (-7)       /* invalid option or argument */
#define GX_E_LIMIT (-8)     /* model exceeds a supported limit */
#define GX_E_STATE (-9)     /* a reached local state has no offsets in the gpf */
#define GX_E_CAPACITY (-10) /* frontier or candidate buffer exhausted */
#define GX_E_SCRATCH (-11)  /* input larger than the no-allocation scratch holds */

/* properties (cu:27) */
#define GX_PROP_NONE 0
#define GX_PROP_DEADLOCK 1
#define GX_PROP_SAFETY 2
#define GX_PROP_LIVENESS 3

/* Partial-order reduction methods for gx_options.por_method.  The reference
 * keeps one method per branch (README.md lines 24-26: ample-por, cample-por,
 * stubborn-por); this snapshot's source holds cample (cu:1340-1675).  AMPLE
 * and STUBBORN have no source here: they select among the same successors
 * (local moves of a process set C and the enabled rules inside C) by the
 * published approaches:
 *   CAMPLE    process clusters (cu:1616-1672);
 *   AMPLE     the local moves of one process whose current labels have no
 *             rule containing it (the fewest local successors, lowest index);
 *   STUBBORN  closure of {p} under d(p) = p + each rule of p's labels that
 *             every member has + the lowest member lacking the label of each
 *             other rule (a necessary enabling set); the smallest closure that
 *             meets a process with an enabled successor.
 * With the cycle proviso, AMPLE and STUBBORN require a new successor in the
 * selected set and otherwise expand fully. */
#define GX_POR_CAMPLE 0
#define GX_POR_AMPLE 1
#define GX_POR_STUBBORN 2

typedef struct gx_model gx_model;

/* Parse <path_noext>.gpf (cu:1814-1815, 1885-1983).  `por` adds the extra
 * state-vector bit the reference loader adds (cu:1912). */
int gx_model_load_gpf(const char *path_noext, int por, gx_model **out);

/* Same, from gpf text in memory. */
int gx_model_load_gpf_text(const char *text, size_t len, int por, gx_model **out);

void gx_model_free(gx_model *m);

typedef struct gx_model_info {
  uint32_t nr_procs;          /* cu:1899-1902 */
  uint32_t bits_act;          /* cu:1903-1906 */
  uint32_t proc_nrstates;     /* cu:1907-1910 */
  uint32_t bits_statevector;  /* cu:1911-1914, +1 under POR */
  uint32_t sv_nints;          /* cu:1923 */
  uint32_t liveness;          /* gpf line 1, cu:1888-1898 */
  uint32_t max_buf_ints;      /* cu:1935-1938 */
  uint32_t nr_trans;          /* cu:1951-1954 */
  uint32_t state_bits;        /* firstbit[N]: bits used by process fields */
  uint32_t key_words;         /* 64-bit words of the device state key */
} gx_model_info;

int gx_model_info_get(const gx_model *m, gx_model_info *out);

/* Exploration options: the reference's -b -t -k -q -d/-p --por
 * --cycle-proviso (cu:1785-1883) as orthogonal fields.  The reference's
 * flag coupling (-d forces the proviso off, -p forces it on) is applied by
 * the CLI, not here. */
typedef struct gx_options {
  uint32_t blocks;           /* -b: workgroups per launch (0 = auto) */
  uint32_t threads;          /* -t: threads per workgroup, multiple of 64 (0 = auto) */
  uint32_t kernel_iters;     /* -k: BFS levels enqueued per host sync (0 = 1) */
  uint32_t bucket_bytes;     /* hash bucket width: 64, 128 or 256 (0 = 64 for slots of <= 2 words, else 128) */
  uint64_t q_ints;           /* -q: hash table size in 32-bit ints (0 = auto) */
  int32_t property;          /* GX_PROP_* */
  int32_t por;               /* --por */
  int32_t cycle_proviso;     /* --cycle-proviso */
  int32_t verbosity;         /* -v: 1 prints each synchronised level, 2 the state counts, 3 every
                                stored state (cu:2128-2140), on stdout */
  uint64_t hash_seed;        /* seeds the bucket hash; counts do not depend on it */
  uint64_t cand_capacity;    /* successor buffer entries (0 = auto) */
  uint64_t frontier_capacity;/* frontier entries per level buffer (0 = auto) */
  uint64_t max_states;       /* stop after the level that reaches this many states (0 = none) */
  uint32_t expand_kernel;    /* successor kernel: 0 auto, 1 process groups, 2 one state per lane */
  uint32_t insert_blocks;    /* FIND-OR-PUT workgroups of 256 threads per launch (0 = auto: up to 64 per CU);
                                -b/-t set the successor kernel's geometry */
  uint32_t por_method;       /* GX_POR_* when por is set (0 = cample, the reference's method) */
  uint32_t dedup;            /* 1: candidates equal to one of the 127 before them on the same
                                wave skip the visited set (the reference's LDS cache, cu:210-265);
                                0 (default): every candidate is looked up.  Counts do not depend
                                on it.  On the 20-process ring it drops 15 % of the lookups but no
                                FIND-OR-PUT time: the lookups it drops hit the caches anyway. */
  uint32_t reserved2;
} gx_options;

void gx_options_default(gx_options *o);

typedef struct gx_result {
  uint64_t states;        /* distinct states in the visited set (cu:2168-2172) */
  uint64_t transitions;   /* successors generated (under POR: ample ones) */
  uint64_t levels;        /* BFS levels expanded */
  uint64_t lookups;       /* candidates offered to the global table (FIND-OR-PUT) */
  uint64_t inserts;       /* candidates that were new */
  uint64_t hash_launches; /* FIND-OR-PUT kernel launches */
  uint64_t table_bytes;   /* bytes of the visited set */
  int32_t violation;      /* deadlock / safety violation found (cu:2150-2162) */
  int32_t table_full;     /* cu:2164-2166 "problem with hash table" */
  int32_t truncated;      /* stopped by max_states */
  int32_t reserved;
  double t_explore_s;     /* wall time of the exploration (excludes parse/alloc/init) */
  double t_hash_s;        /* summed FIND-OR-PUT kernel time (HIP events) */
  double t_expand_s;      /* summed successor-generation kernel time (HIP events) */
  double t_init_s;        /* device allocation + table initialisation */
} gx_result;

/* Explore the state space on the current HIP device: replaces the device
 * setup, BFS driver loop and count of main() (cu:1989-2142, 2168-2172). */
int gx_explore(const gx_model *m, const gx_options *o, gx_result *r);

/* Reusable exploration: create allocates the visited set (-q) and the
 * frontier and successor buffers once; every run resets the visited set
 * and explores from the initial state (cu:2091-2092 init, 2099-2142
 * driver).  Runs of one explorer must not overlap. */
typedef struct gx_explorer gx_explorer;
int gx_explorer_create(const gx_model *m, const gx_options *o, gx_explorer **out);
int gx_explorer_run(gx_explorer *e, gx_result *r);
/* Writes the visited set of the last run to `path` (stdout when NULL) in the
 * reference's --dump / -v 3 format: one line per state, "new: " for states
 * not yet expanded, the local states "p0,p1,...", then the packed 32-bit
 * vector words; finally "nr. of states in hash table: S (U unexplored
 * states)".  Replaces the --dump block cu:2174-2185 (format cu:544-568,
 * 594-612).  GX_E_NOFILE if the file cannot be opened, GX_E_ARG before a
 * run. */
int gx_explorer_dump(gx_explorer *e, const char *path);
void gx_explorer_free(gx_explorer *e);

const char *gx_strerror(int status);

/* ------------------------------------------------------------------ */
/* Per-job options of a batch instance or a served job: the reference's -d/-p
 * --por --cycle-proviso (cu:1785-1883) for one model, as in gx_options. */
typedef struct gx_job_options {
  int32_t property;        /* GX_PROP_* */
  int32_t por;             /* partial-order reduction */
  int32_t cycle_proviso;   /* with por: the cycle proviso */
  uint32_t por_method;     /* GX_POR_* with por */
} gx_job_options;

/* ------------------------------------------------------------------ */
/* Batched exploration: n independent models (plain method, each with its own
 * property) explored together.  Every instance owns a partition of one
 * visited set and its own frontiers; one round expands one BFS level of
 * every unfinished instance, and `levels_per_sync` rounds are captured in a
 * hipGraph and replayed until all instances stop.  Per instance the result
 * equals gx_explore's.  No reference counterpart: the reference explores one
 * model per process (cu:1738-2188). */
typedef struct gx_batch_options {
  uint32_t threads;          /* threads per workgroup (0 = 256) */
  uint32_t blocks;           /* workgroups per launch (0 = auto) */
  uint32_t levels_per_sync;  /* rounds per graph replay (0 = 1) */
  uint32_t bucket_bytes;     /* 64, 128 or 256 (0 = auto) */
  uint64_t q_ints;           /* total visited-set size in 32-bit ints, split over instances */
  uint64_t hash_seed;
  uint64_t cand_capacity;    /* successor buffer entries (0 = auto) */
  uint32_t use_graph;        /* 1: replay captured rounds; 0: launch them one by one */
  uint32_t insert_blocks;    /* FIND-OR-PUT workgroups of 256 threads per round (0 = 8 per CU) */
  uint32_t time_kernels;     /* 1: launch rounds one by one (no graph) with HIP events around the
                                successor and FIND-OR-PUT launches (gx_batch_stats.t_*_s) */
} gx_batch_options;

typedef struct gx_batch_stats {
  double t_explore_s;        /* wall time from the first round to the last */
  double t_init_s;           /* allocation, table init, initial states */
  uint64_t rounds;           /* rounds run (levels plus deferred rounds) */
  uint64_t graph_launches;   /* host synchronisations */
  uint64_t lookups;          /* FIND-OR-PUT operations, all instances */
  uint64_t states;           /* sum over instances */
  uint64_t transitions;      /* sum over instances */
  uint64_t cand_entries;     /* successor-buffer entries scanned by FIND-OR-PUT (successors and the
                                released tails of reservations) */
  double t_expand_s;         /* with time_kernels: successor kernels' time, summed over rounds */
  double t_insert_s;         /* with time_kernels: FIND-OR-PUT kernel time, summed over rounds */
  uint64_t insert_launches;  /* with time_kernels: FIND-OR-PUT launches timed */
} gx_batch_stats;

void gx_batch_options_default(gx_batch_options *o);

/* ------------------------------------------------------------------ */
/* Serving: a continuous-batching request loop.  Jobs (model + property +
 * estimated state count) are queued by gx_server_submit into one of two
 * lanes by their estimate (small jobs, large jobs); the lanes have their own
 * slots, rounds and stream and run beside each other on the device.  Every
 * gx_server_step works each lane whose last replay has finished (waiting up
 * to about a millisecond for one): it retires finished jobs, admits waiting
 * jobs while an instance slot and a contiguous run of visited-set pages for
 * the job's estimated footprint are free, and starts a replay of
 * `rounds_per_step` BFS rounds of the lane's running jobs without waiting
 * for it.  A job whose partition (or its frontier, half the partition's
 * slots) fills up is restarted with twice the pages, so results equal
 * gx_explore's whatever the estimate.
 * Visited sets and frontiers of running jobs stay on the device between
 * steps.  No reference counterpart (the reference runs one model per
 * process). */
typedef struct gx_server gx_server;

typedef struct gx_server_options {
  uint32_t slots;            /* concurrent jobs (0 = 256) */
  uint32_t key_words;        /* key width of the pool: models up to 63*key_words state bits */
  uint32_t max_procs;        /* largest N accepted */
  uint32_t threads;          /* threads per workgroup */
  uint32_t blocks;           /* workgroups per launch (0 = auto) */
  uint32_t rounds_per_step;  /* BFS rounds per step (one graph replay; default 16) */
  uint64_t table_bytes;      /* visited-set pool */
  uint64_t page_bytes;       /* allocation granule of the pool */
  uint64_t cand_capacity;    /* successor buffer entries (0 = auto) */
  uint64_t hash_seed;
  uint32_t model_words;      /* largest compiled model accepted (32-bit words) */
  uint32_t por_jobs;         /* 1: POR jobs accepted (each round also runs the POR kernel) */
  uint32_t insert_blocks;    /* FIND-OR-PUT workgroups of 256 threads per round (0 = auto) */
  uint32_t large_slots;      /* slots of the large-job lane (at most a quarter of `slots`) */
  uint64_t large_states;     /* jobs estimated at this many states or more go to the large-job
                                lane (0: every job to the small-job lane) */
  uint32_t large_cus;        /* compute units the large-job lane's kernels may use (0 = seven
                                eighths; the device's count or more = all); the others stay free
                                for the small-job lane */
  uint32_t reserved;
} gx_server_options;

typedef struct gx_server_status {
  uint32_t running;          /* jobs on the device */
  uint32_t waiting;          /* jobs queued */
  uint32_t finished_pending; /* results not yet polled */
  uint32_t reserved;
  uint64_t steps, rounds, pages_used, pages_total;
} gx_server_status;

typedef struct gx_job_result {
  uint64_t job_id;
  gx_result result;          /* states, transitions, levels, violation, table_full */
  double t_submit, t_admit, t_done;  /* steady-clock seconds */
  int32_t status;            /* GX_OK or an error of this job */
  int32_t restarts;          /* re-admissions after a full partition */
} gx_job_result;

void gx_server_options_default(gx_server_options *o);
int gx_server_create(const gx_server_options *o, gx_server **out);
int gx_server_submit(gx_server *s, const gx_model *m, int32_t property, uint64_t est_states,
                     uint64_t *job_id);
/* Same with the job's options: POR jobs need por_jobs in the server options
 * (GX_E_ARG otherwise); with the cycle proviso, state_bits + 32 must fit the
 * pool's key_words (GX_E_LIMIT otherwise). */
int gx_server_submit_job(gx_server *s, const gx_model *m, const gx_job_options *job, uint64_t est_states,
                         uint64_t *job_id);
int gx_server_step(gx_server *s, gx_server_status *status);
int gx_server_poll(gx_server *s, gx_job_result *out, int max, int *n);
void gx_server_free(gx_server *s);

/* weights (optional, n entries): relative visited-set partition sizes, e.g.
 * estimated state counts; NULL splits the table evenly.  An instance whose
 * partition fills reports table_full like gx_explore. */
int gx_explore_batch(const gx_model *const *models, const int32_t *properties,
                     const uint64_t *weights, int n, const gx_batch_options *o,
                     gx_result *results, gx_batch_stats *stats);

/* Same with options per instance: plain instances run the state-per-lane
 * kernel, POR instances (cample, ample or stubborn, with or without the cycle
 * proviso; the reference's gather_por, cu:1191-1733) the process-group
 * kernel, both in every round.  Per instance the result equals gx_explore's
 * with the same options.  With the cycle proviso a slot carries a 32-bit
 * level stamp after the state bits, so the batch's key width grows to
 * ceil((state_bits + 32) / 63) words where needed. */
int gx_explore_batch_jobs(const gx_model *const *models, const gx_job_options *jobs,
                          const uint64_t *weights, int n, const gx_batch_options *o,
                          gx_result *results, gx_batch_stats *stats);

/* ------------------------------------------------------------------ */
/* Visited-set access (FIND-OR-PUT on its own), the hash-insert phase of the
 * exploration (cu:361-450 FINDORPUT_WARP, cu:656-704 init/count).  Keys are
 * `key_words` 64-bit words whose bit 63 is reserved: the library sets it on
 * every word, so keys differing only in bit 63 are the same key and words of
 * zero are valid. */
typedef struct gx_table gx_table;

int gx_table_create(uint64_t q_ints, uint32_t key_words, uint32_t bucket_bytes,
                    uint64_t hash_seed, gx_table **out);
void gx_table_free(gx_table *t);

/* Offers n keys (n * key_words words, host memory) to the table.  is_new
 * (optional, n bytes) receives 1 for keys that were inserted by this call.
 * *inserted receives the number of new keys.  Returns GX_E_CAPACITY when
 * the table is full. */
int gx_table_insert_host(gx_table *t, const uint64_t *keys, uint64_t n, uint8_t *is_new,
                         uint64_t *inserted);

/* Same with keys already resident in device memory (a device pointer; read
 * only, bit 63 is set on the fly); timed with HIP events into *kernel_ms when
 * non-NULL. */
int gx_table_insert_device(gx_table *t, const uint64_t *dev_keys, uint64_t n,
                           uint64_t *inserted, float *kernel_ms);

/* gx_table_insert_host without any device allocation: no counterpart in the
 * reference; it is for callers that cannot allocate (serving loops,
 * harnesses that forbid hipMalloc).  Keys, is-new bytes and the control block
 * are staged in a fixed scratch array of the library, one per device (about
 * 10.3 MiB, held until the process ends), so at most GX_NOALLOC_MAX_KEYS keys
 * per call (GX_E_SCRATCH beyond).  The table must live on the current device.
 * Calls on one device share its scratch: the library serialises them, and
 * each call has finished with the scratch when it returns. */
#define GX_NOALLOC_MAX_KEYS 262144u
int gx_table_insert_host_noalloc(gx_table *t, const uint64_t *keys, uint64_t n, uint8_t *is_new,
                                 uint64_t *inserted);

/* Number of occupied slots (cu:687-704). */
int gx_table_count(gx_table *t, uint64_t *count);

/* Number of visible HIP devices (0 when none; never fails). */
int gx_device_count(int *count);

/* Device buffer helpers for callers without their own allocator. */
int gx_device_alloc(uint64_t bytes, void **dev_ptr);
int gx_device_free(void *dev_ptr);
int gx_copy_to_device(void *dev_dst, const void *host_src, uint64_t bytes);
int gx_copy_to_host(void *host_dst, const void *dev_src, uint64_t bytes);

/* Measured bandwidth figures beside the published HBM peak, timed with HIP
 * events over `reps` launches (best launch shape of a small sweep):
 * gx_bench_copy   streaming copy of `bytes` bytes device to device (GB/s,
 *                 bytes read + bytes written);
 * gx_bench_read   streaming read of `bytes` bytes (GB/s);
 * gx_bench_random `lookups` reads of random 64-byte buckets of a table of
 *                 `table_bytes` bytes, four lanes of 16 bytes per bucket as
 *                 the FIND-OR-PUT reads them (lookups per second). */
int gx_bench_copy(uint64_t bytes, int reps, double *gbps);
int gx_bench_read(uint64_t bytes, int reps, double *gbps);
int gx_bench_random(uint64_t table_bytes, uint64_t lookups, int reps, double *lookups_per_s);

#ifdef __cplusplus
}
#endif
#endif
