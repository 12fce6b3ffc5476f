/*
 * gx_oracle.h -- CPU oracle for the GPUexplore gather/explore path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the exploration
 * semantics of the reference (GPUexplore/src/GPUexplore.cu, cited as cu:N
 * below).  It is the checker the HIP path is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped: the product library
 * (gpuexplore_amd/csrc) does not link, load or call anything here.
 *
 * Parity pins (what fixes each function independently of this code):
 *   orc_load_gpf        hand-computed gpf fixtures under tests/golden/
 *   orc_get/orc_set     exhaustive brute force against a naive bit array
 *   orc_successors      a separate pure-Python enumerator written from the
 *                       definitions (tests/naive_enum.py) on tiny models, and
 *                       closed forms (independent cycles, chains, barriers)
 *   orc_explore (full)  same + counts invariant under the hash seed
 *   orc_explore (POR)   closed form for independent chains (N*k+1 states),
 *                       reduced set is a subset of the full set, deadlock
 *                       verdict preserved; brute-force recomputation of the
 *                       cluster rules in Python on tiny models
 *   proviso (B6)        brute-force Python restatement on tiny models
 *
 * Semantics followed (where the reference is schedule-dependent or skips
 * work, the intended semantics below are used instead):
 *   - every enabled synchronisation rule fires once (cu:1480-1488);
 *   - T_p(a) is the concatenation of the targets of p's a-entries;
 *   - level-synchronous BFS that stops at the end of the level in which the
 *     first violation is found;
 *   - cycle proviso: a successor is "old" iff it was visited at a level <=
 *     the level being expanded.
 */
#ifndef GX_ORACLE_H
#define GX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_PROCS 32
#define ORC_MAX_WORDS 9

typedef struct orc_model {
  uint32_t liveness;          /* gpf line 1 */
  uint32_t nr_procs;          /* N */
  uint32_t bits_act;
  uint32_t proc_nrstates;
  uint32_t bits_statevector;  /* +1 under POR, as the reference loader does */
  uint32_t sv_nints;
  uint32_t *firstbit;         /* N+1 */
  uint32_t *bits_state;       /* N   */
  uint32_t nbits_offset;
  uint32_t max_buf_ints;
  uint32_t *proc_offsets_start; /* N+1 */
  uint32_t *proc_offsets;
  uint32_t nr_trans;
  uint32_t *proc_trans;
  uint32_t nbits_syncbits_offset;
  uint32_t nr_syncbits_offsets;
  uint32_t *syncbits_offsets;
  uint32_t nr_syncbits;
  uint32_t *syncbits;
  uint32_t words;             /* key width in 32-bit words: ceil(fb[N]/32) */
} orc_model;

enum { ORC_PROP_NONE = 0, ORC_PROP_DEADLOCK = 1, ORC_PROP_SAFETY = 2, ORC_PROP_LIVENESS = 3 };

enum {
  ORC_OK = 0,
  ORC_ERR_NOFILE = -1,
  ORC_ERR_FORMAT = -2,
  ORC_ERR_RANGE = -3,
  ORC_ERR_NOMEM = -4,
  ORC_ERR_LIMIT = -5,
  ORC_ERR_ARG = -6
};

/* partial-order reduction methods (the reference keeps each on its own
 * branch, README.md lines 24-26; this snapshot holds cample only) */
enum { ORC_POR_CAMPLE = 0, ORC_POR_AMPLE = 1, ORC_POR_STUBBORN = 2 };

typedef struct orc_options {
  int property;       /* ORC_PROP_* */
  int por;
  int cycle_proviso;
  uint64_t max_states; /* 0 = unbounded; else stop after the level that crosses it */
  uint64_t hash_seed;
  int por_method;     /* ORC_POR_*, used when por is set */
  uint64_t expected_states; /* presizes the visited set (0: grow from a small table) */
  int progress;       /* orc_explore_mt: a line on stderr after each level */
} orc_options;

typedef struct orc_result {
  uint64_t states;
  uint64_t transitions;
  uint64_t levels;
  uint64_t full_successors; /* successors generated by full expansion (POR: both ample and not) */
  int violation;
  int truncated;            /* max_states bound reached */
  double seconds;
} orc_result;

int orc_load_gpf(const char *path_noext, int por, orc_model **out);
void orc_free(orc_model *m);

uint32_t orc_get(const orc_model *m, const uint32_t *vec, uint32_t p);
void orc_set(const orc_model *m, uint32_t *vec, uint32_t p, uint32_t x);

/* All successors of `state` by full expansion, in generation order
 * (locals of process 0..N-1, then rules by ascending label).  Writes up to
 * `cap` vectors of m->words words each; returns the total count (may exceed
 * cap) or a negative error.  *deadlock is set per cu:1134-1151. */
long orc_successors(const orc_model *m, const uint32_t *state, uint32_t *out,
                    size_t cap, int *deadlock);

/* cample cluster selected for `state` without the cycle proviso (bitmask of
 * processes; all-ones if no reduction applies). */
uint64_t orc_por_cluster(const orc_model *m, const uint32_t *state);

/* process set selected for `state` by ORC_POR_* `method` without the cycle
 * proviso (all-ones if no reduction applies) */
uint64_t orc_por_select(const orc_model *m, const uint32_t *state, int method);

int orc_explore(const orc_model *m, const orc_options *opt, orc_result *res);

/* orc_explore on `nthreads` threads for the plain method (no POR):
 * identical counts, sharded visited set; ORC_ERR_ARG under POR. */
int orc_explore_mt(const orc_model *m, const orc_options *opt, int nthreads, orc_result *res);

#ifdef __cplusplus
}
#endif
#endif
