/*
 * gx_oracle.c -- CPU oracle for the GPUexplore gather/explore path.
 *
 * TEST INFRASTRUCTURE ONLY (see gx_oracle.h).  Plain C, single thread.
 * Every decode step follows the reference kernel's arithmetic on the raw gpf
 * arrays (cu:N = GPUexplore/src/GPUexplore.cu line N); nothing is
 * precompiled, so this stays independent of the product's model compiler.
 */
#include "gx_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* bit helpers (cu:153-161, 205): (x >> start) & ((1 << len) - 1), with
 * len == 32 handled instead of being undefined. */
static inline uint32_t getbits(uint32_t x, uint32_t start, uint32_t len) {
  if (len == 0 || start >= 32) return 0;
  uint64_t mask = (len >= 32) ? 0xFFFFFFFFull : ((1ull << len) - 1ull);
  return (uint32_t)(((uint64_t)x >> start) & mask);
}

/* ------------------------------------------------------------------ */
/* gpf loader (cu:1885-1983).  fgets into a 256-byte buffer, then atoi,
 * which on glibc is (int)strtol: values up to 2^32-1 wrap into uint32. */
typedef struct {
  FILE *fp;
  int eof;
} rdr;

static int rd(rdr *r, uint32_t *v) {
  char buf[256];
  if (!fgets(buf, sizeof buf, r->fp)) {
    r->eof = 1;
    return ORC_ERR_FORMAT;
  }
  *v = (uint32_t)strtol(buf, NULL, 10);
  return ORC_OK;
}

static int rd_arr(rdr *r, uint32_t n, uint32_t **out) {
  uint32_t *a = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  if (!a) return ORC_ERR_NOMEM;
  for (uint32_t i = 0; i < n; i++) {
    if (rd(r, &a[i]) != ORC_OK) {
      free(a);
      return ORC_ERR_FORMAT;
    }
  }
  *out = a;
  return ORC_OK;
}

void orc_free(orc_model *m) {
  if (!m) return;
  free(m->firstbit);
  free(m->bits_state);
  free(m->proc_offsets_start);
  free(m->proc_offsets);
  free(m->proc_trans);
  free(m->syncbits_offsets);
  free(m->syncbits);
  free(m);
}

int orc_load_gpf(const char *path_noext, int por, orc_model **out) {
  char fn[4096];
  snprintf(fn, sizeof fn, "%s.gpf", path_noext); /* cu:1814-1815 */
  FILE *fp = fopen(fn, "r");
  if (!fp) return ORC_ERR_NOFILE; /* cu:1984-1986 */
  orc_model *m = (orc_model *)calloc(1, sizeof *m);
  rdr r = {fp, 0};
  int rc = ORC_ERR_FORMAT;
  uint32_t n_offsets;
  if (!m) {
    fclose(fp);
    return ORC_ERR_NOMEM;
  }
  if (rd(&r, &m->liveness)) goto fail;              /* cu:1888 */
  if (rd(&r, &m->nr_procs)) goto fail;              /* cu:1899 */
  if (rd(&r, &m->bits_act)) goto fail;              /* cu:1903 */
  if (rd(&r, &m->proc_nrstates)) goto fail;         /* cu:1907 */
  if (rd(&r, &m->bits_statevector)) goto fail;      /* cu:1911 */
  m->bits_statevector += por ? 1u : 0u;             /* cu:1912 */
  if (m->nr_procs == 0 || m->nr_procs > ORC_MAX_PROCS) {
    rc = ORC_ERR_LIMIT;
    goto fail;
  }
  if ((rc = rd_arr(&r, m->nr_procs + 1, &m->firstbit))) goto fail; /* cu:1915-1921 */
  m->sv_nints = (m->bits_statevector + 31) / 32;                   /* cu:1923 */
  if ((rc = rd_arr(&r, m->nr_procs, &m->bits_state))) goto fail;   /* cu:1924-1930 */
  rc = ORC_ERR_FORMAT;
  if (rd(&r, &m->nbits_offset)) goto fail;          /* cu:1931 */
  if (rd(&r, &m->max_buf_ints)) goto fail;          /* cu:1935 */
  if ((rc = rd_arr(&r, m->nr_procs + 1, &m->proc_offsets_start))) goto fail; /* cu:1939-1944 */
  n_offsets = m->proc_offsets_start[m->nr_procs];
  if ((rc = rd_arr(&r, n_offsets, &m->proc_offsets))) goto fail;  /* cu:1945-1950 */
  rc = ORC_ERR_FORMAT;
  if (rd(&r, &m->nr_trans)) goto fail;              /* cu:1951 */
  if ((rc = rd_arr(&r, m->nr_trans, &m->proc_trans))) goto fail;  /* cu:1955-1960 */
  rc = ORC_ERR_FORMAT;
  if (rd(&r, &m->nbits_syncbits_offset)) goto fail; /* cu:1962 */
  if (rd(&r, &m->nr_syncbits_offsets)) goto fail;   /* cu:1965 */
  if ((rc = rd_arr(&r, m->nr_syncbits_offsets, &m->syncbits_offsets))) goto fail; /* cu:1968-1973 */
  rc = ORC_ERR_FORMAT;
  if (rd(&r, &m->nr_syncbits)) goto fail;           /* cu:1974 */
  if ((rc = rd_arr(&r, m->nr_syncbits, &m->syncbits))) goto fail; /* cu:1977-1982 */
  fclose(fp);
  /* consistency needed for safe decoding */
  for (uint32_t p = 0; p < m->nr_procs; p++) {
    if (m->firstbit[p + 1] < m->firstbit[p] || m->firstbit[p + 1] - m->firstbit[p] > 32) {
      orc_free(m);
      return ORC_ERR_RANGE;
    }
    if (m->bits_state[p] == 0 || m->bits_state[p] > 31) {
      orc_free(m);
      return ORC_ERR_RANGE;
    }
  }
  if (m->firstbit[m->nr_procs] > 32u * ORC_MAX_WORDS || m->nbits_offset == 0 ||
      m->nbits_offset > 32 || m->nbits_syncbits_offset == 0 || m->nbits_syncbits_offset > 32 ||
      m->bits_act > 30) {
    orc_free(m);
    return ORC_ERR_RANGE;
  }
  m->words = (m->firstbit[m->nr_procs] + 31) / 32;
  if (m->words == 0) m->words = 1;
  *out = m;
  return ORC_OK;
fail:
  fclose(fp);
  orc_free(m);
  return rc;
}

/* ------------------------------------------------------------------ */
/* vector codec: GETSTATEVECTORSTATE / SETSTATEVECTORSTATE (cu:162-178).
 * Field p occupies bits [fb[p], fb[p+1]) of the little-endian word array
 * and may straddle two words. */
uint32_t orc_get(const orc_model *m, const uint32_t *v, uint32_t p) {
  uint32_t lo = m->firstbit[p], hi = m->firstbit[p + 1], w = hi - lo;
  if (w == 0) return 0;
  uint64_t x = v[lo / 32];
  if (lo / 32 != (hi - 1) / 32) x |= (uint64_t)v[lo / 32 + 1] << 32;
  x >>= (lo % 32);
  return (uint32_t)(x & ((w >= 32) ? 0xFFFFFFFFull : ((1ull << w) - 1ull)));
}

void orc_set(const orc_model *m, uint32_t *v, uint32_t p, uint32_t x) {
  uint32_t lo = m->firstbit[p], hi = m->firstbit[p + 1], w = hi - lo;
  if (w == 0) return;
  int straddle = (lo / 32 != (hi - 1) / 32);
  uint64_t mask = ((w >= 32) ? 0xFFFFFFFFull : ((1ull << w) - 1ull)) << (lo % 32);
  uint64_t cur = v[lo / 32];
  if (straddle) cur |= (uint64_t)v[lo / 32 + 1] << 32;
  cur = (cur & ~mask) | (((uint64_t)x << (lo % 32)) & mask);
  v[lo / 32] = (uint32_t)cur;
  if (straddle) v[lo / 32 + 1] = (uint32_t)(cur >> 32);
}

/* ------------------------------------------------------------------ */
/* transition offsets of process p in local state s (cu:883-897) */
static int proc_range(const orc_model *m, uint32_t p, uint32_t s, uint32_t *o1, uint32_t *o2) {
  uint32_t nb = m->nbits_offset, k = 32 / nb;
  uint32_t index = s / k, pos = s - index * k;
  uint64_t at = (uint64_t)m->proc_offsets_start[p] + index;
  uint32_t n_off = m->proc_offsets_start[m->nr_procs];
  if (at >= n_off) return ORC_ERR_RANGE;
  uint32_t tmp = m->proc_offsets[at];
  *o1 = getbits(tmp, pos * nb, nb);
  if (pos == k - 1) {
    if (at + 1 >= n_off) return ORC_ERR_RANGE;
    *o2 = getbits(m->proc_offsets[at + 1], 0, nb);
  } else {
    *o2 = getbits(tmp, (pos + 1) * nb, nb);
  }
  if (*o1 > *o2 || *o2 > m->nr_trans) return ORC_ERR_RANGE;
  return ORC_OK;
}

/* transition entry fields (cu:143, 157-159) */
static inline uint32_t entry_sync(uint32_t t) { return t & 1u; }
static inline uint32_t entry_label(const orc_model *m, uint32_t t) { return getbits(t, 1, m->bits_act); }
static inline uint32_t entry_capacity(const orc_model *m, uint32_t p) {
  return (31u - m->bits_act) / m->bits_state[p];
}
/* target i of entry t for process p, as stored: state id + 1, 0 = end */
static inline uint32_t entry_target(const orc_model *m, uint32_t t, uint32_t i, uint32_t p) {
  return getbits(t, 1 + m->bits_act + i * m->bits_state[p], m->bits_state[p]);
}

/* rules of label a: syncbits ints [so1, so2) (cu:999-1008, 1466-1476);
 * N-bit rule fields, floor(32/N) per int (cu:205, 1480-1487) */
static int label_rules(const orc_model *m, uint32_t a, uint32_t *so1, uint32_t *so2) {
  uint32_t nb = m->nbits_syncbits_offset, k = 32 / nb;
  uint32_t i = a / k, pos = a - i * k;
  if (i >= m->nr_syncbits_offsets) return ORC_ERR_RANGE;
  uint32_t l = m->syncbits_offsets[i];
  *so1 = getbits(l, pos * nb, nb);
  if (pos == k - 1) {
    if (i + 1 >= m->nr_syncbits_offsets) return ORC_ERR_RANGE;
    *so2 = getbits(m->syncbits_offsets[i + 1], 0, nb);
  } else {
    *so2 = getbits(l, (pos + 1) * nb, nb);
  }
  if (*so1 > *so2 || *so2 > m->nr_syncbits) return ORC_ERR_RANGE;
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* per-state decoded view */
#define MAXT 4096
typedef struct {
  uint32_t n;
  uint32_t sp[ORC_MAX_PROCS], o1[ORC_MAX_PROCS], o2[ORC_MAX_PROCS];
  uint32_t nlabels;
  uint32_t labels[MAXT]; /* distinct sync labels present, ascending */
} sview;

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

static int view_state(const orc_model *m, const uint32_t *s, sview *v) {
  v->n = m->nr_procs;
  v->nlabels = 0;
  for (uint32_t p = 0; p < v->n; p++) {
    v->sp[p] = orc_get(m, s, p);
    int rc = proc_range(m, p, v->sp[p], &v->o1[p], &v->o2[p]);
    if (rc) return rc;
    for (uint32_t e = v->o1[p]; e < v->o2[p]; e++) {
      uint32_t t = m->proc_trans[e];
      if (entry_sync(t)) {
        if (v->nlabels >= MAXT) return ORC_ERR_LIMIT;
        v->labels[v->nlabels++] = entry_label(m, t);
      }
    }
  }
  qsort(v->labels, v->nlabels, sizeof(uint32_t), cmp_u32);
  uint32_t u = 0;
  for (uint32_t i = 0; i < v->nlabels; i++)
    if (u == 0 || v->labels[u - 1] != v->labels[i]) v->labels[u++] = v->labels[i];
  v->nlabels = u;
  return ORC_OK;
}

/* T_p(a): concatenated targets (value id+1 decoded to id) of p's a-entries */
static uint32_t sync_targets(const orc_model *m, const sview *v, uint32_t p, uint32_t a,
                             uint32_t *out, uint32_t cap) {
  uint32_t n = 0, c = entry_capacity(m, p);
  for (uint32_t e = v->o1[p]; e < v->o2[p]; e++) {
    uint32_t t = m->proc_trans[e];
    if (!entry_sync(t) || entry_label(m, t) != a) continue;
    for (uint32_t i = 0; i < c; i++) {
      uint32_t x = entry_target(m, t, i, p);
      if (x == 0) break;
      if (n < cap) out[n] = x - 1;
      n++;
    }
  }
  return n;
}

static int has_label(const orc_model *m, const sview *v, uint32_t p, uint32_t a) {
  for (uint32_t e = v->o1[p]; e < v->o2[p]; e++) {
    uint32_t t = m->proc_trans[e];
    if (entry_sync(t) && entry_label(m, t) == a) return 1;
  }
  return 0;
}

/* successor visitor: kind 0 = local move of process `who`, kind 1 = rule `rule` */
typedef int (*succ_fn)(void *ctx, const uint32_t *t, int kind, uint32_t who, uint64_t rule);

typedef struct {
  int any_local_entry;    /* cu:910-913: a local entry exists */
  int any_enabled_rule;   /* cu:1033-1035 */
  uint64_t cluster0[ORC_MAX_PROCS];   /* cu:1347, 1616-1624 */
  uint64_t local_succ;    /* processes with >= 1 local successor */
  uint64_t in_enabled;    /* processes in some enabled rule with >= 1 successor */
  uint32_t nloc[ORC_MAX_PROCS];       /* local successors per process */
  uint64_t stub[ORC_MAX_PROCS];       /* stubborn-set relation d(p), see select_stubborn */
} sinfo;

/* Enumerate successors of s.  Locals of processes in `allow`, then for every
 * sync label in ascending order every enabled rule r with r subset of `allow`
 * (cu:904-949, 951-1130; POR ample filter cu:1409, 1541).  Fills `info` from
 * the full (unfiltered) enabledness. */
static int enumerate(const orc_model *m, const uint32_t *s, const sview *v, uint64_t allow,
                     succ_fn fn, void *ctx, sinfo *info) {
  uint32_t W = m->words, N = m->nr_procs;
  uint32_t t[ORC_MAX_WORDS];
  static __thread uint32_t tg[ORC_MAX_PROCS][MAXT]; /* per thread: orc_explore_mt */
  uint32_t tn[ORC_MAX_PROCS];
  memset(info, 0, sizeof *info);
  for (uint32_t p = 0; p < N; p++) info->cluster0[p] = info->stub[p] = 1ull << p;
  /* local successors */
  for (uint32_t p = 0; p < N; p++) {
    uint32_t c = entry_capacity(m, p);
    for (uint32_t e = v->o1[p]; e < v->o2[p]; e++) {
      uint32_t ent = m->proc_trans[e];
      if (entry_sync(ent)) continue;
      info->any_local_entry = 1;
      for (uint32_t i = 0; i < c; i++) {
        uint32_t x = entry_target(m, ent, i, p);
        if (x == 0) break;
        info->local_succ |= 1ull << p;
        info->nloc[p]++;
        if (!((allow >> p) & 1)) continue;
        memcpy(t, s, W * sizeof(uint32_t));
        orc_set(m, t, p, x - 1);
        if (fn) {
          int rc = fn(ctx, t, 0, p, 0);
          if (rc) return rc;
        }
      }
    }
  }
  /* synchronising successors, labels ascending */
  for (uint32_t li = 0; li < v->nlabels; li++) {
    uint32_t a = v->labels[li];
    uint64_t en = 0;
    for (uint32_t p = 0; p < N; p++)
      if (has_label(m, v, p, a)) en |= 1ull << p;
    uint32_t so1, so2;
    int rc = label_rules(m, a, &so1, &so2);
    if (rc) return rc;
    uint32_t per = 32 / N;
    for (uint32_t j = so1; j < so2; j++) {
      for (uint32_t f = 0; f < per; f++) {
        uint64_t r = getbits(m->syncbits[j], f * N, N);
        if (r == 0) continue;
        /* a member that has the label: a rule all members have joins it
         * whole; otherwise its lowest member without the label (the moves
         * that can enable the rule) */
        const uint64_t lack = r & ~en;
        for (uint32_t p = 0; p < N; p++)
          if (((r >> p) & 1) && ((en >> p) & 1)) {
            info->cluster0[p] |= r;
            info->stub[p] |= lack ? (lack & (~lack + 1)) : r;
          }
        if (r & ~en) continue; /* not enabled */
        info->any_enabled_rule = 1;
        /* product of target lists, lowest process varies fastest (cu:1084-1113) */
        uint32_t mem[ORC_MAX_PROCS], nm = 0;
        uint64_t prod = 1;
        for (uint32_t p = 0; p < N; p++) {
          if (!((r >> p) & 1)) continue;
          tn[p] = sync_targets(m, v, p, a, tg[p], MAXT);
          if (tn[p] > MAXT) return ORC_ERR_LIMIT;
          mem[nm++] = p;
          prod *= tn[p];
        }
        if (prod == 0) continue;
        info->in_enabled |= r;
        if (r & ~allow) continue;
        uint32_t dig[ORC_MAX_PROCS] = {0};
        memcpy(t, s, W * sizeof(uint32_t));
        for (uint32_t i = 0; i < nm; i++) orc_set(m, t, mem[i], tg[mem[i]][0]);
        for (;;) {
          if (fn) {
            rc = fn(ctx, t, 1, 0, r);
            if (rc) return rc;
          }
          uint32_t i = 0;
          for (; i < nm; i++) {
            uint32_t p = mem[i];
            if (++dig[i] < tn[p]) {
              orc_set(m, t, p, tg[p][dig[i]]);
              break;
            }
            dig[i] = 0;
            orc_set(m, t, p, tg[p][0]);
          }
          if (i == nm) break;
        }
      }
    }
  }
  return ORC_OK;
}

typedef struct {
  uint32_t *out;
  size_t cap, n;
  uint32_t W;
} collect_ctx;

static int collect(void *c, const uint32_t *t, int kind, uint32_t who, uint64_t rule) {
  (void)kind;
  (void)who;
  (void)rule;
  collect_ctx *cc = (collect_ctx *)c;
  if (cc->n < cc->cap) memcpy(cc->out + cc->n * cc->W, t, cc->W * sizeof(uint32_t));
  cc->n++;
  return 0;
}

long orc_successors(const orc_model *m, const uint32_t *state, uint32_t *out, size_t cap,
                    int *deadlock) {
  sview v;
  sinfo info;
  int rc = view_state(m, state, &v);
  if (rc) return rc;
  collect_ctx cc = {out, cap, 0, m->words};
  rc = enumerate(m, state, &v, ~0ull, collect, &cc, &info);
  if (rc) return rc;
  if (deadlock) *deadlock = !(info.any_local_entry || info.any_enabled_rule); /* cu:1134-1151 */
  return (long)cc.n;
}

/* ------------------------------------------------------------------ */
/* cample cluster selection (cu:1633-1672) from cluster0 and prov */
static uint64_t select_cluster(uint32_t N, const uint64_t *cluster0, uint64_t prov) {
  uint64_t c[ORC_MAX_PROCS], fin[ORC_MAX_PROCS];
  uint64_t all = (N >= 64) ? ~0ull : ((1ull << N) - 1ull);
  for (uint32_t p = 0; p < N; p++)
    c[p] = ((prov >> p) & 1) ? cluster0[p] : (cluster0[p] & ~(1ull << p)); /* cu:1635-1637 */
  for (uint32_t p = 0; p < N; p++) {
    uint64_t K = c[p], done = 0;
    int ok = 0;
    /* closure: K is closed under i in K => c(i) subset of K (cu:1641-1649) */
    while (K & ~done) {
      uint32_t i = (uint32_t)__builtin_ctzll(K & ~done);
      done |= 1ull << i;
      if ((c[i] >> i) & 1) ok = 1;
      K |= c[i];
    }
    fin[p] = ok ? K : 0; /* cu:1651-1655 */
  }
  int best = (int)N;
  uint64_t C = all; /* cu:1658-1659 */
  for (uint32_t p = 0; p < N; p++) {
    int pc = __builtin_popcountll(fin[p]);
    if (fin[p] && pc < best) { /* first minimal, strictly smaller than N (cu:1661) */
      best = pc;
      C = fin[p];
    }
  }
  return C;
}

/* Ample-set method: the local moves of one process p whose state has no
 * synchronising entry with a rule containing p (cluster0(p) = {p}): no
 * other process can enable a transition dependent on them.  Candidates must
 * be in `cand` (a local successor; a new one under the cycle proviso); the
 * candidate with the fewest local successors is taken, the lowest on ties;
 * none gives full expansion. */
static uint64_t select_ample(uint32_t N, const sinfo *info, uint64_t cand) {
  uint64_t all = (N >= 64) ? ~0ull : ((1ull << N) - 1ull);
  uint32_t best = 0xFFFFFFFFu;
  uint64_t C = all;
  for (uint32_t p = 0; p < N; p++) {
    if (!((cand >> p) & 1) || info->cluster0[p] != (1ull << p)) continue;
    if (info->nloc[p] < best) {
      best = info->nloc[p];
      C = 1ull << p;
    }
  }
  return C;
}

/* Stubborn-set method over processes.  d(p) = {p} plus, for every rule r
 * containing p of a label p has: r itself when every member has the label,
 * else r's lowest member without it (a necessary enabling set).  K(p) is
 * the closure of {p} under d; the transitions of K(p) (local moves of its
 * processes, rules meeting it) form a stubborn set: an enabled rule meeting
 * K lies inside K, a disabled one has a blocking process in K.  K(p) is
 * valid when it meets `target` (processes with an enabled transition; with
 * a new successor under the cycle proviso); the smallest valid K is taken,
 * the lowest p on ties; none, or one of all N processes, gives full
 * expansion. */
static uint64_t select_stubborn(uint32_t N, const sinfo *info, uint64_t target) {
  uint64_t all = (N >= 64) ? ~0ull : ((1ull << N) - 1ull);
  int best = (int)N;
  uint64_t C = all;
  for (uint32_t p = 0; p < N; p++) {
    uint64_t K = 1ull << p, done = 0;
    while (K & ~done) {
      uint32_t i = (uint32_t)__builtin_ctzll(K & ~done);
      done |= 1ull << i;
      K |= info->stub[i];
    }
    if (!(K & target)) continue;
    int pc = __builtin_popcountll(K);
    if (pc < best) {
      best = pc;
      C = K;
    }
  }
  return C;
}

uint64_t orc_por_select(const orc_model *m, const uint32_t *state, int method) {
  sview v;
  sinfo info;
  if (view_state(m, state, &v)) return 0;
  if (enumerate(m, state, &v, 0, NULL, NULL, &info)) return 0;
  if (method == ORC_POR_AMPLE) return select_ample(m->nr_procs, &info, info.local_succ);
  if (method == ORC_POR_STUBBORN) return select_stubborn(m->nr_procs, &info, info.local_succ | info.in_enabled);
  return select_cluster(m->nr_procs, info.cluster0, info.local_succ | info.in_enabled);
}

uint64_t orc_por_cluster(const orc_model *m, const uint32_t *state) {
  sview v;
  sinfo info;
  if (view_state(m, state, &v)) return 0;
  if (enumerate(m, state, &v, 0, NULL, NULL, &info)) return 0;
  return select_cluster(m->nr_procs, info.cluster0, info.local_succ | info.in_enabled);
}

/* ------------------------------------------------------------------ */
/* visited set: open addressing, key = W words.  With levels, lvl[slot] =
 * level + 1 (0 = empty); the cycle proviso needs them.  Without (plain
 * exploration and POR without the proviso) a bitmap marks occupied slots:
 * 4 * W bytes and one bit per slot, so a presized set for the 945 M states of
 * the 24-process POR workload explored unreduced fits in host memory. */
typedef struct {
  uint32_t W;
  uint64_t cap, n, mask, seed;
  uint32_t *keys;
  uint32_t *lvl; /* NULL: level-free, occ marks the slots */
  uint64_t *occ;
} vset;

static inline uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

static inline uint64_t vhash(const vset *s, const uint32_t *k) {
  uint64_t h = s->seed ^ 0x9e3779b97f4a7c15ull;
  for (uint32_t i = 0; i < s->W; i++) h = mix64(h ^ k[i]) + i;
  return h;
}

/* capacity: a power of two holding `expect` keys at load <= 1/2 */
static uint64_t vset_cap_for(uint64_t expect, uint64_t min_cap) {
  uint64_t c = min_cap;
  while (c < 2 * expect) c <<= 1;
  return c;
}

static int vset_init(vset *s, uint32_t W, uint64_t cap, uint64_t seed, int levels) {
  s->W = W;
  s->cap = cap;
  s->mask = cap - 1;
  s->n = 0;
  s->seed = seed;
  s->keys = (uint32_t *)malloc(cap * W * sizeof(uint32_t));
  s->lvl = levels ? (uint32_t *)calloc(cap, sizeof(uint32_t)) : NULL;
  s->occ = levels ? NULL : (uint64_t *)calloc((cap + 63) / 64, sizeof(uint64_t));
  return (s->keys && (s->lvl || s->occ)) ? ORC_OK : ORC_ERR_NOMEM;
}

static void vset_free(vset *s) {
  free(s->keys);
  free(s->lvl);
  free(s->occ);
}

/* stored level + 1 of slot i (1 for an occupied slot of a level-free set),
 * 0 if empty */
static inline uint32_t vset_at(const vset *s, uint64_t i) {
  if (s->lvl) return s->lvl[i];
  return (uint32_t)((s->occ[i >> 6] >> (i & 63)) & 1u);
}

static inline void vset_mark(vset *s, uint64_t i, uint32_t level_plus_1) {
  if (s->lvl)
    s->lvl[i] = level_plus_1;
  else
    s->occ[i >> 6] |= 1ull << (i & 63);
}

/* returns stored level+1 if present (1 without levels), 0 if absent; *slot
 * gets the slot */
static uint32_t vset_probe(const vset *s, const uint32_t *k, uint64_t *slot) {
  uint64_t i = vhash(s, k) & s->mask;
  for (;;) {
    uint32_t l = vset_at(s, i);
    if (l == 0) {
      *slot = i;
      return 0;
    }
    if (memcmp(s->keys + i * s->W, k, s->W * sizeof(uint32_t)) == 0) {
      *slot = i;
      return l;
    }
    i = (i + 1) & s->mask;
  }
}

static int vset_grow(vset *s) {
  vset t;
  int rc = vset_init(&t, s->W, s->cap * 2, s->seed, s->lvl != NULL);
  if (rc) {
    vset_free(&t);
    return rc;
  }
  for (uint64_t i = 0; i < s->cap; i++) {
    const uint32_t l = vset_at(s, i);
    if (!l) continue;
    uint64_t sl;
    vset_probe(&t, s->keys + i * s->W, &sl);
    memcpy(t.keys + sl * s->W, s->keys + i * s->W, s->W * sizeof(uint32_t));
    vset_mark(&t, sl, l);
    t.n++;
  }
  vset_free(s);
  *s = t;
  return ORC_OK;
}

/* insert with level; returns 1 if new, 0 if present, <0 on error */
static int vset_put(vset *s, const uint32_t *k, uint32_t level) {
  if ((s->n + 1) * 2 > s->cap) {
    int rc = vset_grow(s);
    if (rc) return rc;
  }
  uint64_t sl;
  if (vset_probe(s, k, &sl)) return 0;
  memcpy(s->keys + sl * s->W, k, s->W * sizeof(uint32_t));
  vset_mark(s, sl, level + 1);
  s->n++;
  return 1;
}

/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t *v;
  uint64_t n, cap;
  uint32_t W;
} vlist;

static int vlist_push(vlist *l, const uint32_t *k) {
  if (l->n == l->cap) {
    uint64_t nc = l->cap ? l->cap * 2 : 1024;
    uint32_t *nv = (uint32_t *)realloc(l->v, nc * l->W * sizeof(uint32_t));
    if (!nv) return ORC_ERR_NOMEM;
    l->v = nv;
    l->cap = nc;
  }
  memcpy(l->v + l->n * l->W, k, l->W * sizeof(uint32_t));
  l->n++;
  return ORC_OK;
}

typedef struct {
  const orc_model *m;
  const orc_options *o;
  vset *set;
  vlist *next;
  uint32_t level;      /* level being expanded */
  uint64_t transitions, full;
  int violation;
  int check_safety;    /* safety checked in on_succ (full expansion only) */
  /* proviso bookkeeping */
  uint64_t local_new;  /* processes with a new local successor */
  uint64_t rules_new[MAXT];
  uint64_t nrules_new;
} bfs_ctx;

static int is_violating(const orc_model *m, const uint32_t *t) {
  return orc_get(m, t, m->nr_procs - 1) == 1; /* cu:926-935, 1063-1069 */
}

/* ample / full successor: count it and insert it */
static int on_succ(void *c, const uint32_t *t, int kind, uint32_t who, uint64_t rule) {
  (void)kind;
  (void)who;
  (void)rule;
  bfs_ctx *b = (bfs_ctx *)c;
  b->transitions++;
  if (b->check_safety && is_violating(b->m, t)) b->violation = 1;
  int rc = vset_put(b->set, t, b->level + 1);
  if (rc < 0) return rc;
  if (rc == 1) return vlist_push(b->next, t);
  return ORC_OK;
}

/* full expansion visitor under POR: safety check on every successor
 * (cu:1384-1393, 1519-1525) */
static int on_full(void *c, const uint32_t *t, int kind, uint32_t who, uint64_t rule) {
  (void)kind;
  (void)who;
  (void)rule;
  bfs_ctx *b = (bfs_ctx *)c;
  b->full++;
  if (b->o->property == ORC_PROP_SAFETY && is_violating(b->m, t)) b->violation = 1;
  return ORC_OK;
}

/* cycle proviso visitor: a successor is new unless visited at level <= L */
static int on_proviso(void *c, const uint32_t *t, int kind, uint32_t who, uint64_t rule) {
  bfs_ctx *b = (bfs_ctx *)c;
  uint64_t sl;
  uint32_t l = vset_probe(b->set, t, &sl);
  int isnew = !(l != 0 && (l - 1) <= b->level);
  if (!isnew) return ORC_OK;
  if (kind == 0) {
    b->local_new |= 1ull << who;
  } else {
    for (uint64_t i = 0; i < b->nrules_new; i++)
      if (b->rules_new[i] == rule) return ORC_OK;
    if (b->nrules_new < MAXT) b->rules_new[b->nrules_new++] = rule;
  }
  return ORC_OK;
}

int orc_explore(const orc_model *m, const orc_options *o, orc_result *res) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  memset(res, 0, sizeof *res);
  uint32_t W = m->words;
  vset set;
  vlist cur = {NULL, 0, 0, W}, nxt = {NULL, 0, 0, W};
  /* levels only for the cycle proviso; presized from o->expected_states */
  int rc = vset_init(&set, W, vset_cap_for(o->expected_states, 1u << 16), o->hash_seed,
                     o->por && o->cycle_proviso);
  if (rc) {
    vset_free(&set);
    return rc;
  }
  uint32_t s0[ORC_MAX_WORDS] = {0}; /* all processes in state 0 (cu:673-676) */
  vset_put(&set, s0, 0);
  vlist_push(&cur, s0);
  bfs_ctx b;
  memset(&b, 0, sizeof b);
  b.m = m;
  b.o = o;
  b.set = &set;
  b.next = &nxt;
  uint32_t level = 0;
  while (cur.n > 0) {
    b.level = level;
    for (uint64_t i = 0; i < cur.n; i++) {
      const uint32_t *s = cur.v + i * W;
      sview v;
      sinfo info;
      rc = view_state(m, s, &v);
      if (rc) goto out;
      if (!o->por) {
        /* safety on every generated successor (cu:926-935, 1063-1069) */
        b.check_safety = (o->property == ORC_PROP_SAFETY);
        rc = enumerate(m, s, &v, ~0ull, on_succ, &b, &info);
        if (rc) goto out;
      } else {
        /* full expansion pass: enabledness, clusters, safety */
        rc = enumerate(m, s, &v, ~0ull, on_full, &b, &info);
        if (rc) goto out;
        uint64_t prov, lnew = 0;
        if (!o->cycle_proviso) {
          prov = info.local_succ | info.in_enabled; /* cu:1395-1398, 1527-1530 */
        } else {
          sinfo tmp;
          b.local_new = 0;
          b.nrules_new = 0;
          rc = enumerate(m, s, &v, ~0ull, on_proviso, &b, &tmp);
          if (rc) goto out;
          prov = lnew = b.local_new;
          for (uint64_t r = 0; r < b.nrules_new; r++) prov |= b.rules_new[r];
        }
        uint64_t C;
        if (o->por_method == ORC_POR_AMPLE)
          C = select_ample(m->nr_procs, &info, o->cycle_proviso ? lnew : info.local_succ);
        else if (o->por_method == ORC_POR_STUBBORN)
          C = select_stubborn(m->nr_procs, &info, prov);
        else
          C = select_cluster(m->nr_procs, info.cluster0, prov);
        sinfo info2;
        rc = enumerate(m, s, &v, C, on_succ, &b, &info2);
        if (rc) goto out;
      }
      if (o->property == ORC_PROP_DEADLOCK && !(info.any_local_entry || info.any_enabled_rule))
        b.violation = 1; /* cu:1134-1151, 1679-1696 */
    }
    level++;
    res->levels = level;
    /* swap */
    vlist t = cur;
    cur = nxt;
    nxt = t;
    nxt.n = 0;
    if (b.violation) break; /* level-granular early stop */
    if (o->max_states && set.n >= o->max_states) {
      res->truncated = 1;
      break;
    }
  }
  rc = ORC_OK;
out:
  res->states = set.n;
  res->transitions = b.transitions;
  res->full_successors = o->por ? b.full : b.transitions;
  res->violation = b.violation;
  clock_gettime(CLOCK_MONOTONIC, &t1);
  res->seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  vset_free(&set);
  free(cur.v);
  free(nxt.v);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Multi-threaded plain BFS (the CPU baseline at several cores): the same
 * level-synchronous semantics as orc_explore without POR.  The visited set
 * is cut into shards, each an open-addressing vset behind a mutex; threads
 * take frontier chunks from an atomic counter and keep private next-level
 * lists, joined between levels.  Counts equal orc_explore's. */
#include <pthread.h>
#include <stdatomic.h>

#define MT_SHARDS 4096u

typedef struct {
  vset set;
  pthread_mutex_t mu;
} mt_shard;

typedef struct {
  const orc_model *m;
  const orc_options *o;
  mt_shard *shards;
  const uint32_t *cur;
  uint64_t ncur;
  _Atomic uint64_t next_chunk;
  uint32_t level;
} mt_shared;

typedef struct {
  mt_shared *sh;
  vlist next;
  uint64_t transitions;
  int violation, err;
} mt_worker;

static uint32_t mt_shard_of(const mt_shared *sh, const uint32_t *k) {
  uint64_t h = sh->o->hash_seed ^ 0x2545f4914f6cdd1dull;
  for (uint32_t i = 0; i < sh->m->words; i++) h = mix64(h ^ k[i]) + 0x9e37u * i;
  return (uint32_t)(h >> 52) & (MT_SHARDS - 1u);
}

static int mt_on_succ(void *c, const uint32_t *t, int kind, uint32_t who, uint64_t rule) {
  (void)kind;
  (void)who;
  (void)rule;
  mt_worker *w = (mt_worker *)c;
  mt_shared *sh = w->sh;
  w->transitions++;
  if (sh->o->property == ORC_PROP_SAFETY && is_violating(sh->m, t)) w->violation = 1;
  mt_shard *s = &sh->shards[mt_shard_of(sh, t)];
  pthread_mutex_lock(&s->mu);
  int rc = vset_put(&s->set, t, sh->level + 1);
  pthread_mutex_unlock(&s->mu);
  if (rc < 0) return rc;
  if (rc == 1) return vlist_push(&w->next, t);
  return ORC_OK;
}

static void *mt_run(void *arg) {
  mt_worker *w = (mt_worker *)arg;
  mt_shared *sh = w->sh;
  const uint32_t W = sh->m->words;
  for (;;) {
    uint64_t c0 = atomic_fetch_add(&sh->next_chunk, 1024u);
    if (c0 >= sh->ncur) break;
    uint64_t c1 = c0 + 1024u < sh->ncur ? c0 + 1024u : sh->ncur;
    for (uint64_t i = c0; i < c1; i++) {
      const uint32_t *s = sh->cur + i * W;
      sview v;
      sinfo info;
      int rc = view_state(sh->m, s, &v);
      if (!rc) rc = enumerate(sh->m, s, &v, ~0ull, mt_on_succ, w, &info);
      if (rc) {
        w->err = rc;
        return NULL;
      }
      if (sh->o->property == ORC_PROP_DEADLOCK && !(info.any_local_entry || info.any_enabled_rule))
        w->violation = 1; /* cu:1134-1151 */
    }
  }
  return NULL;
}

int orc_explore_mt(const orc_model *m, const orc_options *o, int nthreads, orc_result *res) {
  if (o->por || nthreads < 1 || nthreads > 256) return ORC_ERR_ARG;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  memset(res, 0, sizeof *res);
  const uint32_t W = m->words;
  int rc = ORC_OK;
  mt_shared sh;
  memset(&sh, 0, sizeof sh);
  sh.m = m;
  sh.o = o;
  sh.shards = (mt_shard *)calloc(MT_SHARDS, sizeof(mt_shard));
  mt_worker *ws = (mt_worker *)calloc((size_t)nthreads, sizeof(mt_worker));
  pthread_t *tids = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  vlist cur = {NULL, 0, 0, W};
  if (!sh.shards || !ws || !tids) {
    rc = ORC_ERR_NOMEM;
    goto out;
  }
  for (uint32_t i = 0; i < MT_SHARDS; i++) {
    /* level-free shards, presized from o->expected_states */
    rc = vset_init(&sh.shards[i].set, W, vset_cap_for(o->expected_states / MT_SHARDS, 64), o->hash_seed, 0);
    if (rc) goto out;
    pthread_mutex_init(&sh.shards[i].mu, NULL);
  }
  uint32_t s0[ORC_MAX_WORDS] = {0}; /* cu:673-676 */
  vset_put(&sh.shards[mt_shard_of(&sh, s0)].set, s0, 0);
  vlist_push(&cur, s0);
  for (int t = 0; t < nthreads; t++) {
    ws[t].sh = &sh;
    ws[t].next.W = W;
  }
  uint64_t states = 1, transitions = 0;
  int violation = 0;
  uint32_t level = 0;
  while (cur.n > 0) {
    sh.cur = cur.v;
    sh.ncur = cur.n;
    sh.level = level;
    atomic_store(&sh.next_chunk, 0);
    for (int t = 0; t < nthreads; t++) {
      ws[t].next.n = 0;
      if (pthread_create(&tids[t], NULL, mt_run, &ws[t])) {
        rc = ORC_ERR_NOMEM;
        for (int u = 0; u < t; u++) pthread_join(tids[u], NULL);
        goto out;
      }
    }
    uint64_t nnext = 0;
    for (int t = 0; t < nthreads; t++) {
      pthread_join(tids[t], NULL);
      if (ws[t].err) rc = ws[t].err;
      nnext += ws[t].next.n;
    }
    if (rc) goto out;
    /* next frontier: the threads' lists, concatenated */
    vlist nxt = {NULL, 0, 0, W};
    if (nnext) {
      nxt.v = (uint32_t *)malloc(nnext * W * sizeof(uint32_t));
      if (!nxt.v) {
        rc = ORC_ERR_NOMEM;
        goto out;
      }
      nxt.cap = nnext;
      for (int t = 0; t < nthreads; t++) {
        memcpy(nxt.v + nxt.n * W, ws[t].next.v, ws[t].next.n * W * sizeof(uint32_t));
        nxt.n += ws[t].next.n;
      }
    }
    transitions = 0;
    for (int t = 0; t < nthreads; t++) {
      transitions += ws[t].transitions;
      violation |= ws[t].violation;
    }
    states += nnext;
    free(cur.v);
    cur = nxt;
    level++;
    res->levels = level;
    if (o->progress) {
      fprintf(stderr, "oracle: level %u done, %llu states\n", level, (unsigned long long)states);
      fflush(stderr);
    }
    if (violation) break; /* level-granular early stop */
    if (o->max_states && states >= o->max_states) {
      res->truncated = 1;
      break;
    }
  }
  res->states = states;
  res->transitions = transitions;
  res->full_successors = transitions;
  res->violation = violation;
out:
  clock_gettime(CLOCK_MONOTONIC, &t1);
  res->seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  if (sh.shards) {
    for (uint32_t i = 0; i < MT_SHARDS; i++) {
      vset_free(&sh.shards[i].set);
      pthread_mutex_destroy(&sh.shards[i].mu);
    }
  }
  if (ws)
    for (int t = 0; t < nthreads; t++) free(ws[t].next.v);
  free(sh.shards);
  free(ws);
  free(tids);
  free(cur.v);
  return rc;
}
