/*
 * gx_oracle_cli.c -- command-line front end of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Usage:
 *   gx_oracle_cli <model-without-.gpf> [-d|-p] [--por] [--cycle-proviso]
 *                 [--max-states N] [--seed S] [--threads T] [--expect N]
 * Prints one key=value line: states, transitions, levels, violation, seconds.
 * --threads T (plain method only) runs orc_explore_mt on T threads; --expect N
 * presizes the visited set for N states.
 * Unlike the reference CLI (cu:1851-1874) the options are orthogonal here.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gx_oracle.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <model> [-d|-p] [--por] [--por-method cample|ample|stubborn] [--cycle-proviso] [--max-states N] [--threads T] [--expect N]\n", argv[0]);
    return 1;
  }
  orc_options o;
  memset(&o, 0, sizeof o);
  int threads = 1;
  for (int i = 2; i < argc; i++) {
    if (!strcmp(argv[i], "-d")) o.property = ORC_PROP_DEADLOCK;
    else if (!strcmp(argv[i], "-p")) o.property = ORC_PROP_SAFETY;
    else if (!strcmp(argv[i], "--por")) o.por = 1;
    else if (!strcmp(argv[i], "--cycle-proviso")) o.cycle_proviso = 1;
    else if (!strcmp(argv[i], "--por-method") && i + 1 < argc) {
      const char *v = argv[++i];
      o.por = 1;
      if (!strcmp(v, "cample")) o.por_method = ORC_POR_CAMPLE;
      else if (!strcmp(v, "ample")) o.por_method = ORC_POR_AMPLE;
      else if (!strcmp(v, "stubborn")) o.por_method = ORC_POR_STUBBORN;
      else {
        fprintf(stderr, "unknown POR method %s\n", v);
        return 1;
      }
    }
    else if (!strcmp(argv[i], "--max-states") && i + 1 < argc) o.max_states = strtoull(argv[++i], NULL, 10);
    else if (!strcmp(argv[i], "--seed") && i + 1 < argc) o.hash_seed = strtoull(argv[++i], NULL, 10);
    else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--expect") && i + 1 < argc) o.expected_states = strtoull(argv[++i], NULL, 10);
    else {
      fprintf(stderr, "unknown option %s\n", argv[i]);
      return 1;
    }
  }
  orc_model *m = NULL;
  int rc = orc_load_gpf(argv[1], o.por, &m);
  if (rc) {
    fprintf(stderr, "load failed: %d\n", rc);
    return 1;
  }
  orc_result r;
  rc = threads > 1 ? orc_explore_mt(m, &o, threads, &r) : orc_explore(m, &o, &r);
  if (rc) {
    fprintf(stderr, "explore failed: %d\n", rc);
    orc_free(m);
    return 1;
  }
  printf("states=%llu transitions=%llu levels=%llu full=%llu violation=%d truncated=%d seconds=%.6f\n",
         (unsigned long long)r.states, (unsigned long long)r.transitions,
         (unsigned long long)r.levels, (unsigned long long)r.full_successors, r.violation,
         r.truncated, r.seconds);
  orc_free(m);
  return 0;
}
