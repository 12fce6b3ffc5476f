// gx_cli.cpp -- the `GPUexplore <model> [OPTIONS]` command (cu:1738-2188).
//
// Flag syntax, flag coupling, messages and exit codes follow the
// reference's main(): -d sets the property to deadlock and switches the
// cycle proviso off, -p sets it to safety and switches it on, and
// --cycle-proviso only counts while no property was given (cu:1851-1874).
// Exploration itself goes through the C ABI (libgpuexplore.so).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>

#include "../../include/gpuexplore.h"

static const char *kHelp =
    "Usage: GPUexplore <model> [OPTIONS]\n"
    "Run state-space exploration on model (do not include the file extension).\n"
    "options:\n"
    "  -d                 Check for deadlocks\n"
    "  -p                 Check a safety property (should be embedded in the model)\n"
    "  --por              Apply partial-order reduction\n"
    "  --cycle-proviso    Apply the cycle proviso during partial-order reduction\n"
    "  --por-method NAME  Partial-order reduction by cample (default), ample or stubborn sets;\n"
    "                     implies --por\n"
    "  -k NUM             Run NUM BFS levels per host synchronisation (default 1)\n"
    "  -b NUM             Run the kernels on NUM workgroups (default: fill the GPU)\n"
    "  -t NUM             Use NUM threads per workgroup, a multiple of 64 (default 256)\n"
    "  -q NUM             Allocate NUM integers for the hash table\n"
    "  --bucket BYTES     Hash bucket width: 64, 128 or 256 (default 64 up to 126-bit states)\n"
    "  --seed NUM         Hash seed (state counts do not depend on it)\n"
    "  --dump FILE        Dump the state space to FILE after completing the exploration\n"
    "  -v NUM             Change the verbosity:\n"
    "                        0 - minimal output\n"
    "                        1 - print sequence number of each kernel launch\n"
    "                        2 - print number of states in the hash table after each kernel launch\n"
    "                        3 - print state vectors after each kernel launch\n"
    "  -h, --help         Show this help message\n";

static bool is_help(const char *a) { return !strcmp(a, "--help") || !strcmp(a, "-h") || !strcmp(a, "-?"); }

int main(int argc, char **argv) {
  const auto start = std::chrono::steady_clock::now();
  if (argc == 1) {  // cu:1805-1808
    fprintf(stderr, "ERROR: No input network given!\n");
    fputs(kHelp, stdout);
    return 1;
  }
  if (is_help(argv[1])) {  // cu:1809-1812
    fputs(kHelp, stdout);
    return 0;
  }
  gx_options o;
  gx_options_default(&o);
  int property = GX_PROP_NONE, por = 0, proviso = 0, verbosity = 0;
  const char *dump_file = nullptr;
  for (int i = 2; i < argc;) {  // cu:1817-1883
    const char *a = argv[i];
    const bool has_val = i + 1 < argc;
    if (is_help(a)) {
      fputs(kHelp, stdout);
      return 0;
    } else if (!strcmp(a, "-k") && has_val) {
      o.kernel_iters = static_cast<uint32_t>(atoi(argv[i + 1]));
      i += 2;
    } else if (!strcmp(a, "-b") && has_val) {
      o.blocks = static_cast<uint32_t>(atoi(argv[i + 1]));
      i += 2;
    } else if (!strcmp(a, "-t") && has_val) {
      o.threads = static_cast<uint32_t>(atoi(argv[i + 1]));
      i += 2;
    } else if (!strcmp(a, "-q") && has_val) {
      o.q_ints = static_cast<uint64_t>(atoll(argv[i + 1]));
      i += 2;
    } else if (!strcmp(a, "-v") && has_val) {
      verbosity = atoi(argv[i + 1]);
      if (verbosity > 3) verbosity = 3;
      i += 2;
    } else if (!strcmp(a, "-d")) {
      property = GX_PROP_DEADLOCK;
      proviso = 0;
      i += 1;
    } else if (!strcmp(a, "-p")) {
      property = GX_PROP_SAFETY;
      proviso = 1;
      i += 1;
    } else if (!strcmp(a, "--por")) {
      por = 1;
      i += 1;
    } else if (!strcmp(a, "--cycle-proviso")) {
      if (property == GX_PROP_NONE) proviso = 1;
      i += 1;
    } else if (!strcmp(a, "--por-method") && has_val) {
      // one binary for the reference's per-method branches (README.md lines 24-26)
      const char *v = argv[i + 1];
      if (!strcmp(v, "cample")) {
        o.por_method = GX_POR_CAMPLE;
      } else if (!strcmp(v, "ample")) {
        o.por_method = GX_POR_AMPLE;
      } else if (!strcmp(v, "stubborn")) {
        o.por_method = GX_POR_STUBBORN;
      } else {
        fprintf(stderr, "ERROR: unknown partial-order reduction method %s\n", v);
        fputs(kHelp, stdout);
        return 1;
      }
      por = 1;
      i += 2;
    } else if (!strcmp(a, "--dump") && has_val) {
      dump_file = argv[i + 1];
      i += 2;
    } else if (!strcmp(a, "--bucket") && has_val) {
      o.bucket_bytes = static_cast<uint32_t>(atoi(argv[i + 1]));
      i += 2;
    } else if (!strcmp(a, "--seed") && has_val) {
      o.hash_seed = static_cast<uint64_t>(strtoull(argv[i + 1], nullptr, 10));
      i += 2;
    } else {
      fprintf(stderr, "ERROR: unrecognized option %s\n", a);
      fputs(kHelp, stdout);
      return 1;
    }
  }
  gx_model *m = nullptr;
  int rc = gx_model_load_gpf(argv[1], por, &m);
  if (rc == GX_E_NOFILE) {  // cu:1984-1986
    fprintf(stderr, "ERROR: input network does not exist!\n");
    return 1;
  }
  if (rc != GX_OK) {
    fprintf(stderr, "ERROR: %s\n", gx_strerror(rc));
    return 1;
  }
  gx_model_info info;
  gx_model_info_get(m, &info);
  if (property == GX_PROP_SAFETY) {  // cu:1888-1898
    fprintf(stdout, "Property to check is %sa liveness property\n", info.liveness == 0 ? "not " : "");
    if (info.liveness == 1) property = GX_PROP_LIVENESS;
  }
  fprintf(stdout, "nr of procs: %u\n", info.nr_procs);
  fprintf(stdout, "nr of bits for transition label: %u\n", info.bits_act);
  fprintf(stdout, "min. nr. of proc. states that fit in 32-bit integer: %u\n", info.proc_nrstates);
  fprintf(stdout, "number of bits needed for a state vector: %u\n", info.bits_statevector);
  fprintf(stdout, "maximum label-bounded branching factor: %u\n", info.max_buf_ints);
  fprintf(stdout, "total number of transition entries in network: %u\n", info.nr_trans);

  o.property = property;
  o.por = por;
  o.cycle_proviso = proviso;
  o.verbosity = verbosity;
  gx_result r;
  memset(&r, 0, sizeof r);
  gx_explorer *ex = nullptr;
  rc = gx_explorer_create(m, &o, &ex);
  if (rc == GX_OK) rc = gx_explorer_run(ex, &r);
  if (rc == GX_E_NODEVICE) {  // cu:2000-2003
    fprintf(stderr, "ERROR: No HIP compatible GPU detected!\n");
    gx_explorer_free(ex);
    gx_model_free(m);
    return 1;
  }
  if (rc != GX_OK && rc != GX_E_CAPACITY) {
    fprintf(stderr, "ERROR: %s\n", gx_strerror(rc));
    gx_explorer_free(ex);
    gx_model_free(m);
    return 1;
  }
  const auto stop = std::chrono::steady_clock::now();
  fprintf(stdout, "global mem queue size: %llu, number of entries: %llu\n",
          static_cast<unsigned long long>(r.table_bytes), static_cast<unsigned long long>(r.table_bytes / 4));
  fprintf(stdout, "Run time: %f\n", std::chrono::duration<double>(stop - start).count());  // cu:2146
  fprintf(stdout, "Exploration time %f\n", r.t_explore_s);                              // cu:2148
  if (r.violation) {  // cu:2150-2162
    if (property == GX_PROP_DEADLOCK) printf("deadlock detected!\n");
    else if (property == GX_PROP_SAFETY) printf("safety property violation detected!\n");
    else if (property == GX_PROP_LIVENESS) printf("liveness property violation detected!\n");
  }
  if (r.table_full || rc == GX_E_CAPACITY) fprintf(stderr, "ERROR: problem with hash table\n");  // cu:2164-2166
  fprintf(stdout, "nr. of states in hash table: %llu\n", static_cast<unsigned long long>(r.states));  // cu:2172
  fprintf(stdout, "nr. of transitions: %llu\n", static_cast<unsigned long long>(r.transitions));
  fprintf(stdout, "nr. of BFS levels: %llu\n", static_cast<unsigned long long>(r.levels));
  if (r.t_explore_s > 0)
    fprintf(stdout, "new states/s: %.0f\n", static_cast<double>(r.inserts) / r.t_explore_s);
  if (dump_file) {  // cu:2174-2185
    FILE *t = fopen(dump_file, "w");
    if (t) {
      fclose(t);
      fprintf(stdout, "Dumping state space to file...\n");
      fflush(stdout);
      const int drc = gx_explorer_dump(ex, dump_file);
      if (drc != GX_OK) fprintf(stderr, "ERROR: %s\n", gx_strerror(drc));
    } else {
      fprintf(stderr, "Could not open file to dump the state space\n");
    }
  }
  gx_explorer_free(ex);
  gx_model_free(m);
  return 0;  // cu:2187
}
