// randombenchmark for one MI355X through the drop-in multiply.h.
//
// Same flags and report as the reference driver (randombenchmark.cpp:17-99,
// counters.cpp:61-74): -s size, -b block size, -r recursion levels, -m / -k
// memory budget (MB / KB), -p pattern; inputs 2*drand48()-1 in storage
// order, A then B; default nrec = floor(log2(s / 512)) (MIN_STRASSEN = 512).
// Extra flags: --mode fp64|crt|leaf, --moduli N, --warmup N.
// The timed call is the host-pointer multiply(), so it includes the copies
// to and from the device; bench.py measures device-resident throughput.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "caps/multiply.h"

static int find_option(int argc, char **argv, const char *option) {
  for (int i = 1; i < argc; i++)
    if (strcmp(argv[i], option) == 0) return i;
  return -1;
}
static int read_int(int argc, char **argv, const char *option, int def) {
  int i = find_option(argc, argv, option);
  return (i >= 0 && i < argc - 1) ? atoi(argv[i + 1]) : def;
}
static char *read_string(int argc, char **argv, const char *option, char *def) {
  int i = find_option(argc, argv, option);
  return (i >= 0 && i < argc - 1) ? argv[i + 1] : def;
}

int main(int argc, char **argv) {
  const int MIN_STRASSEN = 512;
  int size = read_int(argc, argv, "-s", -1);
  if (size == -1) {
    printf("Must specify the matrix size desired with -s\n");
    return -1;
  }
  int bs = read_int(argc, argv, "-b", 1);
  int nrec = read_int(argc, argv, "-r", -1);
  int mem = read_int(argc, argv, "-m", 1536);
  char *pattern = read_string(argc, argv, "-p", NULL);
  mem = read_int(argc, argv, "-k", mem * 1024);
  char *mode_s = read_string(argc, argv, "--mode", (char *)"fp64");
  int moduli = read_int(argc, argv, "--moduli", 0);
  int warmup = read_int(argc, argv, "--warmup", 1);
  caps_set_available_memory((long long)mem * 128);  // KB -> doubles
  if (nrec == -1) {
    int rsize = size / MIN_STRASSEN;
    nrec = 0;
    while (rsize > 1) {
      nrec += 1;
      rsize /= 2;
    }
    printf("Setting nrec=%d\n", nrec);
  }
  int mode = CAPS_MODE_FP64;
  if (!strcmp(mode_s, "crt")) mode = CAPS_MODE_CRT_INT8;
  if (!strcmp(mode_s, "leaf")) mode = CAPS_MODE_OZAKI_INT8;
  printf("Benchmarking size %d\n", size);
  MatDescriptor desc = {size, nrec, 1, 1, 1, 1, bs};
  verifyDescriptor(desc);
  long long N = numEntriesPerProc(desc);
  std::vector<double> A(N), B(N), C(N);
  for (long long i = 0; i < N; i++) A[i] = 2 * drand48() - 1;
  for (long long i = 0; i < N; i++) B[i] = 2 * drand48() - 1;
  for (int w = 0; w < warmup; w++) multiply_mode(A.data(), B.data(), C.data(), desc, pattern, mode, moduli);
  caps_reset_counters();
  double t0 = read_timer();
  multiply_mode(A.data(), B.data(), C.data(), desc, pattern, mode, moduli);
  double t = read_timer() - t0;
  caps_counters c;
  caps_get_counters(&c);
  long long flops = c.additions + c.multiplications;
  printf("peak memory usage %.2f MB (device workspace)\n", c.peak_workspace_mb);
  printf("Execution pattern %s\n", c.pattern);
  printf("flops=%lld (%lld+%lld)\n", flops, c.additions - c.multiplications, 2 * c.multiplications);
  printf("time %f actual Gflop/s %f effective Gflop/s %f\n", t, flops / 1.e9 / t,
         2. * size * (double)size * size / 1.e9 / t);
  return 0;
}
