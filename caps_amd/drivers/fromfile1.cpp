// fromfile1 for one MI355X through the drop-in multiply.h (fromfile1.cpp:11-164):
// reads n, A and B as row-major text into column-major storage, converts to
// the CAPS layout (distributeFrom1Proc), multiplies, converts back
// (collectTo1Proc) and prints "Maximum deviation" against a -c check file or
// the product to -o / stdout.  Flags: -i -o -c -b -r -p, plus --mode.
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
  char *inFile = read_string(argc, argv, "-i", NULL);
  char *outFile = read_string(argc, argv, "-o", NULL);
  char *checkFile = read_string(argc, argv, "-c", NULL);
  int bs = read_int(argc, argv, "-b", -1);
  int nrec = read_int(argc, argv, "-r", -1);
  char *pattern = read_string(argc, argv, "-p", NULL);
  char *mode_s = read_string(argc, argv, "--mode", (char *)"fp64");
  if (!inFile) {
    printf("Specify input file\n");
    return -1;
  }
  if (bs == -1 || nrec == -1) {
    printf("Specify parameters with -b and -r.\n");
    return -1;
  }
  FILE *f = fopen(inFile, "r");
  if (!f) {
    printf("Error opening file: %s\n", inFile);
    return -1;
  }
  MatDescriptor desc = {0, nrec, 1, 1, 1, 1, bs};
  if (fscanf(f, "%d", &desc.lda) != 1) return -1;
  const int n = desc.lda;
  std::vector<double> IA((size_t)n * n), IB((size_t)n * n);
  int reads = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) reads += fscanf(f, "%lf", &IA[i + (size_t)j * n]);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) reads += fscanf(f, "%lf", &IB[i + (size_t)j * n]);
  fclose(f);
  if (reads != 2 * n * n) {
    printf("Error reading file: %s, only read %d entries\n", inFile, reads);
    return -1;
  }
  verifyDescriptor(desc);
  int mode = CAPS_MODE_FP64;
  if (!strcmp(mode_s, "crt")) mode = CAPS_MODE_CRT_INT8;
  if (!strcmp(mode_s, "leaf")) mode = CAPS_MODE_OZAKI_INT8;
  std::vector<double> A((size_t)n * n), B((size_t)n * n), C((size_t)n * n), OC((size_t)n * n);
  distributeFrom1Proc(desc, A.data(), IA.data());
  distributeFrom1Proc(desc, B.data(), IB.data());
  double t0 = read_timer();
  multiply_mode(A.data(), B.data(), C.data(), desc, pattern, mode);
  double t = read_timer() - t0;
  collectTo1Proc(desc, OC.data(), C.data());
  if (checkFile) {
    FILE *cf = fopen(checkFile, "r");
    if (!cf) {
      printf("Error opening verification file\n");
      return -1;
    }
    double err = 0, cc;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        if (fscanf(cf, "%lf", &cc) != 1) {
          printf("Error reading verification file\n");
          return -1;
        }
        err = fmax(err, fabs(OC[i + (size_t)j * n] - cc));
      }
    fclose(cf);
    printf("Maximum deviation %e\n", err);
  } else {
    FILE *o = outFile ? fopen(outFile, "w") : stdout;
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < n; j++) fprintf(o, "%.18le ", OC[i + (size_t)j * n]);
      fprintf(o, "\n");
    }
    if (outFile) fclose(o);
  }
  printf("time %f\n", t);
  return 0;
}
