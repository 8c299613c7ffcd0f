/*
 * matrix.h - drop-in for the reference's matrix.h (matrix.h:8-47) without
 * MPI: the MatDescriptor layout contract, DESC_SIZE, verifyDescriptor and
 * numEntriesPerProc.  Storage layout (matrix.h:20-44): nrec levels of
 * quadrants in the order 11, 21, 12, 22, then bs x bs blocks column-major
 * between and within blocks.  On one device nproc = nprocr = nprocc =
 * nproc_summa = 1.
 */
#ifndef MAT_DESCRIPTOR_H
#define MAT_DESCRIPTOR_H

#include <stdio.h>
#include <stdlib.h>

#include "../caps_amd.h"
#include "library.h"

struct MatDescriptor {
  int lda;
  int nrec;
  int nproc;
  int nprocr;
  int nprocc;
  int nproc_summa;
  int bs;
};

const int DESC_SIZE = 7;  // number of ints that make up a MatDescriptor
static_assert(sizeof(MatDescriptor) == DESC_SIZE * sizeof(int), "MatDescriptor is 7 ints");
static_assert(sizeof(MatDescriptor) == sizeof(caps_desc), "same layout as caps_desc");

/* The reference asserts (matrix.cpp:3-7); this prints the reason and aborts. */
inline void verifyDescriptor(MatDescriptor desc) {
  if (caps_verify_desc(reinterpret_cast<const caps_desc *>(&desc)) != CAPS_OK) {
    fprintf(stderr, "verifyDescriptor: %s\n", caps_last_error());
    abort();
  }
}

inline long long numEntriesPerProc(MatDescriptor desc) {
  long long lda = desc.lda;
  return lda * lda / desc.nproc / desc.nproc_summa;
}

#endif
