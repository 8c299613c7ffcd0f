/*
 * multiply.h - drop-in for the reference's multiply entry point
 * (multiply.h:13): same name, argument order, by-value descriptor and
 * default pattern.  A, B, C are host arrays of numEntriesPerProc(desc)
 * doubles in the CAPS layout; A and B are read, C is overwritten.  The
 * product runs on the current HIP device (one MI355X): 'b' pattern steps
 * are batched levels, 'd' / 'h' steps recursive levels, NULL lets the
 * library's planner choose.  Like the reference (multiply.cpp:94-104), an
 * invalid pattern or descriptor prints a message and exits; callers that
 * want error codes use caps_multiply from caps_amd.h directly.
 *
 * The split-precision modes are reached with multiply_mode().
 */
#ifndef MULTIPLY_H
#define MULTIPLY_H

#include <stdio.h>
#include <stdlib.h>

#include "matrix.h"

inline void multiply_mode(double *A, double *B, double *C, MatDescriptor desc,
                          const char *pattern, int mode, int moduli = 0) {
  caps_opts o = {mode, moduli, 0, 0};
  int rc = caps_multiply(A, B, C, reinterpret_cast<const caps_desc *>(&desc), pattern, &o);
  if (rc != CAPS_OK) {
    printf("%s\n", caps_last_error());
    exit(-1);
  }
}

inline void multiply(double *A, double *B, double *C, MatDescriptor desc, char *pattern = NULL) {
  multiply_mode(A, B, C, desc, pattern, CAPS_MODE_FP64);
}

/* distributeFrom1Proc / collectTo1Proc for one process
 * (communication.cpp:657-659, 706-708): column-major <-> CAPS layout. */
inline void distributeFrom1Proc(MatDescriptor desc, double *O, double *I) {
  if (caps_to_layout(I, O, reinterpret_cast<const caps_desc *>(&desc)) != CAPS_OK) {
    printf("%s\n", caps_last_error());
    exit(-1);
  }
}

inline void collectTo1Proc(MatDescriptor desc, double *O, double *I) {
  if (caps_from_layout(I, O, reinterpret_cast<const caps_desc *>(&desc)) != CAPS_OK) {
    printf("%s\n", caps_last_error());
    exit(-1);
  }
}

#endif
