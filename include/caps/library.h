/*
 * library.h - drop-in for the reference's library.h (library.h:1-8):
 * SEVEN and read_timer().  The reference defines a max(a,b) macro before its
 * include guard (library.h:1); it is available here only on request
 * (CAPS_LIBRARY_MAX_MACRO) because the macro breaks std::max.
 */
#ifdef CAPS_LIBRARY_MAX_MACRO
#ifndef max
#define max(a, b) (((a) > (b)) ? (a) : (b))
#endif
#endif

#ifndef LIBRARY_H
#define LIBRARY_H

#include <sys/time.h>

const int SEVEN = 7;

/* seconds since the first call (library.cpp:5-19) */
inline double read_timer() {
  static bool initialized = false;
  static struct timeval start;
  struct timeval end;
  if (!initialized) {
    gettimeofday(&start, 0);
    initialized = true;
  }
  gettimeofday(&end, 0);
  return (end.tv_sec - start.tv_sec) + 1.0e-6 * (end.tv_usec - start.tv_usec);
}

#endif
