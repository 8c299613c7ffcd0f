// Per-element Winograd add steps, shared by the FP64 kernels (V = double or
// a pair of doubles) and the residue kernels (V = int, reduced mod m after).
//
// Each functor evaluates exactly the expression tree of the reference's add
// routines (multiply.cpp:110-210, 272, 277, 409-451); in FP64 that fixes the
// rounding, in residue arithmetic it is exact in any order.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace caps {

template <int OP>
struct Ew;

template <>
struct Ew<EW_BFS_PRE> {
  static constexpr int NIN = 8, NOUT = 8;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    const V &A11 = x[0], &A21 = x[1], &A12 = x[2], &A22 = x[3];
    const V &B11 = x[4], &B21 = x[5], &B12 = x[6], &B22 = x[7];
    V T3 = A21 + A22;
    V T4 = T3 - A11;
    V T5 = A11 - A21;
    V T6 = A12 - T4;
    V S5 = B22 - B12;
    V S3 = B12 - B11;
    V S4 = B22 - S3;
    V S7 = S4 - B21;
    y[0] = T3; y[1] = T4; y[2] = T5; y[3] = T6;
    y[4] = S3; y[5] = S4; y[6] = S5; y[7] = S7;
  }
};

template <>
struct Ew<EW_BFS_POST> {
  static constexpr int NIN = 7, NOUT = 4;
  template <typename V>
  __device__ static void apply(const V *q, V *c) {
    V C11 = q[0] + q[1];
    V U1 = q[0] + q[3];
    V U2 = U1 + q[4];
    V C21 = U2 - q[6];
    V C22 = U2 + q[2];
    V U3 = U1 + q[2];
    V C12 = U3 + q[5];
    c[0] = C11; c[1] = C21; c[2] = C12; c[3] = C22;
  }
};

// tripleSubMatrices(S5=B22-B12, S3=B12-B11, S4=B22-S3) then
// addSubMatrices(T3=A21+A22, T5=A11-A21)            (multiply.cpp:421, 424)
template <>
struct Ew<EW_DFS_PRE1> {
  static constexpr int NIN = 6, NOUT = 5;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    const V &A11 = x[0], &A21 = x[1], &A22 = x[2];
    const V &B11 = x[3], &B12 = x[4], &B22 = x[5];
    V S5 = B22 - B12;
    V S3 = B12 - B11;
    V S4 = B22 - S3;
    V T3 = A21 + A22;
    V T5 = A11 - A21;
    y[0] = S5; y[1] = S3; y[2] = S4; y[3] = T5; y[4] = T3;
  }
};

// subMatrices(T4 = T3 - A11)                          (multiply.cpp:430)
template <>
struct Ew<EW_DFS_PRE2> {
  static constexpr int NIN = 2, NOUT = 1;
  template <typename V>
  __device__ static void apply(const V *x, V *y) { y[0] = x[0] - x[1]; }
};

// subMatrices(T6 = A12 - T4); subMatrices(S7 = S4 - B21)   (434, 436)
template <>
struct Ew<EW_DFS_PRE3> {
  static constexpr int NIN = 4, NOUT = 2;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    y[0] = x[0] - x[1];
    y[1] = x[2] - x[3];
  }
};

// tripleAddMatrices(U1=Q1+Q4, U2=U1+Q5, U3=U1+Q3) then
// addSubMatrices(C22=U2+Q3, C21=U2-Q7)                     (444, 445)
template <>
struct Ew<EW_DFS_POST1> {
  static constexpr int NIN = 5, NOUT = 3;
  template <typename V>
  __device__ static void apply(const V *x, V *y) {
    const V &Q1 = x[0], &Q4 = x[1], &Q5 = x[2], &Q3 = x[3], &Q7 = x[4];
    V U1 = Q1 + Q4;
    V U2 = U1 + Q5;
    V U3 = U1 + Q3;
    y[0] = U3;
    y[1] = U2 + Q3;
    y[2] = U2 - Q7;
  }
};

// addMatrices(C11 = Q1 + Q2)                              (448)
template <>
struct Ew<EW_DFS_POST2> {
  static constexpr int NIN = 2, NOUT = 1;
  template <typename V>
  __device__ static void apply(const V *x, V *y) { y[0] = x[0] + x[1]; }
};

// addMatrices(C12 = U3 + Q6)                              (451)
template <>
struct Ew<EW_DFS_POST3> {
  static constexpr int NIN = 2, NOUT = 1;
  template <typename V>
  __device__ static void apply(const V *x, V *y) { y[0] = x[0] + x[1]; }
};


}  // namespace caps
