// Device helpers of the split-precision modes: deterministic norm bounds,
// power-of-two scale exponents, symmetric residues and the CRT.
#pragma once
#include <hip/hip_runtime.h>

#include "ozaki.h"

namespace caps {

constexpr int NOSCALE = -100000;  // exponent of an all-zero row / column

// running (exponent e, sum s) with sum of squares = s * 4^e
__device__ __forceinline__ void sq_acc(double a, int &e, double &s) {
  if (a == 0.0) return;
  int ea = ilogb(a);
  if (ea > e) {
    s = (e == NOSCALE) ? 0.0 : ldexp(s, 2 * (e - ea));
    e = ea;
  }
  double t = ldexp(a, -e);
  s = fma(t, t, s);
}

// (exponent, sum of squares) of 8 values: scaled by their largest exponent,
// so the 8 squares add without rescaling steps
__device__ __forceinline__ void sq_block8(const double (&v)[8], int &e, double &s) {
  int E = NOSCALE;
#pragma unroll
  for (int k = 0; k < 8; k++) E = v[k] != 0.0 ? max(E, ilogb(v[k])) : E;
  e = E;
  s = 0.0;
  if (E == NOSCALE) return;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const double t = ldexp(v[k], -E);
    s = fma(t, t, s);
  }
}

__device__ __forceinline__ void sq_merge(int &e, double &s, int e2, double s2) {
  if (e2 == NOSCALE) return;
  if (e == NOSCALE) {
    e = e2;
    s = s2;
    return;
  }
  if (e2 > e) {
    s = ldexp(s, 2 * (e - e2)) + s2;
    e = e2;
  } else {
    s = s + ldexp(s2, 2 * (e2 - e));
  }
}

// largest integer t with 2^t * 2^e * sqrt(s) * (1 + 2^-20) <= L
__device__ __forceinline__ int scale_exponent(int e, double s, double L) {
  if (e == NOSCALE || s == 0.0) return 0;
  const double nm = sqrt(s) * (1.0 + 0x1p-20);  // rounding of s covered for K < 2^30
  int t = ilogb(L) - e - ilogb(nm);
  if (ldexp(nm, e + t) > L) t -= 1;
  return t;
}

// residue of a small integer (|x| <= 8192) modulo mi (2 <= mi <= 256) in
// [-mi/2, mi/2], in integer arithmetic: with c22 = round(2^22 / mi),
// q = (x c22 + 2^21) >> 22 rounds x c22 / 2^22, which is within 8192 * 2^-23
// = 2^-10 of x / mi; that never crosses a half-integer, since x / mi is at
// least 1 / (2 mi) >= 2^-9 away from one for odd mi (for mi = 256, c22 is
// exact and a tie rounds up: -128).  So q = round(x / mi) and r = x - q mi.
// Two mad24 and a shift (|x c22| < 2^28).
__device__ __forceinline__ int mod_k(int x, int c22, int mi) {
  const int q = (__mul24(x, c22) + (1 << 21)) >> 22;
  return __mul24(q, -mi) + x;
}

// the low bytes of four ints as one word, byte h from b_h.  v_perm_b32
// (bytes 0-3 of the second operand are selectors 0-3, of the first 4-7, 0x0c
// gives 0): written this way the compiler keeps the byte shifts out of the
// mad24 arithmetic that produced the values (folding them turns mad24 into
// 32-bit multiplies).
__device__ __forceinline__ unsigned pack4_lo(int b0, int b1, int b2, int b3) {
  const unsigned lo = __builtin_amdgcn_perm((unsigned)b1, (unsigned)b0, 0x0c0c0400u);
  const unsigned hi = __builtin_amdgcn_perm((unsigned)b3, (unsigned)b2, 0x04000c0cu);
  return lo | hi;
}

// symmetric residue of the FP64 integer x (|x| <= 2^63) modulo mi: the FP64
// quotient estimate x * inv is within |x / mi| 2^-52 <= 11 of x / mi, so r =
// x - q mi is an integer with |r| < 12 mi <= 3072, exact in one fma; mod_k
// finishes
__device__ __forceinline__ int sym_residue(double x, int mi, double inv, int c22) {
  const double r = fma(-rint(x * inv), (double)mi, x);
  return mod_k((int)r, c22, mi);
}

// residue of a small integer (|x| < 2^22) modulo mi in [-mi/2, mi/2]: the
// FP32 quotient estimate x * inv is within |x/mi| 2^-23 < 1 / (2 mi) of x / mi,
// which is never a half-integer for odd mi (2x = odd * mi is impossible), so
// rint finds the nearest integer; for mi = 256 (exact inverse) a tie gives
// +-128, the same INT8 value after the wrap.  Every step is exact: integers
// below 2^24 in FP32.
__device__ __forceinline__ int mod_small(int x, float mf, float inv) {
  const float xf = (float)x;
  return (int)fmaf(-rintf(xf * inv), mf, xf);
}

// byte q of a packed INT8 residue word from an INT32 product sum that started
// at 128: xb = x + 128, |x| <= 2^28 (K <= 2^14 products of INT8 residues).
//   y = xb + (xb >> 14) (c14 - 2^14) = (xb >> 14) c14 + (xb & 16383) = xb
//   (mod m), |y| < 2^21.1, in one mad24 (c14m = c14 - 2^14);
//   k = floor(y / m + 1/2 - 128 / m) from fma(y, 1/m, cl) in FP32, cl = 1/2 -
//   128/m: the error (< 1.2e-3) is below the distance 1/(2m) of that argument
//   to an integer for odd m, and every step is exact for m = 256;
//   y - k m = r + 128, r = x mod m in [-m/2, m/2): a byte value in [0, 255]
//   that cvt_pk_u8 converts and places at byte q of `word`.  XOR 0x80 per
//   byte then gives the INT8 residue r.
__device__ __forceinline__ unsigned res_byte128(int xb, int c14m, float mf, float inv, float cl,
                                                int q, unsigned word) {
  const int y = __mul24(xb >> 14, c14m) + xb;
  const float yf = (float)y;
  const float k = floorf(fmaf(yf, inv, cl));
  return __builtin_amdgcn_cvt_pk_u8_f32(fmaf(-k, mf, yf), q, word);
}

// symmetric residue of an INT32 product sum (|x| < 2^31) modulo mi
__device__ __forceinline__ int mod_i32(int x, int mi, float inv) {
  int q = __float2int_rn((float)x * inv);
  int r = x - q * mi;
  const int hi = (mi - 1) >> 1, lo = -(mi >> 1);
  r = r > hi ? r - mi : r;
  r = r < lo ? r + mi : r;
  return r;
}

// S_p += r_l * W_l[p] with pieces W_l[p] < 2^37: exact for |r_l| <= 2^11 (a
// residue, or a sum congruent to it), every partial sum below 14 * 2^48 < 2^52
template <int NP>
__device__ __forceinline__ void crt_acc(double (&S)[NP], int r, int l,
                                        const OzakiConsts *__restrict__ kc) {
  const double rd = (double)r;
#pragma unroll
  for (int p = 0; p < NP; p++) S[p] = fma(rd, kc->w[l][p], S[p]);
}

// correctly rounded conversion of a signed 128-bit integer to FP64
__device__ __forceinline__ double i128_to_double(__int128 x) {
  const bool neg = x < 0;
  const unsigned __int128 a = neg ? (unsigned __int128)0 - (unsigned __int128)x : (unsigned __int128)x;
  const unsigned long long hi = (unsigned long long)(a >> 64), lo = (unsigned long long)a;
  const int bits = hi ? 128 - __clzll(hi) : (lo ? 64 - __clzll(lo) : 0);
  double r;
  if (bits <= 53) {
    r = (double)lo;  // exact
  } else {
    const int sh = bits - 53;
    unsigned __int128 q = a >> sh;
    const unsigned __int128 rem = a - (q << sh);
    const unsigned __int128 half = (unsigned __int128)1 << (sh - 1);
    if (rem > half || (rem == half && (q & 1))) q += 1;  // round half to even
    r = ldexp((double)(unsigned long long)q, sh);      // q <= 2^53: exact
  }
  return neg ? -r : r;
}

// s + e == a + b exactly (Knuth's TwoSum; no ordering of a and b needed)
__device__ __forceinline__ double two_sum(double a, double b, double &e) {
  const double s = a + b;
  const double bv = s - a;
  e = (a - (s - bv)) + (b - bv);
  return s;
}

// a + b rounded to odd: the sum itself when it is a double, else the one of
// its two neighbouring doubles whose last significand bit is 1 (a + b != 0
// whenever the rounded sum is inexact, so the step below never starts at 0)
__device__ __forceinline__ double add_round_odd(double a, double b) {
  double e;
  const double s = two_sum(a, b, e);
  long long bits = __double_as_longlong(s);
  if (e != 0.0 && !(bits & 1)) bits += ((e > 0.0) == (s > 0.0)) ? 1 : -1;
  return __longlong_as_double(bits);
}

// X = sum_p S_p 2^(37p) reduced to (-M/2, M/2): the quotient k = round(S/M)
// from an FP64 estimate (|S| < 2^14.8 M, so the estimate of S/M is within
// 2^-36 of it: exact unless X were within 2^-36 M of M/2, which the scaling
// bound, |X| <= (1/2 - 2^-31) M, excludes), D_p = S_p - k M_p exact (|k| <
// 2^15, |D_p| < 2^53), and X = sum_p D_p 2^(37p) rounded once to FP64:
//  * two pieces: D_1 2^37 and D_0 are doubles, one rounded addition;
//  * three pieces: the correctly rounded sum of the three doubles a = D_2
//    2^74, b = D_1 2^37, c = D_0 by rounding to odd (Boldo and Melquiond,
//    "Emulation of FMA and correctly rounded sums: proved algorithms using
//    rounding to odd", IEEE Trans. Computers 57(4), 2008): (uh, ul) =
//    TwoSum(b, c), (th, tl) = TwoSum(a, uh), v = RO(tl + ul), X = RN(th + v);
//  * four pieces: X assembled exactly in 128 bits (|X| < M/2 <= 2^127).
template <int NP>
__device__ __forceinline__ double crt_finish(const double (&S)[NP],
                                             const OzakiConsts *__restrict__ kc) {
  double approx = S[NP - 1];
#pragma unroll
  for (int p = NP - 2; p >= 0; p--) approx = fma(approx, 0x1p37, S[p]);
  const double k = rint(approx * kc->invM);
  double D[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) D[p] = fma(-k, kc->Mp[p], S[p]);
  if constexpr (NP == 1) {
    return D[0];
  } else if constexpr (NP == 2) {
    return fma(D[1], 0x1p37, D[0]);  // D_1 2^37 exact: one rounding
  } else if constexpr (NP == 3) {
    double ul, tl;
    const double uh = two_sum(D[1] * 0x1p37, D[0], ul);
    const double th = two_sum(D[2] * 0x1p74, uh, tl);
    return th + add_round_odd(tl, ul);
  } else {
    __int128 X = 0;
#pragma unroll
    for (int p = NP - 1; p >= 0; p--) X = X * ((__int128)1 << OZ_PIECE_BITS) + (__int128)(long long)D[p];
    return i128_to_double(X);
  }
}

// offset of element (i, j) of an n x n matrix in the CAPS layout (nproc = 1):
// nrec quadrant levels in the order 11, 21, 12, 22, then bs x bs blocks
// column-major between and within blocks (matrix.h:20-44)
__device__ __forceinline__ long long caps_elem_offset(int n, int nrec, int bs, int i, int j) {
  long long off = 0;
  long long s = n;
  for (int l = 0; l < nrec; l++) {
    long long h = s >> 1;
    int bi = i >= h, bj = j >= h;
    off += (long long)(bi + 2 * bj) * h * h;
    i -= bi * (int)h;
    j -= bj * (int)h;
    s = h;
  }
  long long nb = s / bs;
  return off + ((long long)(i / bs) + (long long)(j / bs) * nb) * bs * bs + (i % bs) +
         (long long)(j % bs) * bs;
}

// offset of (li, lj) inside an m x m leaf stored as bs x bs blocks
__device__ __forceinline__ long long leaf_elem_offset(int m, int bs, int li, int lj) {
  const long long nb = m / bs;
  return ((long long)(li / bs) + (long long)(lj / bs) * nb) * bs * bs + (li % bs) +
         (long long)(lj % bs) * bs;
}

// leaf index L (CAPS order: base-4 digits, quadrant 11, 21, 12, 22 per level)
// -> leaf row block bi and leaf column block bj
__device__ __forceinline__ void leaf_coords(int L, int nrec, int &bi, int &bj) {
  bi = 0;
  bj = 0;
  for (int l = 0; l < nrec; l++) {
    const int q = (L >> (2 * (nrec - 1 - l))) & 3;
    bi = 2 * bi + (q & 1);
    bj = 2 * bj + (q >> 1);
  }
}

}  // namespace caps
