// Host side of the split-precision leaf: moduli and CRT constants.
//
// The moduli are pairwise coprime and at most 256, so every symmetric residue
// fits in INT8 ([-128, 127] for 256, [-(m-1)/2, (m-1)/2] for odd m).  Taking
// them greedily downward from 256 maximises log2(M) for a given count.
//
// CRT weights W_l = (M / m_l) * ((M / m_l)^-1 mod m_l) are < M and satisfy
// W_l = 1 (mod m_l), W_l = 0 (mod m_j), j != l; they are split into base-2^37
// digits so that residue * digit (< 2^44) and a sum of up to 20 such (< 2^49)
// are exact in FP64.
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ozaki.h"

namespace caps {

namespace {

// little-endian base-2^32 unsigned big integer, just enough for the constants
struct Big {
  std::vector<uint32_t> d;
  explicit Big(uint64_t v = 0) {
    while (v) {
      d.push_back((uint32_t)v);
      v >>= 32;
    }
  }
  void mul(uint32_t x) {
    uint64_t carry = 0;
    for (auto &w : d) {
      uint64_t t = (uint64_t)w * x + carry;
      w = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) d.push_back((uint32_t)carry);
  }
  uint32_t divmod(uint32_t x) {  // *this /= x, returns remainder
    uint64_t r = 0;
    for (size_t i = d.size(); i-- > 0;) {
      uint64_t cur = (r << 32) | d[i];
      d[i] = (uint32_t)(cur / x);
      r = cur % x;
    }
    while (!d.empty() && d.back() == 0) d.pop_back();
    return (uint32_t)r;
  }
  uint32_t mod(uint32_t x) const {
    uint64_t r = 0;
    for (size_t i = d.size(); i-- > 0;) r = ((r << 32) | d[i]) % x;
    return (uint32_t)r;
  }
  uint64_t bits(int pos, int len) const {  // bits [pos, pos+len), len <= 48
    uint64_t out = 0;
    for (int b = 0; b < len; b++) {
      int p = pos + b;
      size_t w = (size_t)p / 32;
      if (w < d.size() && ((d[w] >> (p % 32)) & 1u)) out |= 1ull << b;
    }
    return out;
  }
  int bitlen() const {
    if (d.empty()) return 0;
    int hb = 31;
    while (!((d.back() >> hb) & 1u)) hb--;
    return (int)(d.size() - 1) * 32 + hb + 1;
  }
  double to_double() const {  // rounded toward zero to 53 bits
    int bl = bitlen();
    if (bl <= 53) return (double)bits(0, bl < 1 ? 1 : bl);
    return std::ldexp((double)bits(bl - 53, 53), bl - 53);
  }
};

int inverse_mod(int a, int m) {  // a^-1 mod m, gcd(a, m) = 1
  int t = 0, nt = 1, r = m, nr = a % m;
  while (nr) {
    int q = r / nr;
    int tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  return t < 0 ? t + m : t;
}

std::vector<int> make_moduli() {
  std::vector<int> out;
  for (int c = 256; c >= 2 && (int)out.size() < OZ_MAX_MODULI; c--) {
    bool ok = true;
    for (int o : out)
      if (std::gcd(c, o) != 1) {
        ok = false;
        break;
      }
    if (ok) out.push_back(c);
  }
  return out;
}

}  // namespace

const int *ozaki_moduli_table() {
  static std::vector<int> t = make_moduli();
  return t.data();
}

int ozaki_npieces(int nmod) {
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return OZ_MAX_PIECES;
  Big M(1);
  for (int l = 0; l < nmod; l++) M.mul((uint32_t)ozaki_moduli_table()[l]);
  return (M.bitlen() + OZ_PIECE_BITS - 1) / OZ_PIECE_BITS;
}

bool ozaki_make_consts(int nmod, OzakiConsts &k) {
  if (nmod < 2 || nmod > OZ_MAX_MODULI) return false;
  memset(&k, 0, sizeof(k));
  const int *mods = ozaki_moduli_table();
  k.nmod = nmod;
  Big M(1);
  for (int l = 0; l < nmod; l++) {
    k.m[l] = mods[l];
    {
      int c = (1 << 14) % mods[l];
      if (c > (mods[l] - 1) / 2) c -= mods[l];
      k.c14[l] = c;
    }
    k.c22[l] = ((1 << 23) / mods[l] + 1) >> 1;  // round(2^22 / m)
    k.inv_m[l] = 1.0 / (double)mods[l];
    M.mul((uint32_t)mods[l]);
  }
  const int mbits = M.bitlen();
  k.npieces = (mbits + OZ_PIECE_BITS - 1) / OZ_PIECE_BITS;
  if (k.npieces > OZ_MAX_PIECES) return false;
  for (int p = 0; p < k.npieces; p++)
    k.Mp[p] = (double)M.bits(p * OZ_PIECE_BITS, OZ_PIECE_BITS);
  for (int l = 0; l < nmod; l++) {
    Big Ml = M;
    Ml.divmod((uint32_t)mods[l]);                  // M / m_l
    int y = inverse_mod((int)Ml.mod((uint32_t)mods[l]), mods[l]);
    Ml.mul((uint32_t)y);                          // W_l < M
    for (int p = 0; p < k.npieces; p++)
      k.w[l][p] = (double)Ml.bits(p * OZ_PIECE_BITS, OZ_PIECE_BITS);
  }
  const double Md = M.to_double();  // <= M
  k.invM = 1.0 / Md;
  k.log2M = std::log2(Md);
  // L^2 <= (M/2)(1 - 2^-30): keeps |X| at least 2^-31 M away from M/2, far
  // beyond the 2^-36 M error of the quotient estimate in the reconstruction
  // (inputs congruent to the residues with |r| <= 2^11).
  const double D = Md * 0.5 * (1.0 - std::ldexp(1.0, -30));
  k.L = std::sqrt(D) * (1.0 - std::ldexp(1.0, -40));
  return true;
}

}  // namespace caps
