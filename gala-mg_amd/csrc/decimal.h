// Exact decimal -> binary64 conversion for the device text ingest.
//
// The edge-list weight token is a decimal number,
//   [+-]? ( D+ ( '.' D* )? | '.' D+ ) ( [eE] [+-]? D+ )?
// whose value is rounded to the nearest double, ties to even -- the result
// C's strtod gives for the same token.  Two methods:
//   * S < 10^15 and |E| <= 22 (value = S * 10^E): S and 10^|E| are exact
//     doubles, so one IEEE multiply or divide rounds the exact value once
//     (Clinger, "How to read floating point numbers accurately", PLDI 1990);
//   * otherwise exact integer arithmetic on 32-bit limbs: X = S * 5^E when
//     E >= 0, or X = floor(S * 2^t / 5^-E) with the remainder kept as a sticky
//     bit when E < 0; then X * 2^b is rounded to 53 bits (or to the subnormal
//     grid) with round-half-even.
// At most DEC_MAX_SIG significant digits (between the first and the last
// nonzero digit) are accepted; a longer token, or one that rounds to an
// infinity, is a format error.  Written for both the host and the device
// compiler so that the conversion can be checked against strtod on the CPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GALA_HD __host__ __device__
#else
#define GALA_HD
#endif

namespace gala {
namespace dec {

constexpr int DEC_MAX_SIG = 40;  // 10^40 < 2^133: five 32-bit limbs
constexpr int XL = 32;           // limbs of the working integer (1024 bits)

enum : int { DEC_OK = 0, DEC_BAD = 1, DEC_RANGE = 2 };

GALA_HD inline bool is_ws(uint8_t c) {
  return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}
GALA_HD inline bool is_digit(uint8_t c) { return c >= '0' && c <= '9'; }

struct Big {
  uint32_t l[XL];
  int n;  // limbs in use; l[n-1] != 0 unless n == 0
};

GALA_HD inline void big_mul_small(Big& x, uint32_t m, uint32_t add) {
  uint64_t carry = add;
  for (int i = 0; i < x.n; i++) {
    uint64_t t = (uint64_t)x.l[i] * m + carry;
    x.l[i] = (uint32_t)t;
    carry = t >> 32;
  }
  if (carry && x.n < XL) x.l[x.n++] = (uint32_t)carry;
}

// x /= d; returns whether the remainder was nonzero
GALA_HD inline bool big_div_small(Big& x, uint32_t d) {
  uint64_t r = 0;
  for (int i = x.n - 1; i >= 0; i--) {
    uint64_t cur = (r << 32) | x.l[i];
    x.l[i] = (uint32_t)(cur / d);
    r = cur % d;
  }
  while (x.n > 0 && x.l[x.n - 1] == 0) x.n--;
  return r != 0;
}

GALA_HD inline int big_bitlen(const Big& x) {
  if (x.n == 0) return 0;
  uint32_t top = x.l[x.n - 1];
  int b = 0;
  while (top) { b++; top >>= 1; }
  return 32 * (x.n - 1) + b;
}

GALA_HD inline int big_bit(const Big& x, int pos) {
  if (pos < 0 || pos >= 32 * x.n) return 0;
  return (x.l[pos >> 5] >> (pos & 31)) & 1;
}

// bits [s, s + 64) of x (s >= 0)
GALA_HD inline uint64_t big_shr64(const Big& x, int s) {
  uint64_t r = 0;
  for (int k = 0; k < 64; k++) r |= (uint64_t)big_bit(x, s + k) << k;
  return r;
}

// any bit below position s set
GALA_HD inline bool big_any_below(const Big& x, int s) {
  int full = s >> 5;
  for (int i = 0; i < full && i < x.n; i++)
    if (x.l[i]) return true;
  if (full < x.n && (s & 31)) return (x.l[full] & ((1u << (s & 31)) - 1)) != 0;
  return false;
}

GALA_HD inline void big_shl(Big& x, int t) {  // caller guarantees room
  int w = t >> 5, b = t & 31;
  if (x.n == 0) return;
  int nn = x.n + w + 1;
  if (nn > XL) nn = XL;
  for (int i = nn - 1; i >= 0; i--) {
    int src = i - w;
    uint32_t hi = (src >= 0 && src < x.n) ? x.l[src] : 0u;
    uint32_t lo = (src - 1 >= 0 && src - 1 < x.n) ? x.l[src - 1] : 0u;
    x.l[i] = b ? (hi << b) | (lo >> (32 - b)) : hi;
  }
  x.n = nn;
  while (x.n > 0 && x.l[x.n - 1] == 0) x.n--;
}

GALA_HD inline uint32_t pow5(int r) {  // 5^r for r <= 13 (5^13 < 2^31)
  uint32_t p = 1;
  for (int i = 0; i < r; i++) p *= 5;
  return p;
}

GALA_HD inline double bits_to_double(uint64_t u) {
  union { uint64_t u; double d; } c;
  c.u = u;
  return c.d;
}

// exactly S * 10^E rounded to a double (S > 0 given by limbs; nsig digits)
GALA_HD inline int round_big(Big& S, int nsig, int64_t E, bool neg, double* out) {
  const uint64_t sgn = neg ? (1ull << 63) : 0;
  if (nsig + E - 1 >= 309) return DEC_RANGE;          // >= 10^309 > DBL_MAX
  if (nsig + E <= -324) { *out = bits_to_double(sgn); return DEC_OK; }  // < 1e-324: rounds to 0
  int64_t b;
  bool sticky = false;
  if (E >= 0) {
    for (int64_t k = E; k > 0; k -= 13) big_mul_small(S, pow5(k >= 13 ? 13 : (int)k), 0);
    b = E;
  } else {
    const int64_t k = -E;  // <= 363 here
    const int bits5 = (int)(k * 2.3219280949) + 1;  // >= bit length of 5^k
    int t = 67 + bits5 - big_bitlen(S);
    if (t < 0) t = 0;
    big_shl(S, t);
    for (int64_t r = k; r > 0; r -= 13) sticky |= big_div_small(S, pow5(r >= 13 ? 13 : (int)r));
    b = -(int64_t)t - k;
  }
  const int L = big_bitlen(S);
  const int64_t elead = L - 1 + b;
  if (elead > 1023) return DEC_RANGE;
  int64_t u = elead - 52;
  if (u < -1074) u = -1074;
  const int64_t s = u - b;
  uint64_t m;
  if (s <= 0) {
    m = big_shr64(S, 0) << (-s);  // exact: X has at most 53 bits here
  } else {
    m = s >= L + 1 ? 0 : big_shr64(S, (int)s);
    const int guard = big_bit(S, (int)(s - 1));
    const bool rest = sticky || big_any_below(S, (int)(s - 1));
    if (guard && (rest || (m & 1))) m++;
  }
  if (m == (1ull << 53)) { m >>= 1; u++; }
  if (m >= (1ull << 52) && u + 52 > 1023) return DEC_RANGE;
  uint64_t bitsv = m < (1ull << 52) ? m : ((uint64_t)(u + 52 + 1023) << 52) | (m - (1ull << 52));
  *out = bits_to_double(sgn | bitsv);
  return DEC_OK;
}

// Parse the token [p, q) (no whitespace inside).  Returns DEC_OK with *out set,
// DEC_BAD for a token outside the grammar or with more than DEC_MAX_SIG
// significant digits, DEC_RANGE when it rounds to an infinity.
template <class Get>
GALA_HD inline int parse_decimal(Get get, int64_t p, int64_t q, double* out) {
  bool neg = false;
  if (p < q && (get(p) == '+' || get(p) == '-')) { neg = get(p) == '-'; p++; }
  Big S;
  S.n = 0;
  int nsig = 0, zpend = 0;
  int64_t nint = 0, nfrac = 0;  // digit counts before and after the point
  bool point = false;
  for (; p < q; p++) {
    const uint8_t c = get(p);
    if (c == '.') {
      if (point) return DEC_BAD;
      point = true;
      continue;
    }
    if (!is_digit(c)) break;
    if (point) nfrac++; else nint++;
    const uint32_t d = c - '0';
    if (d == 0) {
      if (nsig > 0) zpend++;
      continue;
    }
    if (nsig + zpend + 1 > DEC_MAX_SIG) return DEC_BAD;
    for (int z = 0; z < zpend; z++) big_mul_small(S, 10, 0);
    if (S.n == 0) S.l[S.n++] = d; else big_mul_small(S, 10, d);
    nsig += zpend + 1;
    zpend = 0;
  }
  if (nint + nfrac == 0) return DEC_BAD;
  int64_t ex = 0;
  if (p < q) {
    const uint8_t c = get(p);
    if (c != 'e' && c != 'E') return DEC_BAD;
    p++;
    bool eneg = false;
    if (p < q && (get(p) == '+' || get(p) == '-')) { eneg = get(p) == '-'; p++; }
    if (p >= q) return DEC_BAD;
    for (; p < q; p++) {
      const uint8_t e = get(p);
      if (!is_digit(e)) return DEC_BAD;
      if (ex < 100000000) ex = ex * 10 + (e - '0');
    }
    if (eneg) ex = -ex;
  }
  const uint64_t sgn = neg ? (1ull << 63) : 0;
  if (nsig == 0) { *out = bits_to_double(sgn); return DEC_OK; }
  const int64_t E = ex - nfrac + zpend;  // value = S * 10^E
  if (nsig <= 15 && E >= -22 && E <= 22) {
    const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                            1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    uint64_t s = 0;
    for (int i = S.n - 1; i >= 0; i--) s = (s << 32) | S.l[i];
    const double v = E >= 0 ? (double)s * p10[E] : (double)s / p10[-E];
    *out = neg ? -v : v;
    return DEC_OK;
  }
  return round_big(S, nsig, E, neg, out);
}

// Integer token [p, q): [+-]? D+ .  Magnitude saturates at 2^40.
template <class Get>
GALA_HD inline bool parse_int(Get get, int64_t p, int64_t q, bool* neg, uint64_t* mag) {
  *neg = false;
  if (p < q && (get(p) == '+' || get(p) == '-')) { *neg = get(p) == '-'; p++; }
  if (p >= q) return false;
  uint64_t v = 0;
  for (; p < q; p++) {
    const uint8_t c = get(p);
    if (!is_digit(c)) return false;
    if (v < (1ull << 40)) v = v * 10 + (c - '0');
  }
  *mag = v;
  return true;
}

}  // namespace dec
}  // namespace gala
