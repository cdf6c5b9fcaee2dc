// pair_math.h -- integer helpers of the fused pair kernel, shared with host
// code (and its tests): the low-half window prefilter and the division by
// an invariant divisor.
//
// Window prefilter.
// A binary-rooted candidate c = op(a, b) of size sb is a hit when c, or a
// chain of unary nodes above it, can equal the target t: with U unary nodes
// still fitting (U = max_size - sb >= 1) that is
//   c in [t - floor(U/2), t + ceil(U/2)]       (even chains, c = t - k), or
//   c in [-t - ceil(U/2), -t + floor(U/2)]     (odd chains,  c = k - t),
// two windows of span U.  The prefilter works on the low 32 bits of every
// value (a value inside a window [lo, lo + U] has lo32(c) - lo32(lo) <= U)
// and folds the seven operator instances of a pair (a, b) into one running
// minimum of distances; the pair may hit when that minimum is at most
// psp = 2 * ceil(U/2) (U, or U + 1 when U is odd).  The test is a necessary
// condition only: pairs that pass are re-checked exactly.
//
//  - a ^ b:  (a ^ b) - lo_e for the even window; for the odd one the value's
//    complement, ~(a ^ b) = (~a) ^ b, lies in [~(lo_o + U), ~lo_o], so
//    ((~a) ^ b) + lo_o + U + 1 <= U.  With ~a staged per A value each
//    distance is one xor-add.
//  - a * b:  the product's distances to both window starts.
//  - a + b, a - b, b - a: as windows on b, both parities of the three
//    instances are the intervals of radius h = ceil(U/2) around +-(t - a)
//    and +-(t + a).  fold(x) = min(x, -x) is the circular distance of x to 0
//    in Z/2^32, so |fold(b) - fold(c)| <= dist(b, c) and dist(b, -c)
//    (triangle inequality): b within h of +-c gives
//    fold(b) - (fold(c) - h) in [0, 2h].  Two subtractions cover all six
//    windows (fold(b) is formed once per B value, fold(t -+ a) - h once per
//    A value).
// The & and | instances are checked separately, only for the A values whose
// bits allow them to reach a window, and only against the windows they can
// reach.
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#define MBA_HD __host__ __device__
#else
#define MBA_HD
#endif

namespace mba {

// Circular distance of x to 0 in Z/2^32.
MBA_HD inline uint32_t fold32(uint32_t x) {
  const uint32_t n = 0u - x;
  return x < n ? x : n;
}

MBA_HD inline uint32_t min3_u32(uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t m = a < b ? a : b;
  return m < c ? m : c;
}

// Launch-uniform constants of the window prefilter (low halves).
struct WinPre {
  uint32_t le;   // lo32 of the even window's start
  uint32_t lo;   // lo32 of the odd window's start
  uint32_t kx1;  // -le: (a ^ b) + kx1 is the even distance
  uint32_t kx2;  // lo + U + 1: ((~a) ^ b) + kx2 is the odd distance
  uint32_t t;    // lo32 of the target
  uint32_t h;    // ceil(U / 2)
  uint32_t psp;  // threshold 2h
};

MBA_HD inline WinPre win_pre(uint32_t le, uint32_t lo, uint32_t span) {
  WinPre w;
  w.le = le;
  w.lo = lo;
  w.kx1 = 0u - le;
  w.kx2 = lo + span + 1u;
  w.t = le + (span >> 1);
  w.h = (span + 1u) >> 1;
  w.psp = 2u * w.h;
  return w;
}

// (a ^ b) + c.  On the device one v_xad_u32 (the compiler keeps the xor and
// the add apart); a pure ALU instruction, so its operands are plain values.
MBA_HD inline uint32_t xad_u32(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_xad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
#else
  return (a ^ b) + c;
#endif
}

// Per-A values: f1 = fold(t - a) - h, f2 = fold(t + a) - h.
MBA_HD inline void win_stage(const WinPre &w, uint32_t al, uint32_t &f1, uint32_t &f2) {
  f1 = fold32(w.t - al) - w.h;
  f2 = fold32(w.t + al) - w.h;
}

// The five distances of ^ * and the additive instances of the pair (a, b)
// folded into acc; fb = fold32(b), nal = ~a.
MBA_HD inline uint32_t win_pair(const WinPre &w, uint32_t al, uint32_t nal, uint32_t f1,
                                uint32_t f2, uint32_t b, uint32_t fb, uint32_t acc) {
  const uint32_t x1 = xad_u32(al, b, w.kx1);
  const uint32_t x2 = xad_u32(nal, b, w.kx2);
  const uint32_t m = al * b;
  acc = min3_u32(acc, x1, x2);
  acc = min3_u32(acc, m - w.le, m - w.lo);
  return min3_u32(acc, fb - f1, fb - f2);
}

// Whether a & b (AND) or a | b (OR) can land in the window [w, w + span]
// for some b (low halves): every value of the window shares its bits above
// the span's bit length p with w or with w + span, and a & b lies inside a
// (a | b contains a), so one of those two high parts must lie inside a
// (contain a).
template <bool AND>
MBA_HD inline bool window_bitop(uint32_t al, uint32_t w, uint32_t span) {
  int p = 0;
  while (p < 32 && (span >> p) != 0u) ++p;
  if (p >= 32) return true;
  const uint32_t ah = al >> p, h0 = w >> p, h1 = (w + span) >> p;
  return AND ? ((h0 & ~ah) == 0u || (h1 & ~ah) == 0u) : ((ah & ~h0) == 0u || (ah & ~h1) == 0u);
}

// The (instance, window) combinations an A value must be checked for: bit 0
// a & b in the even window, bit 1 a & b in the odd one, bit 2 a | b in the
// even one, bit 3 a | b in the odd one.  Besides the bit test above, a small
// A value needs neither: for 0 <= a <= h = ceil(span/2), b <= a | b <= b + a,
// so a | b in a window W puts b in [W - a, W + a] (W's ends moved out by a),
// which the additive tests already cover (b in W - a from a + b, b in W + a
// from b - a; the two meet as 2a <= span + 1); likewise a & b for
// -h <= a <= -1, as b - (-a - 1) <= a & b <= b.
enum : uint32_t { kAndEven = 1u, kAndOdd = 2u, kOrEven = 4u, kOrOdd = 8u };
MBA_HD inline uint32_t bitop_flags(uint32_t al, uint32_t le, uint32_t lo, uint32_t span) {
  const uint32_t h = (span + 1u) >> 1;
  const bool and_covered = 0u - al <= h && al != 0u, or_covered = al <= h;
  return (!and_covered && window_bitop<true>(al, le, span) ? kAndEven : 0u) |
         (!and_covered && window_bitop<true>(al, lo, span) ? kAndOdd : 0u) |
         (!or_covered && window_bitop<false>(al, le, span) ? kOrEven : 0u) |
         (!or_covered && window_bitop<false>(al, lo, span) ? kOrOdd : 0u);
}

// Distances of x = a & b (or a | b) to the windows its flags allow (two
// flag bits, even then odd), folded into acc.
MBA_HD inline uint32_t win_bitop_dist(const WinPre &w, uint32_t x, uint32_t fl2,
                                      uint32_t acc) {
  const uint32_t de = (fl2 & 1u) ? x - w.le : 0xffffffffu;
  const uint32_t dd = (fl2 & 2u) ? x - w.lo : 0xffffffffu;
  return min3_u32(acc, de, dd);
}

// & and | distances of the pair (a, b) folded into acc, for the
// combinations in `flags` (bitop_flags).
MBA_HD inline uint32_t win_pair_bitops(const WinPre &w, uint32_t al, uint32_t b,
                                       uint32_t flags, uint32_t acc) {
  acc = win_bitop_dist(w, al & b, flags & 3u, acc);
  return win_bitop_dist(w, al | b, flags >> 2, acc);
}

// Quotient of n < 2^32 by an invariant divisor d >= 1 (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, figure 4.1): with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d)
// + 1, q = (t1 + ((n - t1) >> min(l, 1))) >> max(l - 1, 0), t1 = mulhi(m, n).
struct DivMagic {
  uint32_t m;         // the multiplier
  uint32_t sh1, sh2;  // min(l, 1), max(l - 1, 0)
};

inline DivMagic div_magic(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) ++l;
  DivMagic g;
  g.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
  g.sh1 = l < 1 ? l : 1;
  g.sh2 = l > 1 ? l - 1 : 0;
  return g;
}

MBA_HD inline uint32_t div_u32(uint32_t n, const DivMagic &g) {
  const uint32_t t1 = (uint32_t)(((uint64_t)g.m * n) >> 32);
  return (t1 + ((n - t1) >> g.sh1)) >> g.sh2;
}

// floor(n / d), d >= 1, for 64-bit operands without the software 64-bit
// division: 32-bit operands divide directly; otherwise the double-precision
// quotient (within a few units of the true one: each operand is rounded to
// 53 bits) is corrected by whole steps of d.
MBA_HD inline uint64_t udiv64(uint64_t n, uint64_t d) {
  if ((n >> 32) == 0 && (d >> 32) == 0) return (uint32_t)n / (uint32_t)d;
  double qd = (double)n / (double)d;
  uint64_t q = qd >= 18446744073709551615.0 ? ~0ull : (uint64_t)qd;
  // lower q while q * d exceeds n (or overflows 64 bits)
  for (;;) {
    const unsigned __int128 p = (unsigned __int128)q * d;
    if (p <= n) break;
    --q;
  }
  while (n - q * d >= d) ++q;  // q * d <= n here, so no wrap
  return q;
}

}  // namespace mba
