// Host check of the integer helpers of the fused pair kernel
// (simba_amd/csrc/pair_math.h).
//
// Division by an invariant divisor: div_u32(n, div_magic(d)) == n / d on
// every small (n, d), around powers of two and at random; udiv64(n, d) ==
// n / d around 2^32, 2^53, 2^63 and 2^64 and at random.
//
// Window prefilter: for every operator instance and parity,
// a pair whose value lies inside an exact unary-chain window must pass the
// folded low-half test.  Pairs are constructed to land on every offset of
// the windows (the instance inverted for b), plus random pairs checked
// against the exact windows, and the pass rate of random pairs is bounded
// (the test is not vacuous).  Exit status 0 when every check holds.
//
//   g++ -O2 -std=c++17 -I simba_amd/csrc pair_math_check.cpp
#include <cstdint>
#include <cstdio>
#include <random>

#include "pair_math.h"

using mba::WinPre;

namespace {

uint32_t inst_value(int inst, uint32_t a, uint32_t b) {
  switch (inst) {
    case 0: return a & b;
    case 1: return a | b;
    case 2: return a ^ b;
    case 3: return a + b;
    case 4: return a * b;
    case 5: return a - b;
    default: return b - a;
  }
}

// The exact windows of the kernel's targets (targets_for): even
// [t - U/2, t - U/2 + U], odd [-t - (U+1)/2, -t - (U+1)/2 + U].
struct Windows {
  uint32_t le, lo, span;
};

Windows windows(uint32_t t, uint32_t U) {
  return {t - U / 2, (0u - t) - (U + 1) / 2, U};
}

bool in_windows(const Windows &w, uint32_t v) {
  return v - w.le <= w.span || v - w.lo <= w.span;
}

// The kernel's prefilter of one pair: staged per-A values, fold of b, the
// & | branch when the A value's flag is set.
bool prefilter(const Windows &ws, uint32_t a, uint32_t b) {
  const WinPre w = mba::win_pre(ws.le, ws.lo, ws.span);
  uint32_t f1, f2;
  mba::win_stage(w, a, f1, f2);
  uint32_t acc = 0xffffffffu;
  const uint32_t fl = mba::bitop_flags(a, ws.le, ws.lo, ws.span);
  if (fl) acc = mba::win_pair_bitops(w, a, b, fl, acc);
  acc = mba::win_pair(w, a, ~a, f1, f2, b, mba::fold32(b), acc);
  return acc <= w.psp;
}

uint32_t inverse_odd(uint32_t a) {
  uint32_t x = a;
  for (int i = 0; i < 5; ++i) x *= 2u - a * x;
  return x;
}

// A b with inst(a, b) == v, when one exists (the & and | instances: with
// random bits where they are free).
bool solve_b(int inst, uint32_t a, uint32_t v, uint32_t r, uint32_t &b) {
  switch (inst) {
    case 0:  // a & b = v needs v inside a
      if (v & ~a) return false;
      b = v | (r & ~a);
      return true;
    case 1:  // a | b = v needs a inside v
      if (a & ~v) return false;
      b = (v & ~a) | (r & a);
      return true;
    case 2: b = a ^ v; return true;
    case 3: b = v - a; return true;
    case 4: {
      if (a == 0u) {
        if (v != 0u) return false;
        b = r;
        return true;
      }
      const int k = __builtin_ctz(a);
      if (k && (v & ((1u << k) - 1u))) return false;
      const uint32_t inv = inverse_odd(a >> k);
      b = ((v >> k) * inv) | (k ? (r << (32 - k)) : 0u);
      return true;
    }
    case 5: b = a - v; return true;
    default: b = v + a; return true;
  }
}

}  // namespace

int check_division(std::mt19937_64 &rng) {
  long bad = 0, n_checked = 0;
  auto check = [&](uint32_t n, uint32_t d) {
    ++n_checked;
    if (mba::div_u32(n, mba::div_magic(d)) != n / d) {
      if (bad < 10) std::printf("DIV n %u d %u got %u\n", n, d, mba::div_u32(n, mba::div_magic(d)));
      ++bad;
    }
  };
  for (uint32_t d = 1; d <= 300; ++d)
    for (uint32_t n = 0; n <= 5000; ++n) check(n, d);
  for (int e = 0; e < 32; ++e)
    for (int de = -2; de <= 2; ++de) {
      const uint32_t d = (uint32_t)((1ull << e) + de);
      if (d == 0) continue;
      for (int ne = 0; ne < 32; ++ne)
        for (int dn = -3; dn <= 3; ++dn) check((uint32_t)((1ull << ne) + dn), d);
      check(0xffffffffu, d);
      check(0xfffffffeu, d);
    }
  for (int i = 0; i < 2000000; ++i) {
    const uint32_t d = (uint32_t)rng() >> (rng() % 32);
    if (d) check((uint32_t)rng(), d);
  }
  std::printf("divisions checked %ld, wrong %ld\n", n_checked, bad);
  return bad ? 5 : 0;
}

int check_udiv64(std::mt19937_64 &rng) {
  long bad = 0, n_checked = 0;
  auto check = [&](uint64_t n, uint64_t d) {
    if (d == 0) return;
    ++n_checked;
    if (mba::udiv64(n, d) != n / d) {
      if (bad < 10)
        std::printf("UDIV64 n %llu d %llu got %llu\n", (unsigned long long)n,
                    (unsigned long long)d, (unsigned long long)mba::udiv64(n, d));
      ++bad;
    }
  };
  const uint64_t big[] = {0ull, 1ull, 2ull, 3ull, 0xffffffffull, 0x100000000ull, 0x100000001ull,
                          (1ull << 53) - 1, 1ull << 53, (1ull << 53) + 1, 1ull << 63,
                          (1ull << 63) + 1, ~0ull - 1, ~0ull};
  for (uint64_t n : big)
    for (uint64_t d : big) {
      for (int dn = -2; dn <= 2; ++dn) {
        check(n + dn, d);
        check(n, d + dn);
      }
    }
  for (int i = 0; i < 3000000; ++i) {
    const uint64_t n = rng() >> (rng() % 64), d = rng() >> (rng() % 64);
    check(n, d);
    if (d) check(n / d * d, d), check(n / d * d + d - 1 >= n / d * d ? n / d * d + d - 1 : n, d);
  }
  std::printf("64-bit divisions checked %ld, wrong %ld\n", n_checked, bad);
  return bad ? 6 : 0;
}

int main() {
  std::mt19937_64 rng(20261204);
  auto r32 = [&]() { return (uint32_t)rng(); };
  if (int rc = check_division(rng)) return rc;
  if (int rc = check_udiv64(rng)) return rc;
  const uint32_t edges[] = {0u, 1u, 2u, 0x7fffffffu, 0x80000000u, 0x80000001u, 0xfffffffeu,
                            0xffffffffu, 0x55555555u, 0xaaaaaaaau};
  long checked = 0, misses = 0;
  for (int trial = 0; trial < 60000; ++trial) {
    const uint32_t t = trial < 10 * 31 ? edges[trial % 10] : r32();
    const uint32_t U = 1u + (uint32_t)(trial % 31);
    const Windows ws = windows(t, U);
    uint32_t a = (trial / 3) % 7 == 0 ? edges[trial % 10] : r32();
    if (trial % 11 == 0) a = r32() & ~(uint32_t)rng() & 0xffff0000u;  // sparse bits
    if (trial % 13 == 0) a = (uint32_t)(trial / 13 % 81) - 40u;  // small |a|: 0, +-1, ...,
                                                                  // the covered bit-op flags
    for (int inst = 0; inst < 7; ++inst) {
      for (int par = 0; par < 2; ++par) {
        const uint32_t start = par ? ws.lo : ws.le;
        for (uint32_t k = 0; k <= U; ++k) {
          const uint32_t v = start + k;
          uint32_t aa = a;
          if (inst == 0) aa = a | v;   // make & reachable
          if (inst == 1) aa = a & v;   // make | reachable
          uint32_t b;
          if (!solve_b(inst, aa, v, r32(), b)) continue;
          if (inst_value(inst, aa, b) != v) {
            std::printf("construction error inst %d\n", inst);
            return 2;
          }
          ++checked;
          if (!prefilter(ws, aa, b)) {
            if (misses < 10)
              std::printf("MISS inst %d parity %d t %08x U %u a %08x b %08x v %08x\n", inst,
                          par, t, U, aa, b, v);
            ++misses;
          }
        }
      }
    }
  }
  // small A values against random and aimed B values, every window span
  for (int trial = 0; trial < 200000; ++trial) {
    const uint32_t t = r32(), U = 1u + (uint32_t)(trial % 31);
    const Windows ws = windows(t, U);
    const uint32_t a = (uint32_t)(trial % 81) - 40u;
    const uint32_t w0 = (trial & 1) ? ws.lo : ws.le;
    const uint32_t b = w0 + (uint32_t)(trial / 2 % 50) - 20u;  // around a window
    for (int inst = 0; inst < 7; ++inst) {
      if (!in_windows(ws, inst_value(inst, a, b))) continue;
      ++checked;
      if (!prefilter(ws, a, b)) {
        if (misses < 10)
          std::printf("MISS small a inst %d t %08x U %u a %08x b %08x\n", inst, t, U, a, b);
        ++misses;
      }
    }
  }
  // random pairs: any exact hit must pass; count the passes
  long passes = 0, exact = 0, free_passes = 0;
  const long n_rand = 4000000;
  for (long i = 0; i < n_rand; ++i) {
    const uint32_t t = r32(), U = 1u + (uint32_t)(i % 31);
    const Windows ws = windows(t, U);
    const uint32_t a = r32(), b = (i & 1) ? r32() : t - a + (uint32_t)(i % 17) - 8u;
    bool any = false;
    for (int inst = 0; inst < 7; ++inst) any = any || in_windows(ws, inst_value(inst, a, b));
    const bool pf = prefilter(ws, a, b);
    exact += any;
    passes += pf;
    if (i & 1) free_passes += pf && !any;
    if (any && !pf) {
      if (misses < 10) std::printf("MISS random t %08x U %u a %08x b %08x\n", t, U, a, b);
      ++misses;
    }
  }
  std::printf("constructed hits %ld, random pairs %ld (exact hits %ld, passed %ld, "
              "unaimed false passes %ld), misses %ld\n",
              checked, n_rand, exact, passes, free_passes, misses);
  if (misses) return 1;
  if (checked < 1000000) return 3;
  // half of the random pairs are aimed at the additive windows; the
  // unaimed half must almost never pass without a hit (a handful of windows
  // of at most 33 values among 2^32)
  if (free_passes > 20) return 4;
  return 0;
}
