// mba_common.h -- grammar, count table and operator semantics shared by the
// host driver and the gfx950 kernels.
//
// Operator indices follow the reference's op order (globals.cuh:13-15,
// nocache.cu:251): 0 '~', 1 '&', 2 '|', 3 '^', 4 neg ('#'), 5 '+', 6 '-',
// 7 '*'.  Leaves are numbered 0..L-1: variables first, then constants.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace mba {

constexpr int kMaxSize = 32;
constexpr int kMaxLeaves = 16;
constexpr int kMaxSamples = 128;
constexpr uint8_t kTokOp = 32;  // RPN token: < 32 leaf, 32+o operator o

enum Op : int { NOT = 0, AND, OR, XOR, NEG, ADD, SUB, MUL };

__host__ __device__ inline bool is_unary(int op) { return op == NOT || op == NEG; }

template <typename word>
__host__ __device__ inline word apply_op(int op, word l, word r) {
  switch (op) {
    case NOT: return ~l;
    case AND: return l & r;
    case OR: return l | r;
    case XOR: return l ^ r;
    case NEG: return word(0) - l;
    case ADD: return l + r;
    case SUB: return l - r;
    default: return l * r;
  }
}

// Count table T[s][0..8], the reference's compute_n_formulas
// (nocache.cu:117-160): columns 0..7 = trees of size s with that root op,
// column 8 = all trees of size s.  T[1][8] = L and the size-1 op columns are
// zero.  Commutative ops only take left size j <= (s-1)/2; '-' takes every j.
struct CountTable {
  int L = 0, max_size = 0;
  uint64_t t[kMaxSize + 1][9] = {};
  // Returns false if any entry overflows uint64.
  bool build(int leaves, int max_sz) {
    L = leaves;
    max_size = max_sz;
    for (auto &row : t)
      for (auto &x : row) x = 0;
    t[1][8] = (uint64_t)leaves;
    for (int s = 2; s <= max_sz; ++s) {
      unsigned __int128 U = t[s - 1][8], C = 0, NC = 0;
      for (int j = 1; j <= s - 2; ++j) {
        unsigned __int128 p = (unsigned __int128)t[j][8] * t[s - 1 - j][8];
        NC += p;
        if (j <= (s - 1) / 2) C += p;
      }
      unsigned __int128 tot = 2 * U + 5 * C + NC;
      if ((tot >> 64) != 0) return false;
      t[s][NOT] = t[s][NEG] = (uint64_t)U;
      t[s][AND] = t[s][OR] = t[s][XOR] = t[s][ADD] = t[s][MUL] = (uint64_t)C;
      t[s][SUB] = (uint64_t)NC;
      t[s][8] = (uint64_t)tot;
    }
    return true;
  }
  uint64_t total(int s) const { return t[s][8]; }
  // Rank offset of the op block `op` at size s.
  uint64_t op_offset(int s, int op) const {
    uint64_t r = 0;
    for (int o = 0; o < op; ++o) r += t[s][o];
    return r;
  }
  // Rank offset of split j inside the binary op block at size s.
  uint64_t split_offset(int s, int j) const {
    uint64_t r = 0;
    for (int jj = 1; jj < j; ++jj) r += t[jj][8] * t[s - 1 - jj][8];
    return r;
  }
};

// Host unrank (same order as the reference's numberToFormula,
// nocache.cu:23-83): produces RPN tokens.
void unrank_host(const CountTable &T, int size, uint64_t rank, uint8_t *rpn);

// Infix printers: SIMBA (nocache.cu:162-195) and VFB (full-cache.cu:483-509).
std::string infix_simba(const uint8_t *rpn, int size, int n_vars,
                        const uint64_t *consts, int width);
std::string infix_vfb(const uint8_t *rpn, int size, int n_vars,
                      const uint64_t *consts, int width);
// Reference RPN characters: '0'+leaf and ~&|^#+-*.
std::string rpn_chars(const uint8_t *rpn, int size);

}  // namespace mba
