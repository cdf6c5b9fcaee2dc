// expr.cpp -- host-side rank -> tree decoding and expression printing.
#include <cstdio>
#include <cstring>

#include "mba_common.h"

namespace mba {

// Iterative form of the reference's numberToFormula (nocache.cu:23-83):
// a work stack of (rank, size, position); the left child of a binary node of
// children-size cs and left size j occupies [pos, pos+j), the right child
// [pos+j, pos+cs), the operator sits at pos+cs.
void unrank_host(const CountTable &T, int size, uint64_t rank, uint8_t *rpn) {
  struct Item { uint64_t r; int s, pos; };
  Item st[kMaxSize + 1];
  int sp = 0;
  st[sp++] = {rank, size, 0};
  while (sp > 0) {
    Item it = st[--sp];
    if (it.s == 1) {
      rpn[it.pos] = (uint8_t)it.r;
      continue;
    }
    uint64_t n = it.r;
    int op = 0;
    while (n >= T.t[it.s][op]) n -= T.t[it.s][op++];
    int cs = it.s - 1;
    rpn[it.pos + cs] = (uint8_t)(kTokOp + op);
    if (is_unary(op)) {
      st[sp++] = {n, cs, it.pos};
      continue;
    }
    int j = 1;
    for (;; ++j) {
      uint64_t blk = T.t[j][8] * T.t[cs - j][8];
      if (n < blk) break;
      n -= blk;
    }
    uint64_t rs = T.t[cs - j][8];
    st[sp++] = {n / rs, j, it.pos};
    st[sp++] = {n % rs, cs - j, it.pos + j};
  }
}

static std::string leaf_name(int leaf, int n_vars, const uint64_t *consts,
                             int width) {
  char b[48];
  if (leaf < n_vars) {
    snprintf(b, sizeof b, "v%d", leaf);
  } else {
    uint64_t m = width == 64 ? ~0ull : 0xffffffffull;
    uint64_t c = consts[leaf - n_vars] & m;
    if (c >> (width - 1))
      snprintf(b, sizeof b, "-%llu", (unsigned long long)((0 - c) & m));
    else
      snprintf(b, sizeof b, "%llu", (unsigned long long)c);
  }
  return b;
}

static const char kOpCh[8] = {'~', '&', '|', '^', '-', '+', '-', '*'};

static std::string infix(const uint8_t *rpn, int size, int n_vars,
                         const uint64_t *consts, int width, bool vfb) {
  std::vector<std::string> st;
  for (int i = 0; i < size; ++i) {
    uint8_t t = rpn[i];
    if (t < kTokOp) {
      st.push_back(leaf_name(t, n_vars, consts, width));
      continue;
    }
    int op = t - kTokOp;
    std::string o(1, kOpCh[op]);
    if (is_unary(op)) {
      if (st.empty()) return "Error : Incorrect formula";
      std::string x = st.back();
      st.back() = vfb ? o + "(" + x + ")" : "(" + o + x + ")";
    } else {
      if (st.size() < 2) return "Error : Incorrect formula (stack size < 2).";
      std::string r = st.back();
      st.pop_back();
      std::string l = st.back();
      st.back() = vfb ? "(" + l + ")" + o + "(" + r + ")"
                      : "(" + l + " " + o + " " + r + ")";
    }
  }
  if (st.size() != 1) return "Error : Incorrect formula (stack size != 2).";
  return st[0];
}

std::string infix_simba(const uint8_t *rpn, int size, int n_vars,
                        const uint64_t *consts, int width) {
  return infix(rpn, size, n_vars, consts, width, false);
}

std::string infix_vfb(const uint8_t *rpn, int size, int n_vars,
                      const uint64_t *consts, int width) {
  return infix(rpn, size, n_vars, consts, width, true);
}

std::string rpn_chars(const uint8_t *rpn, int size) {
  static const char ops[8] = {'~', '&', '|', '^', '#', '+', '-', '*'};
  std::string s;
  for (int i = 0; i < size; ++i)
    s += rpn[i] < kTokOp ? char('0' + rpn[i]) : ops[rpn[i] - kTokOp];
  return s;
}

}  // namespace mba
