// simba_kernels.hip -- gfx950 kernels of the SIMBA (cache-free, unranking)
// search.
//
// The reference (nocache.cu:85-107, mba.cuh:6-69) gives one thread one rank:
// unrank to RPN in local memory, then a divergent stack evaluation.  Here a
// size level is instead swept as outer products.  Every binary-rooted tree of
// size s is op(X, Y) with X of size j and Y of size m = s-1-j; for an
// unordered split {j, m} the seven operator instances & | ^ + * (X-Y) (Y-X)
// are fused and evaluated on one filter sample from the value tables V[j],
// V[m] (value of every tree of that size on that sample, in rank order).  A
// lane holds RB (4 or 8) values of the B side in VGPRs; the A side is staged
// in LDS and read as wave-uniform broadcasts, so the inner loop is integer
// VALU work with no divergence.  Unary-rooted trees are the composition of a
// chain of ~/neg with a binary-rooted tree or a leaf; a chain of u nodes maps
// x to (-1)^u x + k with |k| <= (u+1)/2, so two windowed compares per value
// (span U = the nodes that still fit) cover every unary tree of every larger
// size.  Hits are rare; they are appended
// to a hit buffer, resolved to (size, rank) on the device (k_hit_screen,
// k_hit_confirm) and verified on all samples (verify_wave).
#include "simba_kernels.h"
#include "pair_math.h"

namespace mba {

// Table memory is device (global) memory.  Its pointer reaches the kernels
// through TableView as a generic pointer; routing it through the global
// address space lets the compiler emit global rather than flat accesses
// (flat loads also count against the LDS/scalar wait counter).
template <typename T>
__device__ __forceinline__ T *global_ptr(const void *p) {
  typedef T __attribute__((address_space(1))) *gp;
  return (T *)(gp)(uintptr_t)p;
}


template <typename word>
struct WordOf;

// ---------------------------------------------------------------------------
// Table-assisted value of the tree (m, r) on the filter sample: unrank
// top-down in the reference's order, but stop at any subtree of size <= Kt and
// read its value from V[size]; then evaluate the remaining skeleton in
// postfix order.
// ---------------------------------------------------------------------------
template <typename word>
__device__ __forceinline__ word build_value(const TableView *__restrict__ tv, int m,
                                            uint64_t r);

template <typename word>
__device__ word value_of(const TableView *__restrict__ tv, int m, uint64_t r) {
  const word *V = global_ptr<const word>(tv->base);
  const int Kt = tv->Kt;
  if (m <= Kt) return V[tv->off[m] + r];
  // one level above the tables: the root's children are tabulated
  if (m - 1 <= Kt) return build_value<word>(tv, m, r);
  uint8_t tok[kMaxSize];     // 0xFF: inside a tabulated subtree
  uint64_t ref[kMaxSize];    // table index for a 0x40|size token
  struct Item { uint64_t r; int s, pos; };
  Item st[kMaxSize];
  for (int i = 0; i < m; ++i) tok[i] = 0xFF;
  int sp = 0;
  st[sp++] = {r, m, 0};
  while (sp > 0) {
    Item it = st[--sp];
    if (it.s <= Kt) {
      tok[it.pos + it.s - 1] = (uint8_t)(0x40 | it.s);
      ref[it.pos + it.s - 1] = it.r;
      continue;
    }
    uint64_t n = it.r;
    int op = 0;
    while (n >= tv->T[it.s][op]) n -= tv->T[it.s][op++];
    const int cs = it.s - 1;
    tok[it.pos + cs] = (uint8_t)(0x80 | op);
    if (is_unary(op)) {
      st[sp++] = {n, cs, it.pos};
      continue;
    }
    int j = 1;
    for (;; ++j) {
      uint64_t blk = tv->T[j][8] * tv->T[cs - j][8];
      if (n < blk) break;
      n -= blk;
    }
    const uint64_t rs = tv->T[cs - j][8];
    const uint64_t a = udiv64(n, rs);
    st[sp++] = {a, j, it.pos};
    st[sp++] = {n - a * rs, cs - j, it.pos + j};
  }
  word vs[kMaxSize];
  int top = -1;
  for (int i = 0; i < m; ++i) {
    const uint8_t t = tok[i];
    if (t == 0xFF) continue;
    if (t & 0x40) {
      vs[++top] = V[tv->off[t & 0x3F] + ref[i]];
    } else {
      const int op = t & 7;
      if (is_unary(op)) {
        vs[top] = apply_op<word>(op, vs[top], word(0));
      } else {
        const word rv = vs[top--];
        vs[top] = apply_op<word>(op, vs[top], rv);
      }
    }
  }
  return vs[0];
}

// ---------------------------------------------------------------------------
// Unrank (size, rank) into reference RPN tokens (numberToFormula,
// nocache.cu:23-83) with a per-lane stack.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unrank_rpn(const TableView *__restrict__ tv, int size,
                                           uint64_t rank, uint8_t *rpn) {
  struct Item { uint64_t r; int s, pos; };
  Item st[kMaxSize];
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
    while (n >= tv->T[it.s][op]) n -= tv->T[it.s][op++];
    const int cs = it.s - 1;
    rpn[it.pos + cs] = (uint8_t)(kTokOp + op);
    if (is_unary(op)) {
      st[sp++] = {n, cs, it.pos};
      continue;
    }
    int j = 1;
    for (;; ++j) {
      const uint64_t blk = tv->T[j][8] * tv->T[cs - j][8];
      if (n < blk) break;
      n -= blk;
    }
    const uint64_t rs = tv->T[cs - j][8];
    const uint64_t a = udiv64(n, rs);
    st[sp++] = {a, j, it.pos};
    st[sp++] = {n - a * rs, cs - j, it.pos + j};
  }
}

// Value of the tree (size, rank) on one sample's leaf values.
template <typename word>
__device__ word eval_tree(const TableView *__restrict__ tv, int size, uint64_t rank,
                          const uint64_t *leaves) {
  uint8_t rpn[kMaxSize];
  unrank_rpn(tv, size, rank, rpn);
  word vs[kMaxSize];
  int top = -1;
  for (int i = 0; i < size; ++i) {
    const uint8_t t = rpn[i];
    if (t < kTokOp) {
      vs[++top] = (word)leaves[t];
    } else {
      const int op = t - kTokOp;
      if (is_unary(op)) {
        vs[top] = apply_op<word>(op, vs[top], word(0));
      } else {
        const word rv = vs[top--];
        vs[top] = apply_op<word>(op, vs[top], rv);
      }
    }
  }
  return vs[0];
}

// ---------------------------------------------------------------------------
// Table build: V[m][r] for r in [0, T[m]).  One lane per rank; the op block
// and split are located by prefix sums over the count table (the reference's
// unrank order, nocache.cu:52-79), the children are read from V[<m].
// ---------------------------------------------------------------------------
template <typename word>
__device__ __forceinline__ word build_value32(const TableView *__restrict__ tv, int m,
                                              uint32_t r) {
  // every count below is at most T[m][8] < 2^32: 32-bit arithmetic is exact
  const word *V = global_ptr<const word>(tv->base);
  uint32_t x = r;
  int op = 0;
  while (x >= (uint32_t)tv->T[m][op]) x -= (uint32_t)tv->T[m][op++];
  if (is_unary(op)) return apply_op<word>(op, V[tv->off[m - 1] + x], word(0));
  const int cs = m - 1;
  int j = 1;
  for (;; ++j) {
    const uint32_t blk = (uint32_t)tv->T[j][8] * (uint32_t)tv->T[cs - j][8];
    if (x < blk) break;
    x -= blk;
  }
  const uint32_t rs = (uint32_t)tv->T[cs - j][8];
  const uint32_t a = x / rs;
  return apply_op<word>(op, V[tv->off[j] + a], V[tv->off[cs - j] + (x - a * rs)]);
}

template <typename word>
__device__ __forceinline__ word build_value(const TableView *__restrict__ tv, int m,
                                            uint64_t r) {
  if (tv->T[m][8] <= 0xffffffffull) return build_value32<word>(tv, m, (uint32_t)r);
  const word *V = global_ptr<const word>(tv->base);
  uint64_t x = r;
  int op = 0;
  while (x >= tv->T[m][op]) x -= tv->T[m][op++];
  if (is_unary(op)) return apply_op<word>(op, V[tv->off[m - 1] + x], word(0));
  const int cs = m - 1;
  int j = 1;
  for (;; ++j) {
    const uint64_t blk = tv->T[j][8] * tv->T[cs - j][8];
    if (x < blk) break;
    x -= blk;
  }
  const uint64_t rs = tv->T[cs - j][8];
  const uint64_t a = x / rs;
  return apply_op<word>(op, V[tv->off[j] + a], V[tv->off[cs - j] + (x - a * rs)]);
}

// The children of V[m][r] as table element offsets (ib unused for a unary
// root) and the root op: build_value without the loads, so a lane can issue
// the loads of several values together.
__device__ __forceinline__ int build_refs(const TableView *__restrict__ tv, int m, uint64_t r,
                                          uint64_t &ia, uint64_t &ib) {
  uint64_t x = r;
  int op = 0;
  while (x >= tv->T[m][op]) x -= tv->T[m][op++];
  if (is_unary(op)) {
    ia = tv->off[m - 1] + x;
    ib = ia;
    return op;
  }
  const int cs = m - 1;
  int j = 1;
  for (;; ++j) {
    const uint64_t blk = tv->T[j][8] * tv->T[cs - j][8];
    if (x < blk) break;
    x -= blk;
  }
  const uint64_t rs = tv->T[cs - j][8];
  uint64_t a;
  if ((x >> 32) == 0 && rs <= 0xffffffffull) a = (uint32_t)x / (uint32_t)rs;
  else a = x / rs;
  ia = tv->off[j] + a;
  ib = tv->off[cs - j] + (x - a * rs);
  return op;
}

template <typename word>
__global__ void __launch_bounds__(256)
k_table_build(const TableView *__restrict__ tv, word *__restrict__ out, int m,
              uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += stride)
    out[r] = build_value<word>(tv, m, r);
}

// One (op, left size) block of V[m]: out[a * nB + b] = op(VA[a], VB[b]) for
// a < nA, b < nB (unary blocks: nB = 1, VB unused).  A tile is 1024 outputs,
// 2^logTB columns by 1024 >> logTB rows, so lanes need no division and the
// stores of a wave are contiguous.
template <typename word>
__global__ void __launch_bounds__(256)
k_table_block(const word *__restrict__ VA, const word *__restrict__ VB, word *__restrict__ out,
              int op, uint64_t nA, uint64_t nB, int logTB, uint64_t nBT, uint64_t ntiles) {
  const uint32_t TB = 1u << logTB;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t at = t / nBT, bt = t - at * nBT;
    for (uint32_t k = threadIdx.x; k < 1024; k += 256) {
      const uint64_t a = at * (1024 >> logTB) + (k >> logTB);
      const uint64_t b = bt * TB + (k & (TB - 1));
      if (a < nA && b < nB) {
        const word x = VA[a];
        out[a * nB + b] = is_unary(op) ? apply_op<word>(op, x, word(0))
                                       : apply_op<word>(op, x, VB[b]);
      }
    }
  }
}

// Value tables of the small sizes m_lo..m_hi by one workgroup, size after
// size: each value from its root op and the smaller sizes' tables
// (build_value); the barrier between sizes orders the workgroup's stores
// before the next size reads them.
template <typename word>
__global__ void __launch_bounds__(1024)
k_table_small(const TableView *__restrict__ tv, int m_lo, int m_hi) {
  for (int m = m_lo; m <= m_hi; ++m) {
    word *out = global_ptr<word>(tv->base) + tv->off[m];
    const uint64_t n = tv->T[m][8];
    for (uint64_t r = threadIdx.x; r < n; r += blockDim.x) out[r] = build_value<word>(tv, m, r);
    __syncthreads();
  }
}

template <typename word>
__device__ __forceinline__ bool in_window(word c, word lo_e, word lo_o, word span) {
  return (word)(c - lo_e) <= span || (word)(c - lo_o) <= span;
}

template <typename word>
__device__ __forceinline__ word inst_value(int inst, word a, word b) {
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

template <typename word>
__device__ __noinline__ void record_hit(HitRec *hits, uint32_t *count, uint32_t cap,
                                        uint64_t a, uint64_t b, word v, int inst,
                                        uint32_t pair) {
  const uint32_t i = atomicAdd(count, 1u);
  if (i < cap) {
    HitRec h;
    h.a = a;
    h.b = b;
    h.v = (uint64_t)v;
    h.inst = (uint32_t)inst;
    h.pair = pair;
    hits[i] = h;
  }
}

// ---------------------------------------------------------------------------
// Fused pair check over a tile range.  Tile t covers A values
// [at*TA, at*TA+TA) x B values [bt*TB, bt*TB+TB) with at = t % nTA.
// SAME: j == m, so (Y - X) is the (X - Y) instance of the swapped pair and is
// skipped.
//
// Each lane holds RB B values; the A values of the tile are wave-uniform
// (scalar registers), so everything that depends on A alone is formed once
// per A value on the scalar unit.  Only the low 32 bits of a result are
// tested in the loop: a value c inside a window [lo, lo + span] (span < 2^32)
// has lo32(c) - lo32(lo) <= span (mod 2^32), and lo32 of & | ^ + - * needs
// only the low halves of the operands (the 64-bit product's low half is one
// 32-bit multiply).  Lanes that pass re-walk their tile with the exact
// full-width test.
//
// Window mode (unary chains above the root still possible, span > 0): the
// seven distances of a pair are folded into one running minimum; the
// additive instances fold A into the window offset, so each costs one
// subtraction per window:
//   a + b - lo = b - (lo - a),  a - b - lo = (a - lo) - b,  b - a - lo = b - (a + lo).
// Exact mode (span = 0: the candidate itself must equal the target t; the
// odd window is empty because no unary node fits above the root): the
// invertible instances become one compare of b against a uniform value
// (b = a ^ t, t - a, a - t, t + a), * is one multiply and a compare, and
// a & b = t (a | b = t) is only possible when t's bits lie inside a's (a's
// inside t's), a uniform test that skips the instance for the whole tile row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lo32(uint32_t x) { return x; }
__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }

__device__ __forceinline__ uint32_t min3u(uint32_t a, uint32_t b, uint32_t c) {
  return min(min(a, b), c);
}

// Moves a wave-uniform value into scalar registers, so that the expressions
// built from it stay on the scalar unit.
__device__ __forceinline__ uint32_t uniform(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}
__device__ __forceinline__ uint64_t uniform(uint64_t x) {
  return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
         __builtin_amdgcn_readfirstlane((uint32_t)x);
}

// Per-A values of a tile, staged in LDS by the workgroup (one A value per
// lane) and read back as wave-uniform broadcasts inside the loop:
//   exact mode:   q0 = {a, a ^ t, t - a, a - t},
//                 q1 = {t + a, bit-op gap, product residue, shift | kind << 8}
//   window mode, every instance (pair_math.h):
//                 q0 = {a, ~a, fold(t - a) - h, fold(t + a) - h},
//                 q1 = {-, -, -, bitop_flags: the & | windows a can reach}
//   window mode, some instances:
//                 q0 = {a, lo_e - a, lo_o - a, a - lo_e},
//                 q1 = {a - lo_o, a + lo_e, a + lo_o, -}
// (low halves; t = lo_e in exact mode).
struct AStage {
  uint4 q0[kPairTA];
  uint4 q1[kPairTA];
  uint64_t full[kPairTA];  // the full-width A values (exact re-walk of a tile)
  uint32_t fmask[kPairTA / 32];  // window mode: bit i % 32 of word i / 32 set
                                 // when A value i has bitop flags
};

// a * b = t (mod 2^32) as a compare: with a = 2^k a' (a' odd), it holds iff
// 2^k divides t and b = (t >> k) * a'^-1 (mod 2^(32-k)), i.e. (b << k) ==
// residue << k.  kind 0: no b fits, 1: compare, 2: every b (a = t = 0).
// The inverse of an odd a' by Newton's iteration x <- x (2 - a' x), which
// doubles the correct low bits from the 3 of x = a' (a'^2 = 1 mod 8).
__device__ __forceinline__ uint32_t mul_residue(uint32_t a, uint32_t t, uint32_t &shift_kind) {
  if (a == 0u) {
    shift_kind = (t == 0u ? 2u : 0u) << 8;
    return 0u;
  }
  const uint32_t k = (uint32_t)__builtin_ctz(a);
  if (k != 0u && (t & ((1u << k) - 1u)) != 0u) {
    shift_kind = 0u;
    return 0u;
  }
  const uint32_t ao = a >> k;
  uint32_t inv = ao;
#pragma unroll
  for (int it = 0; it < 4; ++it) inv *= 2u - ao * inv;
  shift_kind = k | (1u << 8);
  return ((t >> k) * inv) << k;
}

// a & b = t needs t inside a, a | b = t needs a inside t: zero when either
// instance can hit for some b.
__device__ __forceinline__ uint32_t bitops_gap(uint32_t al, uint32_t tl) {
  return min(tl & ~al, al & ~tl);
}

// Staged layout of a launch: the exact-mode layout when every operator
// instance is checked against a single target (span 0), the folded window
// layout when every instance is checked against windows, else the
// per-instance window layout.
enum : int { kStageExact = 0, kStageWindow = 1, kStageInst = 2 };
__device__ __forceinline__ int stage_mode(uint32_t mask, bool same, uint32_t span) {
  const uint32_t full = same ? 0x3fu : 0x7fu;
  if ((mask & full) != full) return kStageInst;
  return span == 0u ? kStageExact : kStageWindow;
}

// Stages A values [a0, a1) (GEN_A: computed per lane from the tables) in
// the layout `mode` (stage_mode).  Returns, for the exact mode, whether any A
// value admits & or | hits.
template <typename word>
__device__ __forceinline__ bool stage_a(AStage &S, const TableView *__restrict__ tv,
                                        const word *__restrict__ A, bool gen_a, int j,
                                        uint64_t a0, uint64_t a1, int mode, uint32_t le,
                                        uint32_t lo, uint32_t span) {
  __syncthreads();  // the previous tile's readers are done
  const uint32_t i = threadIdx.x;
  bool bit = false;
  uint32_t fl = 0u;  // window mode: this A value's bitop flags
  if (a0 + i < a1) {
    const word av = gen_a ? value_of<word>(tv, j, a0 + i) : A[a0 + i];
    S.full[i] = (uint64_t)av;
    const uint32_t al = lo32(av);
    if (mode == kStageExact) {
      const uint32_t gap = bitops_gap(al, le);
      uint32_t shift_kind;
      const uint32_t res = mul_residue(al, le, shift_kind);
      S.q0[i] = make_uint4(al, al ^ le, le - al, al - le);
      S.q1[i] = make_uint4(le + al, gap, res, shift_kind);
      bit = gap == 0u;
    } else if (mode == kStageWindow) {
      const WinPre w = win_pre(le, lo, span);
      uint32_t f1, f2;
      win_stage(w, al, f1, f2);
      fl = bitop_flags(al, le, lo, span);
      S.q0[i] = make_uint4(al, ~al, f1, f2);
      S.q1[i] = make_uint4(0u, 0u, 0u, fl);
    } else {
      S.q0[i] = make_uint4(al, le - al, lo - al, al - le);
      S.q1[i] = make_uint4(al - lo, al + le, al + lo, 0u);
    }
  }
  {  // the flagged A values as bit masks, 32 per word (a wave covers two)
    const uint64_t m = __ballot(fl != 0u);
    if ((i & 63u) == 0u) {
      S.fmask[i >> 5] = (uint32_t)m;
      S.fmask[(i >> 5) + 1] = (uint32_t)(m >> 32);
    }
  }
  return __syncthreads_or(bit) != 0;
}

// ---- B side sources --------------------------------------------------------
// kBDesc: blocks formed from tables, row-band or general tiles chosen per
// block on the device; kBBand / kBGrid: only row-band / only general block
// tiles (the per-instance launcher knows which, and each instance holds one
// path).
enum : int { kBTable = 0, kBDesc = 1, kBGeneric = 2, kBBand = 3, kBGrid = 4 };

template <typename word>
__device__ __forceinline__ uint32_t lo32_at(const word *p, uint64_t i) {
  return lo32(p[i]);
}
template <>
__device__ __forceinline__ uint32_t lo32_at<uint64_t>(const uint64_t *p, uint64_t i) {
  return ((const uint32_t *)p)[2 * i];  // little-endian low half
}

// Row band of block tile bt: bt / ntc, by the invariant-divisor multiply
// when both fit 32 bits (the usual case), else a 64-bit division.
__device__ __forceinline__ uint64_t block_row(const BBlock &bb, uint64_t bt) {
  if (bb.ntc == 1) return bt;
  if ((bt >> 32) == 0 && bb.ntc <= 0xffffffffull) return div_u32((uint32_t)bt, bb.ntc_div);
  return bt / bb.ntc;
}

// Coordinates of the k-th B value of this lane in block tile bt.
template <int RB>
__device__ __forceinline__ bool desc_coords(const BBlock &bb, uint64_t bt, int k, uint64_t &ra,
                                            uint64_t &rb) {
  const uint64_t tr = block_row(bb, bt);
  const uint64_t tc = bt - tr * bb.ntc;
  const uint32_t idx = threadIdx.x + (uint32_t)k * kPairThreads;
  const uint32_t TB = 1u << bb.logTB;
  const uint64_t x = tc * TB + (idx & (TB - 1));
  const uint64_t y = tr * (uint64_t)((kPairThreads * RB) >> bb.logTB) + (idx >> bb.logTB);
  ra = bb.trans ? x : y;
  rb = bb.trans ? y : x;
  return ra < bb.nA && rb < bb.nB;
}

// The path steps of a block (innermost first) on one value.
template <typename word>
__device__ __forceinline__ word apply_steps(const BBlock &bb, const TableView *__restrict__ tv,
                                            const word *V, word v) {
  for (int i = 0; i < bb.nsteps; ++i) {
    const BStep &st = bb.steps[i];
    if (st.kind == 0) {
      if (st.op) v = word(0) - v;
      v = word(v + (word)st.c);
    } else {
      const word c = V[tab_off(tv, st.c)];
      v = st.kind == 1 ? apply_op<word>(st.op, c, v) : apply_op<word>(st.op, v, c);
    }
  }
  return v;
}

template <typename word>
__device__ __forceinline__ word desc_value(const BBlock &bb, const TableView *__restrict__ tv,
                                           const word *V, uint64_t ra, uint64_t rb) {
  word v = V[tab_off(tv, bb.offA) + ra];
  if (bb.op >= 0) v = apply_op<word>(bb.op, v, V[tab_off(tv, bb.offB) + rb]);
  return apply_steps<word>(bb, tv, V, v);
}

__device__ __forceinline__ uint32_t lo_op(int op, uint32_t a, uint32_t b) {
  switch (op) {
    case AND: return a & b;
    case OR: return a | b;
    case XOR: return a ^ b;
    case ADD: return a + b;
    case SUB: return a - b;
    default: return a * b;
  }
}

// The path steps on the low halves of a lane's RB values: an affine step is
// -v = (v ^ ~0) - ~0, so s*v + k = (v ^ sg) + (k - sg) with sg = 0 or ~0; a
// fixed sibling's low half is one uniform load.
template <typename word, int RB>
__device__ __forceinline__ void steps_lo(const BBlock &bb, const TableView *__restrict__ tv,
                                         const word *V, uint32_t *bl) {
  for (int i = 0; i < bb.nsteps; ++i) {
    const BStep &st = bb.steps[i];
    if (st.kind == 0) {
      const uint32_t sg = st.op ? 0xffffffffu : 0u;
      const uint32_t kk = (uint32_t)st.c - sg;
#pragma unroll
      for (int k = 0; k < RB; ++k) bl[k] = (bl[k] ^ sg) + kk;
    } else {
      const uint32_t c = lo32_at<word>(V, tab_off(tv, st.c));
      if (st.kind == 1) {
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = lo_op(st.op, c, bl[k]);
      } else {
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = lo_op(st.op, bl[k], c);
      }
    }
  }
}

// Low 32 bits of the RB B values of this lane in a block tile of the general
// 2-D mapping (blocks whose columns span less than a tile; row-band tiles go
// through desc_band_tile): inner value, then the path steps.
template <typename word, int RB, int OP>
__device__ __forceinline__ void desc_lo_op(const BBlock &bb, const TableView *__restrict__ tv,
                                           const word *V, uint64_t bt, uint32_t *bl,
                                           bool *ok) {
  auto comb = [](uint32_t a, uint32_t b) -> uint32_t {
    return OP == AND ? (a & b) : OP == OR ? (a | b) : OP == XOR ? (a ^ b)
         : OP == ADD ? (a + b) : OP == SUB ? (a - b) : OP == MUL ? (a * b) : a;
  };
  const word *VA = V + tab_off(tv, bb.offA);
  const word *VB = OP >= 0 ? V + tab_off(tv, bb.offB) : V;
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    uint64_t ra, rb;
    ok[k] = desc_coords<RB>(bb, bt, k, ra, rb);
    if (!ok[k]) ra = 0, rb = 0;
    uint32_t v = lo32_at<word>(VA, ra);
    if (OP >= 0) v = comb(v, lo32_at<word>(VB, rb));
    bl[k] = v;
  }
  if (bb.nsteps) steps_lo<word, RB>(bb, tv, V, bl);
}

template <typename word, int RB>
__device__ __forceinline__ void desc_lo(const BBlock &bb, const TableView *__restrict__ tv,
                                        const word *V, uint64_t bt, uint32_t *bl, bool *ok) {
  switch (bb.op) {
    case AND: desc_lo_op<word, RB, AND>(bb, tv, V, bt, bl, ok); break;
    case OR: desc_lo_op<word, RB, OR>(bb, tv, V, bt, bl, ok); break;
    case XOR: desc_lo_op<word, RB, XOR>(bb, tv, V, bt, bl, ok); break;
    case ADD: desc_lo_op<word, RB, ADD>(bb, tv, V, bt, bl, ok); break;
    case SUB: desc_lo_op<word, RB, SUB>(bb, tv, V, bt, bl, ok); break;
    case MUL: desc_lo_op<word, RB, MUL>(bb, tv, V, bt, bl, ok); break;
    default: desc_lo_op<word, RB, -1>(bb, tv, V, bt, bl, ok); break;
  }
}

// Full B value and rank (within size m) of the k-th value of this lane
// (used by the exact re-walk only).
template <typename word, int BSRC, int RB, typename PL>
__device__ __forceinline__ word b_full(const TableView *__restrict__ tv,
                                       const word *__restrict__ B, const PL &p,
                                       uint64_t bt, int k, uint64_t &rank) {
  if (BSRC == kBDesc || BSRC == kBBand || BSRC == kBGrid) {
    uint64_t ra, rb;
    desc_coords<RB>(p.bb, bt, k, ra, rb);
    rank = p.bb.start + p.bb.stride * (ra * p.bb.nB + rb);
    return desc_value<word>(p.bb, tv, global_ptr<const word>(tv->base), ra, rb);
  }
  const uint64_t local = bt * (kPairThreads * RB) + threadIdx.x + (uint64_t)k * kPairThreads;
  if (BSRC == kBGeneric) {
    rank = p.bb.start + p.bb.stride * local;
    return value_of<word>(tv, p.m, rank);
  }
  rank = local;
  if (BSRC == kBTable) return B[local];
  return value_of<word>(tv, p.m, local);
}

// Low halves of this lane's B values of tile bt (and their validity).
template <typename word, int BSRC, int RB, typename PL>
__device__ __forceinline__ void b_lo(const TableView *__restrict__ tv,
                                     const word *__restrict__ B, const PL &p,
                                     uint64_t bt, uint32_t *bl, bool *ok) {
  if (BSRC == kBDesc || BSRC == kBGrid) {
    desc_lo<word, RB>(p.bb, tv, global_ptr<const word>(tv->base), bt, bl, ok);
    return;
  }
  const uint64_t b0 = bt * (kPairThreads * RB) + threadIdx.x;
  const uint64_t n = BSRC == kBGeneric ? p.bb.nA : p.nB;
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const uint64_t b = b0 + (uint64_t)k * kPairThreads;
    ok[k] = b < n;
    if (BSRC == kBTable)
      bl[k] = ok[k] ? lo32_at<word>(B, b) : 0u;
    else
      bl[k] = ok[k] ? lo32(value_of<word>(tv, p.m, p.bb.start + p.bb.stride * b)) : 0u;
  }
}

// Whether lo32(op(a, b)) lies in a window's low half for some operator
// instance (a necessary condition of the exact test, as in the tile loops).
__device__ __forceinline__ bool pair_may_hit(uint32_t a, uint32_t b, uint32_t le, uint32_t lo,
                                             uint32_t span) {
  const uint32_t x_and = a & b, x_or = a | b, x_xor = a ^ b, x_add = a + b, x_mul = a * b;
  const uint32_t x_ab = a - b, x_ba = b - a;
  uint32_t m = min3u(x_and - le, x_and - lo, x_or - le);
  m = min3u(m, x_or - lo, x_xor - le);
  m = min3u(m, x_xor - lo, x_add - le);
  m = min3u(m, x_add - lo, x_mul - le);
  m = min3u(m, x_mul - lo, x_ab - le);
  m = min3u(m, x_ab - lo, x_ba - le);
  return min(m, x_ba - lo) <= span;
}

// ---- the fused checks -------------------------------------------------------
// Exact mode: each compare's lane mask is or-ed into one wave-wide mask on
// the scalar unit (one VALU compare + one scalar or per candidate); the
// product is a compare of b with the A value's residue (mul_residue).
template <bool SAME, int RB>
__device__ __forceinline__ void exact_a(uint64_t &m, const uint4 q0, const uint4 q1,
                                        const uint32_t *bl) {
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const uint32_t b = bl[k];
    m |= __builtin_amdgcn_ballot_w64(b == q0.y);            // a ^ b = t
    m |= __builtin_amdgcn_ballot_w64(b == q0.z);            // a + b = t
    m |= __builtin_amdgcn_ballot_w64(b == q0.w);            // a - b = t
    if (!SAME) m |= __builtin_amdgcn_ballot_w64(b == q1.x);  // b - a = t
  }
  const uint32_t sk = uniform(q1.w), res = q1.z;
  if ((sk >> 8) == 1u) {                                    // a * b = t
    const uint32_t sh = sk & 0xffu;
    if (sh == 0u) {
#pragma unroll
      for (int k = 0; k < RB; ++k) m |= __builtin_amdgcn_ballot_w64(bl[k] == res);
    } else {
#pragma unroll
      for (int k = 0; k < RB; ++k) m |= __builtin_amdgcn_ballot_w64((bl[k] << sh) == res);
    }
  } else if ((sk >> 8) == 2u) {
    m = ~0ull;  // a = t = 0: every b; the re-walk checks them exactly
  }
}

template <bool SAME, int RB>
__device__ __forceinline__ bool exact_tile(const AStage &S, uint32_t na, const uint32_t *bl,
                                           uint32_t tl, bool bitops) {
  uint64_t m = 0;
  uint32_t i;
#pragma unroll 1
  for (i = 0; i < na; ++i) exact_a<SAME, RB>(m, S.q0[i], S.q1[i], bl);
  if (bitops) {  // rare: & and | for the A values that admit them
    for (i = 0; i < na; ++i) {
      if (S.q1[i].y != 0u) continue;
      const uint32_t al = S.q0[i].x;
#pragma unroll
      for (int k = 0; k < RB; ++k) {
        m |= __builtin_amdgcn_ballot_w64((al & bl[k]) == tl);
        m |= __builtin_amdgcn_ballot_w64((al | bl[k]) == tl);
      }
    }
  }
  return (m >> (threadIdx.x & 63)) & 1ull;
}

// Window mode, every instance: the folded prefilter of pair_math.h (ten
// VALU instructions per pair: two xor-adds, a product and its two
// distances, two distances of fold(b), three min3), & and | only for the A
// values whose staged flag says they can reach a window (a wave-uniform
// branch per A value).  The result is compared with w.psp.  (SAME pairs need
// no separate B - A check: its windows are those of the additive tests.)
// The & (AND) or | distances of one A value and a lane's RB B values to the
// windows in f2 (bit 0 even, bit 1 odd; workgroup-uniform): one subtraction
// per B value when a single window is reachable.
template <bool AND, int RB>
__device__ __forceinline__ uint32_t bitop_tile(const WinPre &w, uint32_t al,
                                               const uint32_t *bl, uint32_t f2, uint32_t acc) {
  if (f2 == 3u) {
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const uint32_t x = AND ? (al & bl[k]) : (al | bl[k]);
      acc = min3u(acc, x - w.le, x - w.lo);
    }
  } else {
    const uint32_t ws = f2 == 1u ? w.le : w.lo;
#pragma unroll
    for (int k = 0; k < RB; k += 2) {
      const uint32_t x0 = AND ? (al & bl[k]) : (al | bl[k]);
      const uint32_t x1 = AND ? (al & bl[k + 1]) : (al | bl[k + 1]);
      acc = min3u(acc, x0 - ws, x1 - ws);
    }
  }
  return acc;
}

template <int RB>
__device__ __forceinline__ uint32_t window_tile(const AStage &S, uint32_t na,
                                                const uint32_t *bl, const WinPre &w) {
  static_assert(RB % 2 == 0, "B values per lane come in pairs");
  uint32_t fb[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) fb[k] = fold32(bl[k]);
  uint32_t acc = 0xffffffffu;
  uint32_t fm = 0u;  // flagged A values of the current 32
  for (uint32_t i = 0; i < na; ++i) {
    if ((i & 31u) == 0u) fm = uniform(S.fmask[i >> 5]);
    const uint4 q0 = S.q0[i];
    if ((fm >> (i & 31u)) & 1u) {
      const uint32_t fl = uniform(S.q1[i].w);  // bitop_flags of this A value
      if (fl & (kAndEven | kAndOdd)) acc = bitop_tile<true, RB>(w, q0.x, bl, fl & 3u, acc);
      if (fl & (kOrEven | kOrOdd)) acc = bitop_tile<false, RB>(w, q0.x, bl, fl >> 2, acc);
    }
#pragma unroll
    for (int k = 0; k < RB; ++k) acc = win_pair(w, q0.x, q0.y, q0.z, q0.w, bl[k], fb[k], acc);
  }
  return acc;
}

// One operator instance (block-at-a-time mode, window form, any span); the
// instance is a template parameter so each loop holds only its own check.
template <int INST, int RB>
__device__ __forceinline__ uint32_t inst_tile_t(const AStage &S, uint32_t na,
                                                const uint32_t *bl, uint32_t le, uint32_t lo,
                                                uint32_t acc) {
  for (uint32_t i = 0; i < na; ++i) {
    const uint4 q0 = S.q0[i];
    const uint4 q1 = S.q1[i];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const uint32_t b = bl[k];
      uint32_t d0, d1;
      if (INST == 0) d0 = (q0.x & b) - le, d1 = (q0.x & b) - lo;
      else if (INST == 1) d0 = (q0.x | b) - le, d1 = (q0.x | b) - lo;
      else if (INST == 2) d0 = (q0.x ^ b) - le, d1 = (q0.x ^ b) - lo;
      else if (INST == 3) d0 = b - q0.y, d1 = b - q0.z;
      else if (INST == 4) d0 = q0.x * b - le, d1 = q0.x * b - lo;
      else if (INST == 5) d0 = q0.w - b, d1 = q1.x - b;
      else d0 = b - q1.y, d1 = b - q1.z;
      acc = min3u(acc, d0, d1);
    }
  }
  return acc;
}

template <int RB>
__device__ __forceinline__ uint32_t inst_tile(const AStage &S, uint32_t na, int inst,
                                              const uint32_t *bl, uint32_t le, uint32_t lo,
                                              uint32_t acc) {
  switch (inst) {  // wave-uniform
    case 0: return inst_tile_t<0, RB>(S, na, bl, le, lo, acc);
    case 1: return inst_tile_t<1, RB>(S, na, bl, le, lo, acc);
    case 2: return inst_tile_t<2, RB>(S, na, bl, le, lo, acc);
    case 3: return inst_tile_t<3, RB>(S, na, bl, le, lo, acc);
    case 4: return inst_tile_t<4, RB>(S, na, bl, le, lo, acc);
    case 5: return inst_tile_t<5, RB>(S, na, bl, le, lo, acc);
    default: return inst_tile_t<6, RB>(S, na, bl, le, lo, acc);
  }
}

// Low-half prefilter of a lane's RB B values against the staged A values of
// a tile: whether some pair of this lane may hit (exact, folded-window or
// per-instance window test, by the launch's layout).
template <bool SAME, int RB, int FIX = -1>
__device__ __forceinline__ bool pair_prefilter(const AStage &S, uint32_t na, const uint32_t *bl,
                                               const bool *ok, uint32_t mask, bool bitops,
                                               uint32_t le, uint32_t lo, uint32_t sp) {
  bool any_ok = false;
#pragma unroll
  for (int k = 0; k < RB; ++k) any_ok |= ok[k];
  const uint32_t full = SAME ? 0x3fu : 0x7fu;
  const int mode = FIX >= 0 ? FIX : stage_mode(mask, SAME, sp);
  bool hit;
  if (mode == kStageExact) {
    hit = exact_tile<SAME, RB>(S, na, bl, le, bitops);
  } else if (mode == kStageWindow) {
    const WinPre w = win_pre(le, lo, sp);
    hit = window_tile<RB>(S, na, bl, w) <= w.psp;
  } else {
    uint32_t acc = 0xffffffffu;
    for (int inst = 0; inst < 7; ++inst)
      if ((mask & full) >> inst & 1u) acc = inst_tile<RB>(S, na, inst, bl, le, lo, acc);
    hit = acc <= sp;
  }
  return hit && any_ok;
}

// Rare: some lane passed the prefilter.  Its tile is re-walked with the exact
// test, the A values from LDS (full width): per B value, count its hits (the
// pairs that pass the low-half test of every instance checked in full),
// reserve their records with one atomic and write them.  bfull(k, rank)
// returns the full k-th B value of this lane and its rank.
template <typename word, bool SAME, int RB, typename BFull>
__device__ __forceinline__ void pair_rewalk(const AStage &S, uint32_t na, uint64_t a0,
                                            const bool *ok, uint32_t mask, word lo_e, word lo_o,
                                            word span, uint32_t pair, uint32_t tag,
                                            HitRec *__restrict__ hits,
                                            uint32_t *__restrict__ hit_count, uint32_t hit_cap,
                                            BFull bfull) {
  const uint32_t le = lo32(lo_e), lo = lo32(lo_o), sp = (uint32_t)span;
  const int ninst = SAME ? 6 : 7;
  for (int k = 0; k < RB; ++k) {
    if (!ok[k]) continue;
    uint64_t b;
    const word bv = bfull(k, b);
    const uint32_t b32 = lo32(bv);
    uint32_t n = 0;
    for (uint32_t i = 0; i < na; ++i) {
      const word av = (word)S.full[i];
      if (!pair_may_hit(lo32(av), b32, le, lo, sp)) continue;
      for (int inst = 0; inst < ninst; ++inst)
        n += (mask >> inst & 1u) && in_window(inst_value<word>(inst, av, bv), lo_e, lo_o, span);
    }
    if (n == 0) continue;
    uint32_t r = atomicAdd(hit_count, n);
    for (uint32_t i = 0; i < na && r < hit_cap; ++i) {
      const word av = (word)S.full[i];
      if (!pair_may_hit(lo32(av), b32, le, lo, sp)) continue;
      for (int inst = 0; inst < ninst; ++inst) {
        const word c = inst_value<word>(inst, av, bv);
        if (!((mask >> inst & 1u) && in_window(c, lo_e, lo_o, span))) continue;
        if (r < hit_cap) {
          HitRec h;
          h.a = a0 + i;
          h.b = b;
          h.v = (uint64_t)c;
          h.inst = (uint32_t)(inst | tag);
          h.pair = pair;
          hits[r] = h;
        }
        ++r;
      }
    }
  }
}

// Row-band tile of a B block (its columns span at least a tile): the tile's
// column values are loaded once (low halves) and combined with each of its
// rows (up to bb.rows_pt block rows, one uniform load each), so a column
// value read from memory serves every row of the tile.
template <typename word, bool SAME, int RB, typename PL, int FIX = -1>
__device__ __forceinline__ void desc_band_tile(const TableView *__restrict__ tv,
                                               const AStage &S, const PL &p,
                                               uint32_t na, uint64_t a0, uint64_t bt,
                                               bool bitops, word lo_e, word lo_o, word span,
                                               HitRec *__restrict__ hits,
                                               uint32_t *__restrict__ hit_count,
                                               uint32_t hit_cap, uint32_t tag) {
  constexpr uint32_t tile = kPairThreads * RB;
  const BBlock &bb = p.bb;
  const word *V = global_ptr<const word>(tv->base);
  const uint64_t tr = block_row(bb, bt);
  const uint64_t tc = bt - tr * bb.ntc;
  const uint64_t x0 = tc * tile;
  const uint64_t nX = bb.trans ? bb.nA : bb.nB, nY = bb.trans ? bb.nB : bb.nA;
  const uint64_t left = nX - x0;
  const uint32_t rem = left > tile ? tile : (uint32_t)left;
  const word *col = V + tab_off(tv, bb.trans ? bb.offA : bb.offB) + x0;
  const word *rowp = V + (bb.op < 0 ? 0 : tab_off(tv, bb.trans ? bb.offB : bb.offA));
  uint32_t cl[RB], bl[RB];
  bool ok[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const uint32_t off = threadIdx.x + (uint32_t)k * kPairThreads;
    ok[k] = off < rem;
    cl[k] = lo32_at<word>(col, ok[k] ? off : 0u);
  }
  const uint64_t y0 = tr * (uint64_t)bb.rows_pt;
  const uint32_t nr = (uint32_t)(nY - y0 < (uint64_t)bb.rows_pt ? nY - y0 : bb.rows_pt);
  const uint32_t pair = ((uint32_t)p.sb << 16) | (uint32_t)p.j;
  const uint32_t le = lo32(lo_e), lo = lo32(lo_o), sp = (uint32_t)span;
  // the tile's row values in one register: lane q < nr loads row q, each
  // row is read back with one lane read (up to kBandRows rows)
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t rows_v = bb.op < 0 || lane >= nr ? 0u : lo32_at<word>(rowp, y0 + lane);
#pragma unroll 1
  for (uint32_t r = 0; r < nr; ++r) {  // workgroup-uniform
    const uint64_t y = y0 + r;
    const uint32_t rowv = __builtin_amdgcn_readlane(rows_v, r);
    switch (bb.op) {  // workgroup-uniform; only '-' depends on the order
      case AND:
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = cl[k] & rowv;
        break;
      case OR:
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = cl[k] | rowv;
        break;
      case XOR:
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = cl[k] ^ rowv;
        break;
      case ADD:
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = cl[k] + rowv;
        break;
      case MUL:
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = cl[k] * rowv;
        break;
      case SUB:
        if (bb.trans) {
#pragma unroll
          for (int k = 0; k < RB; ++k) bl[k] = cl[k] - rowv;
        } else {
#pragma unroll
          for (int k = 0; k < RB; ++k) bl[k] = rowv - cl[k];
        }
        break;
      default:
#pragma unroll
        for (int k = 0; k < RB; ++k) bl[k] = cl[k];
        break;
    }
    if (bb.nsteps) steps_lo<word, RB>(bb, tv, V, bl);
    const bool hit = pair_prefilter<SAME, RB, FIX>(S, na, bl, ok, p.mask, bitops, le, lo, sp);
    if (__any(hit) && hit) {
      pair_rewalk<word, SAME, RB>(
          S, na, a0, ok, p.mask, lo_e, lo_o, span, pair, tag, hits, hit_count, hit_cap,
          [&](int k, uint64_t &rank) -> word {
            const uint64_t x = x0 + threadIdx.x + (uint64_t)k * kPairThreads;
            const uint64_t ra = bb.trans ? x : y, rb = bb.trans ? y : x;
            rank = bb.start + bb.stride * (ra * bb.nB + rb);
            return desc_value<word>(bb, tv, V, ra, rb);
          });
    }
  }
}

// One tile of a pair check (shared by the per-instance and batched kernels).
// The A values of the tile must be staged in S (stage_a).  p.mask selects the
// operator instances (bit i: instance i, see HitRec); `tag` is or-ed into
// HitRec::inst above bit 8 (the batch's instance index).
template <typename word, bool SAME, int BSRC, int RB, typename PL, int FIX = -1>
__device__ __forceinline__ void pair_tile(const TableView *__restrict__ tv,
                                          const AStage &S, const word *__restrict__ A,
                                          const word *__restrict__ B, const PL &p,
                                          uint64_t at, uint64_t bt, bool bitops, word lo_e,
                                          word lo_o, word span, HitRec *__restrict__ hits,
                                          uint32_t *__restrict__ hit_count, uint32_t hit_cap,
                                          uint32_t tag) {
  const uint64_t a0 = at * p.ta;
  const uint64_t a1 = a0 + p.ta < p.nA ? a0 + p.ta : p.nA;
  const uint32_t na = (uint32_t)(a1 - a0);
  if (BSRC == kBBand ||
      (BSRC == kBDesc && (1u << p.bb.logTB) == (uint32_t)(kPairThreads * RB))) {
    desc_band_tile<word, SAME, RB, PL, FIX>(tv, S, p, na, a0, bt, bitops, lo_e, lo_o, span,
                                            hits, hit_count, hit_cap, tag);
    return;
  }
  uint32_t bl[RB];
  bool ok[RB];
  b_lo<word, BSRC, RB, PL>(tv, B, p, bt, bl, ok);
  const bool hit = pair_prefilter<SAME, RB, FIX>(S, na, bl, ok, p.mask, bitops, lo32(lo_e),
                                                 lo32(lo_o), (uint32_t)span);
  if (__any(hit) && hit) {
    const uint32_t pair = ((uint32_t)p.sb << 16) | (uint32_t)p.j;
    pair_rewalk<word, SAME, RB>(S, na, a0, ok, p.mask, lo_e, lo_o, span, pair, tag, hits,
                                hit_count, hit_cap, [&](int k, uint64_t &rank) -> word {
                                  return b_full<word, BSRC, RB, PL>(tv, B, p, bt, k, rank);
                                });
  }
}

// FIX: the stage mode when the launcher knows it (kStageWindow: every
// operator instance against windows, the bulk of a search), so the kernel
// holds that path only; -1: chosen per launch on the device.
template <typename word, bool SAME, int BSRC, int RB, int FIX>
__global__ void __launch_bounds__(kPairThreads)
k_pair_check(const TableView *__restrict__ tv, const word *__restrict__ A,
             const word *__restrict__ B, PairLaunch p, uint64_t lo_e64, uint64_t lo_o64,
             uint64_t span64, HitRec *__restrict__ hits, uint32_t *__restrict__ hit_count,
             uint32_t hit_cap) {
  __shared__ AStage S;
  const word lo_e = (word)lo_e64, lo_o = (word)lo_o64, span = (word)span64;
  const int mode = FIX >= 0 ? FIX : stage_mode(p.mask, SAME, (uint32_t)span);
  const uint32_t le = lo32(lo_e), lo = lo32(lo_o);
  uint64_t staged = ~0ull;
  bool bitops = false;
  for (uint64_t t = p.tile_begin + blockIdx.x; t < p.tile_end; t += gridDim.x) {
    const uint64_t at = p.nTA == 1 ? 0 : t % p.nTA, bt = p.nTA == 1 ? t : t / p.nTA;
    // an equal-size split as a triangle: a tile whose every A index exceeds
    // every B index holds only pairs (a, b), a > b, whose mirror (b, a) lies
    // in a kept tile, where all seven instances give both orders' trees
    if (p.tri && at * p.ta > bt * (uint64_t)(kPairThreads * RB) + (kPairThreads * RB - 1))
      continue;
    if (at != staged) {  // workgroup-uniform: every wave takes the same tiles
      const uint64_t a0 = at * p.ta;
      const uint64_t a1 = a0 + p.ta < p.nA ? a0 + p.ta : p.nA;
      bitops = stage_a<word>(S, tv, A, p.gen_a, p.j, a0, a1, mode, le, lo, (uint32_t)span);
      staged = at;
    }
    pair_tile<word, SAME, BSRC, RB, PairLaunch, FIX>(tv, S, A, B, p, at, bt, bitops, lo_e,
                                                     lo_o, span, hits, hit_count, hit_cap, 0u);
  }
}

// Leaves (size-1 trees) against the same windows: exact leaf solutions and
// every unary chain over a leaf.
template <typename word>
__global__ void k_leaf_scan(const TableView *__restrict__ tv, int L, uint64_t lo_e64,
                            uint64_t lo_o64, uint64_t span64, HitRec *hits,
                            uint32_t *hit_count, uint32_t hit_cap) {
  const word *V = global_ptr<const word>(tv->base) + tv->off[1];
  const int i = threadIdx.x;
  if (i < L) {
    const word c = V[i];
    if (in_window(c, (word)lo_e64, (word)lo_o64, (word)span64))
      record_hit<word>(hits, hit_count, hit_cap, (uint64_t)i, 0, c, 7, 1u << 16);
  }
}

// ---------------------------------------------------------------------------
// The reference's per-candidate path (numberToFormula + is_solution_rpn,
// nocache.cu:23-83 / mba.cuh:11-66) for one lane: unrank the tree of `size`
// and `rank` into RPN and evaluate it sample by sample, leaving at the first
// mismatch.  Leaf values [sample][leaf] and outputs are in LDS.
// ---------------------------------------------------------------------------
template <typename word>
__device__ bool tree_matches(const TableView *__restrict__ tv, const word *s_leaf,
                             const word *s_out, int L, int n_samples, int size, uint64_t rank) {
  uint8_t rpn[kMaxSize];
  unrank_rpn(tv, size, rank, rpn);
  for (int s = 0; s < n_samples; ++s) {
    word vs[kMaxSize];
    int top = -1;
    for (int i = 0; i < size; ++i) {
      const uint8_t t = rpn[i];
      if (t < kTokOp) {
        vs[++top] = s_leaf[s * L + t];
      } else {
        const int op = t - kTokOp;
        if (is_unary(op)) {
          vs[top] = apply_op<word>(op, vs[top], word(0));
        } else {
          const word rv = vs[top--];
          vs[top] = apply_op<word>(op, vs[top], rv);
        }
      }
    }
    if (vs[0] != s_out[s]) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Window-hit resolution on the device.  A hit X of size sb with filter-sample
// value x meets an even chain when t_s - x_s is the same offset k on every
// sample s (a chain of unary nodes then computes x + k), an odd chain when
// t_s + x_s is (-x + k); |k| <= U = max_size - sb.  For each parity, the
// shortest chain whose size lies in [min_size, max_size] and, among those,
// the smallest rank (chain_rank) give the hit's candidate (size, rank); a
// candidate that holds on every sample lowers found[inst][size].
// k_hit_screen (a lane per hit) keeps the candidates that also hold on a
// second sample, k_hit_confirm (a wave per candidate) checks every sample.
// ---------------------------------------------------------------------------
constexpr int kVerifyDepth = kMaxSize / 2 + 2;  // max RPN stack depth

template <typename word>
__device__ word eval_rpn_lds(const uint8_t *rpn, int size, const uint64_t *leaves,
                             word (*stk)[64], int lane) {
  int top = -1;
  for (int i = 0; i < size; ++i) {
    const uint8_t t = rpn[i];
    if (t < kTokOp) {
      stk[++top][lane] = (word)leaves[t];
    } else {
      const int op = t - kTokOp;
      if (is_unary(op)) {
        stk[top][lane] = apply_op<word>(op, stk[top][lane], word(0));
      } else {
        const word rv = stk[top--][lane];
        stk[top][lane] = apply_op<word>(op, stk[top][lane], rv);
      }
    }
  }
  return stk[0][lane];
}

__device__ __forceinline__ uint64_t op_offset_dev(const TableView *tv, int s, int op) {
  uint64_t r = 0;
  for (int o = 0; o < op; ++o) r += tv->T[s][o];
  return r;
}

__device__ __forceinline__ uint64_t split_offset_dev(const TableView *tv, int s, int j) {
  uint64_t r = 0;
  for (int jj = 1; jj < j; ++jj) r += tv->T[jj][8] * tv->T[s - 1 - jj][8];
  return r;
}

template <typename word>
__device__ __forceinline__ int64_t signed_of(word x) {
  return sizeof(word) == 4 ? (int64_t)(int32_t)(uint32_t)x : (int64_t)(uint64_t)x;
}

// Rank of the binary-rooted tree of a pair hit at its size sb (order of
// nocache.cu:52-79), or the leaf index of a leaf hit.
__device__ __forceinline__ uint64_t hit_base_rank(const TableView *tv, const HitRec &h) {
  const uint32_t inst = h.inst & 0xffu;
  const int sb = (int)(h.pair >> 16), j = (int)(h.pair & 0xffffu), m = sb - 1 - j;
  if (inst == 7) return h.a;
  if (inst == 6)  // (B - A): '-' with left child of size m
    return op_offset_dev(tv, sb, SUB) + split_offset_dev(tv, sb, m) + h.b * tv->T[j][8] + h.a;
  const int op = inst == 0 ? AND : inst == 1 ? OR : inst == 2 ? XOR
               : inst == 3 ? ADD : inst == 4 ? MUL : SUB;
  return op_offset_dev(tv, sb, op) + split_offset_dev(tv, sb, j) + h.a * tv->T[m][8] + h.b;
}

// Shortest chain of unary nodes over a tree of size sb mapping x to x + k
// (odd = false) or to -x + k (odd = true), of length u >= u_min with
// sb + u <= max_size, and its smallest rank increment (a '~' node adds 0 to
// its child's rank, a neg node at size s adds T[s][0..3], nocache.cu:52-63).
// Number the nodes 1 (innermost) .. u: the chain computes (-1)^u x + c with
// c = -sum over its '~' nodes i of (-1)^(u-i); all '~' gives c = 0 (u even)
// or -1 (u odd), and each '~' replaced by neg moves c by +1 at the positions
// i = u, u-2, ... and by -1 at i = u-1, u-3, ...  So |k| <= u/2 (even) and
// -(u+1)/2 <= k <= (u-1)/2 (odd) are reachable, and the cheapest choice
// replaces the innermost positions of the needed class (T[s] grows with s).
// Returns u, or -1 when no chain fits.
__device__ __forceinline__ int chain_rank(const TableView *tv, int sb, int64_t k, bool odd,
                                          int u_min, int max_size, uint64_t &add) {
  const int64_t ak = k >= 0 ? k : -k;
  const int64_t need = odd ? (k >= 0 ? 2 * k + 1 : 2 * ak - 1) : 2 * ak;
  const int64_t lo = odd ? (u_min | 1) : u_min + (u_min & 1);
  const int64_t u = need > lo ? need : lo;
  if (sb + u > max_size) return -1;
  const int64_t d = k - (odd ? -1 : 0);
  int i = ((d >= 0) == odd) ? 1 : 2;
  uint64_t c = 0;
  for (int64_t cnt = d >= 0 ? d : -d; cnt > 0; --cnt, i += 2)
    c += op_offset_dev(tv, sb + i, NEG);
  add = c;
  return (int)u;
}

// Copies *src into the workgroup's LDS copy (every lane takes part).
__device__ __forceinline__ void stage_view(TableView *dst, const TableView *src) {
  const uint4 *s = (const uint4 *)src;
  uint4 *d = (uint4 *)dst;
  for (uint32_t i = threadIdx.x; i < sizeof(TableView) / 16; i += blockDim.x) d[i] = s[i];
}

// Screen of recorded hits, one lane per hit: the filter-sample windows and
// one more sample.  Each parity that holds on both gives a candidate (size,
// rank) from the filter-sample offset alone; candidates that can still lower
// their slot are listed (hit index, slot | parity << 31, rank) for the check
// on every sample.  A wave whose hits share one instance reads its table view
// from LDS.
template <typename word>
__device__ void screen_one(const TableView *tv, const uint64_t *__restrict__ leafvals,
                           const uint64_t *__restrict__ outputs, const BatchInstData &d,
                           const HitRec &h, uint32_t hi, int min_size, int max_size,
                           const unsigned long long *__restrict__ found,
                           SurvRec *__restrict__ surv, uint32_t *__restrict__ n_surv) {
  const uint32_t I = h.inst >> 8, inst = h.inst & 0xffu;
  const int sb = (int)(h.pair >> 16), j = (int)(h.pair & 0xffffu), m = sb - 1 - j;
  const uint64_t *out = outputs + d.out_off;
  const word x0 = (word)h.v, t0 = (word)out[d.fs];
  const word ke = word(t0 - x0), ko = word(t0 + x0);
  const word wU = (word)(max_size - sb), span = word(2) * wU;
  bool even = word(ke + wU) <= span, odd = word(ko + wU) <= span;
  if (!even && !odd) return;
  if (d.n_samples > 1) {
    const uint32_t s2 = d.fs + 1 < d.n_samples ? d.fs + 1 : 0;
    const uint64_t *ls = leafvals + d.leaf_off + (size_t)s2 * tv->L;
    const word x = inst < 7 ? inst_value<word>((int)inst, eval_tree<word>(tv, j, h.a, ls),
                                                eval_tree<word>(tv, m, h.b, ls))
                            : (word)ls[h.a];
    const word t = (word)out[s2];
    even = even && word(t - x) == ke;
    odd = odd && word(t + x) == ko;
    if (!even && !odd) return;
  }
  const uint64_t rX = hit_base_rank(tv, h);
  const int u_min = min_size > sb ? min_size - sb : 0;
  for (int par = 0; par < 2; ++par) {
    if (!(par ? odd : even)) continue;
    uint64_t add = 0;
    const int u = chain_rank(tv, sb, signed_of<word>(par ? ko : ke), par != 0, u_min, max_size,
                             add);
    if (u < 0) continue;
    const uint32_t slot = I * (kMaxSize + 1) + (uint32_t)(sb + u);
    const uint64_t rank = rX + add;
    if (rank >= found[slot]) continue;
    SurvRec e;
    e.hit = hi;
    e.slot = slot | ((uint32_t)par << 31);
    e.rank = rank;
    surv[atomicAdd(n_surv, 1u)] = e;
  }
}

// Hits [0, min(*n_hits, cap)), 64 per wave, grid-stride.
template <typename word>
__global__ void __launch_bounds__(64)
k_hit_screen(const TableView *__restrict__ tvs, const uint64_t *__restrict__ leafvals,
             const uint64_t *__restrict__ outputs, const BatchInstData *__restrict__ idat,
             const HitRec *__restrict__ hits, const uint32_t *__restrict__ n_hits,
             uint32_t cap, int min_size, int max_size,
             const unsigned long long *__restrict__ found, SurvRec *__restrict__ surv,
             uint32_t *__restrict__ n_surv) {
  __shared__ __align__(16) TableView s_tv;
  const uint32_t n = min(*n_hits, cap);
  for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
    const uint32_t hi = base + threadIdx.x;
    const bool valid = hi < n;
    HitRec h{};
    if (valid) h = hits[hi];
    const uint32_t I = h.inst >> 8;
    const uint32_t I0 = __builtin_amdgcn_readfirstlane(I);  // lane 0 always holds a hit
    const bool shared = __ballot(valid && I != I0) == 0ull;  // wave-uniform
    if (shared) {
      stage_view(&s_tv, tvs + I0);
      __syncthreads();
    }
    if (valid)
      screen_one<word>(shared ? &s_tv : tvs + I, leafvals, outputs, idat[I], h, hi, min_size,
                       max_size, found, surv, n_surv);
    __syncthreads();  // s_tv is restaged by the next round
  }
}

// Check of the listed candidates on every sample, one wave per candidate
// (grid-stride over *n_surv): the instance's table view is staged in LDS,
// lanes 0 and 1 unrank the hit's two subtrees, each lane evaluates its samples
// (RPN stacks in LDS).  A candidate that holds lowers found[slot]; a slot
// lowered in this launch is listed once in `touched` (its flag is cleared by
// k_found_gather).  Candidates that can no longer lower their slot are
// skipped.
template <typename word>
__global__ void __launch_bounds__(64)
k_hit_confirm(const TableView *__restrict__ tvs, const uint64_t *__restrict__ leafvals,
              const uint64_t *__restrict__ outputs, const BatchInstData *__restrict__ idat,
              const HitRec *__restrict__ hits, const SurvRec *__restrict__ surv,
              const uint32_t *__restrict__ n_surv, unsigned long long *__restrict__ found,
              uint32_t *__restrict__ flags, uint32_t *__restrict__ touched,
              uint32_t *__restrict__ n_touched) {
  __shared__ __align__(16) TableView s_tv;
  __shared__ uint8_t s_ra[kMaxSize], s_rb[kMaxSize];
  __shared__ word s_stack[kVerifyDepth][64];
  const int lane = (int)threadIdx.x;
  const uint32_t ns = *n_surv;
  // a contiguous run of candidates per workgroup (consecutive candidates
  // mostly share an instance, whose table view then stays staged)
  const uint32_t per = (ns + gridDim.x - 1) / gridDim.x;
  const uint32_t e0 = blockIdx.x * per, e1 = min(ns, e0 + per);
  uint32_t staged = ~0u;
  for (uint32_t e = e0; e < e1; ++e) {  // workgroup-uniform
    const SurvRec c = surv[e];
    const uint32_t slot = c.slot & 0x7fffffffu;
    const bool odd = (c.slot >> 31) != 0u;
    const unsigned long long f = found[slot];
    if (__ballot(c.rank >= f) != 0ull) continue;
    const HitRec g = hits[c.hit];
    const uint32_t I = g.inst >> 8, inst = g.inst & 0xffu;
    const BatchInstData d = idat[I];
    if (I != staged) {
      stage_view(&s_tv, tvs + I);
      staged = I;
    }
    __syncthreads();
    const int L = s_tv.L;
    const int sb = (int)(g.pair >> 16), j = (int)(g.pair & 0xffffu), m = sb - 1 - j;
    const uint64_t *lv = leafvals + d.leaf_off, *out = outputs + d.out_off;
    const word x0 = (word)g.v, t0 = (word)out[d.fs];
    const word k = odd ? word(t0 + x0) : word(t0 - x0);
    // the two subtrees are unranked side by side (lanes 0 and 1)
    if (lane < 2 && inst < 7) unrank_rpn(&s_tv, lane ? m : j, lane ? g.b : g.a, lane ? s_rb : s_ra);
    __syncthreads();
    bool bad = false;
    for (uint32_t s = (uint32_t)lane; s < d.n_samples; s += 64) {
      const uint64_t *ls = lv + (size_t)s * L;
      word x;
      if (inst < 7) {
        const word xa = eval_rpn_lds<word>(s_ra, j, ls, s_stack, lane);
        const word xb = eval_rpn_lds<word>(s_rb, m, ls, s_stack, lane);
        x = inst_value<word>((int)inst, xa, xb);
      } else {
        x = (word)ls[g.a];
      }
      const word t = (word)out[s];
      bad = bad || (odd ? word(t + x) : word(t - x)) != k;
    }
    if (__ballot(bad) == 0ull && lane == 0) {
      atomicMin(found + slot, (unsigned long long)c.rank);
      if (atomicExch(flags + slot, 1u) == 0u) touched[atomicAdd(n_touched, 1u)] = slot;
    }
    __syncthreads();  // s_tv, s_ra, s_rb and the stacks are reused
  }
}

// Candidates (instance, size, minimum rank) of the slots k_hit_confirm
// listed, [0, *n_touched); clears their flags for the next launch.
__global__ void __launch_bounds__(256)
k_found_gather(const uint32_t *__restrict__ touched, const uint32_t *__restrict__ n_touched,
               const unsigned long long *__restrict__ found, uint32_t *__restrict__ flags,
               VerifyCand *__restrict__ out) {
  const uint32_t n = *n_touched;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t slot = touched[i];
    flags[slot] = 0u;
    VerifyCand c;
    c.inst = slot / (kMaxSize + 1);
    c.size = slot % (kMaxSize + 1);
    c.rank = found[slot];
    out[i] = c;
  }
}

// ---------------------------------------------------------------------------
// SIMBA-RTid order (nocache_rtid.cu:85-95): thread g of a (size, op) block
// of N trees checks rank offset + ((g * C) mod 2^64) mod N.  Threads
// [g0, g1) of the block; the smallest g whose tree matches every sample is
// kept with atomicMin.  The filter-sample value comes from the tables (value_of),
// the full check only runs on a match.
// ---------------------------------------------------------------------------
constexpr uint64_t kRtidC = 2246822507ull;

template <typename word>
__global__ void __launch_bounds__(256)
k_rtid_scan(const TableView *__restrict__ tv, const uint64_t *__restrict__ leafvals,
            const uint64_t *__restrict__ outputs, int n_samples, int size, uint64_t offset,
            uint64_t N, uint64_t g0, uint64_t g1, uint64_t t0,
            unsigned long long *__restrict__ best) {
  __shared__ word s_leaf[kMaxSamples * kMaxLeaves];
  __shared__ word s_out[kMaxSamples];
  const int L = tv->L;
  for (int i = threadIdx.x; i < n_samples * L; i += blockDim.x) s_leaf[i] = (word)leafvals[i];
  for (int i = threadIdx.x; i < n_samples; i += blockDim.x) s_out[i] = (word)outputs[i];
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = g0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < g1; g += stride) {
    const uint64_t r = offset + (g * kRtidC) % N;
    if (value_of<word>(tv, size, r) != (word)t0) continue;
    if (tree_matches<word>(tv, s_leaf, s_out, L, n_samples, size, r))
      atomicMin(best, (unsigned long long)g);
  }
}

// ---------------------------------------------------------------------------
// Batched search: one persistent kernel per size step over the jobs of every
// active instance.  Workgroups pull runs of up to 32 consecutive tiles from a
// device counter; a binary search over the jobs' first tiles locates a run's
// first job, and the run walks forward from there, keeping the job and its
// staged A tile while they do not change.  Every workgroup leaves once the
// counter passes the step's total, so the grid drains.
// ---------------------------------------------------------------------------
// A batch job's pair core with its B block read in place (the job's block
// in the step's block array; every lane of the workgroup reads the same one).
struct BatchPair : PairCore {
  const BBlock &bb;
  __device__ BatchPair(const PairCore &c, const BBlock &b) : PairCore(c), bb(b) {}
};

// The block of a job's B side that the last tile fell in: its B tiles
// [lo, hi) and index (workgroup-uniform), so consecutive tiles of a block
// skip the search.
struct BlockCursor {
  uint64_t lo = 0, hi = 0;
  int k = 0;
};

template <typename word, bool BLOCKS>
__device__ __forceinline__ void batch_pair(AStage &S, const TableView *tv, const BatchJob &J,
                           const BBlock *__restrict__ blocks, uint64_t lt, uint64_t &staged,
                           bool &bitops, BlockCursor &bc, HitRec *hits, uint32_t *hc,
                           uint32_t cap) {
  const word *base = global_ptr<const word>(tv->base);
  const uint64_t at = J.p.nTA == 1 ? 0 : lt % J.p.nTA;
  uint64_t bt = J.p.nTA == 1 ? lt : lt / J.p.nTA;
  // a B side above the tables runs through the blocks [bb, bb + nbb): the
  // last one whose first tile (tile_pre) is at most bt holds tile bt
  const BBlock *bp = blocks + J.bb;  // the table source reads no block
  if (BLOCKS && J.nbb > 1) {
    if (bt < bc.lo || bt >= bc.hi) {
      int lo = 0, hi = (int)J.nbb - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bp[mid].tile_pre <= bt) lo = mid;
        else hi = mid - 1;
      }
      bc.k = lo;
      bc.lo = bp[lo].tile_pre;
      bc.hi = lo + 1 < (int)J.nbb ? bp[lo + 1].tile_pre : ~0ull;
    }
    bp += bc.k;
    bt -= bc.lo;
  }
  const BatchPair p(J.p, *bp);
  const word *A = base + (p.gen_a ? 0 : tv->off[p.j]);
  const word *B = base + (p.gen_b ? 0 : tv->off[p.m]);
  const word lo_e = (word)J.lo_e, lo_o = (word)J.lo_o, span = (word)J.span;
  const uint32_t tag = J.inst << 8;
  if (at != staged) {  // workgroup-uniform; the caller resets `staged` per job
    const uint64_t a0 = at * p.ta;
    const uint64_t a1 = a0 + p.ta < p.nA ? a0 + p.ta : p.nA;
    bitops = stage_a<word>(S, tv, A, p.gen_a, p.j, a0, a1,
                           stage_mode(p.mask, p.j == p.m, (uint32_t)span), lo32(lo_e),
                           lo32(lo_o), (uint32_t)span);
    staged = at;
  }
  // the B side: its table (BLOCKS false), a block formed from tables or a
  // generic rank range (BLOCKS true)
#define BATCH_TILE(SAME_, BS)                                                              \
  pair_tile<word, SAME_, BS, kPairRB, BatchPair>(tv, S, A, B, p, at, bt, bitops, lo_e, lo_o, \
                                                 span, hits, hc, cap, tag)
  if (!BLOCKS) {
    // J.bsub consecutive B tiles per job tile; an equal-size split as a
    // triangle (PairCore::tri) skips the tiles wholly below the diagonal
    const uint32_t sub = J.bsub > 1 ? J.bsub : 1;
    const uint64_t b0 = bt * sub, tbw = (uint64_t)(kPairThreads * kPairRB);
    for (uint32_t q = 0; q < sub && b0 + q < p.nTB; ++q) {
      bt = b0 + q;
      if (p.tri && at * p.ta > bt * tbw + tbw - 1) continue;
      if (p.j == p.m && !p.tri) BATCH_TILE(true, kBTable);
      else BATCH_TILE(false, kBTable);
    }
  } else if (p.j == p.m) {
    if (p.bb.generic) BATCH_TILE(true, kBGeneric);
    else BATCH_TILE(true, kBDesc);
  } else {
    if (p.bb.generic) BATCH_TILE(false, kBGeneric);
    else BATCH_TILE(false, kBDesc);
  }
#undef BATCH_TILE
}

// At least 4 waves per SIMD: 128 registers hold every path of the kernel
// without spills, and the extra waves hide the small tiles' memory latency.
template <typename word, bool BLOCKS>
__global__ void __launch_bounds__(kPairThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_batch_step(const BatchStep *__restrict__ step, const BatchJob *__restrict__ jobs_base,
             const BBlock *__restrict__ blocks, const TableView *__restrict__ tvs,
             unsigned long long *head, HitRec *hits, uint32_t *hit_count, uint32_t hit_cap,
             const int32_t *__restrict__ settled) {
  __shared__ unsigned long long s_t;
  __shared__ AStage S;
  // this kernel's part of the step's tiles (see BatchStep), its own counter
  const uint64_t first = BLOCKS ? step->tiles_table : 0;
  const uint64_t total = BLOCKS ? step->total_tiles : step->tiles_table;
  if (BLOCKS) ++head;
  const int njobs = (int)step->njobs;
  const BatchJob *__restrict__ jobs = jobs_base + step->job_off;
  // tiles per counter grab: up to 32, fewer when the step has few tiles per
  // workgroup, so the tail stays short
  const uint64_t per = (total - first) / ((uint64_t)gridDim.x * 8u);
  const uint64_t K = per < 1 ? 1 : (per > 32 ? 32 : per);
  int jb = 0, jcur = -1;  // workgroup-uniform job cursor
  bool done = false;      // the job's instance is settled
  BlockCursor bc;         // the job's B block of the last tile
  BatchJob J;
  uint64_t staged = ~0ull;  // A tile of job jcur staged in S
  bool bitops = false;
  for (;;) {
    if (threadIdx.x == 0) s_t = first + atomicAdd(head, (unsigned long long)K);
    __syncthreads();
    const uint64_t t0 = s_t;
    __syncthreads();
    if (t0 >= total) break;
    const uint64_t t1 = t0 + K < total ? t0 + K : total;
    if (!(jobs[jb].tile0 <= t0 && (jb + 1 == njobs || jobs[jb + 1].tile0 > t0))) {
      int lo = 0, hi = njobs - 1;  // the last job whose first tile is <= t0
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= t0) lo = mid;
        else hi = mid - 1;
      }
      jb = lo;
    }
    for (uint64_t t = t0; t < t1; ++t) {
      while (jb + 1 < njobs && jobs[jb + 1].tile0 <= t) ++jb;
      if (jb != jcur) {
        J = jobs[jb];
        jcur = jb;
        staged = ~0ull;
        bc.lo = bc.hi = 0;
        // an instance settled at a smaller size needs no more work (the
        // settled sizes change only between steps)
        done = settled && settled[J.inst] > 0;
      }
      if (done) continue;
      const TableView *tv = tvs + J.inst;
      const uint64_t lt = t - J.tile0 + J.p.tile_begin;
      if (!BLOCKS && J.kind == 0) {  // table build: V[m] ranks [lt*kBuildTile, +kBuildTile)
        // The tile's first rank is located once (workgroup-uniform): its
        // (op, left size) block of V[m], the block's children and the
        // rank's row a0 and column b0 in it.  A lane's ranks inside that
        // block are rows and columns of it found with a compare (a block
        // at least kBuildTile wide) or a 32-bit division; ranks past the
        // block's end (a tile spanning blocks) decompose on their own.
        constexpr int kPer = kBuildTile / kPairThreads;
        const int m = J.p.m;
        const word *V = global_ptr<const word>(tv->base);
        word *out = global_ptr<word>(tv->base) + tv->off[m];
        const uint64_t rt = lt * kBuildTile;
        uint64_t x = rt, blk_end = 0, nB = 1, offA, offB;
        int op = 0;
        while (x >= tv->T[m][op]) x -= tv->T[m][op++];
        if (is_unary(op)) {
          offA = tv->off[m - 1];
          offB = offA;
          blk_end = rt - x + tv->T[m][op];
        } else {
          int j = 1;
          for (;; ++j) {
            const uint64_t blk = tv->T[j][8] * tv->T[m - 1 - j][8];
            if (x < blk) break;
            x -= blk;
          }
          nB = tv->T[m - 1 - j][8];
          offA = tv->off[j];
          offB = tv->off[m - 1 - j];
          blk_end = rt - x + tv->T[j][8] * nB;
        }
        const uint64_t a0 = uniform(x / nB), b0 = uniform(x - a0 * nB);
        uint64_t ia[kPer], ib[kPer];
        int opk[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
          const uint32_t d = threadIdx.x + (uint32_t)k * kPairThreads;
          const uint64_t r = rt + d;
          if (r >= J.p.nB) {
            opk[k] = -1;
            ia[k] = ib[k] = tv->off[1];
          } else if (r < blk_end) {
            const uint64_t t = b0 + d;
            uint64_t a, b;
            if (nB >= kBuildTile) {
              const bool wrap = t >= nB;
              a = a0 + (wrap ? 1 : 0);
              b = wrap ? t - nB : t;
            } else {
              const uint32_t q = (uint32_t)t / (uint32_t)nB;
              a = a0 + q;
              b = t - (uint64_t)q * nB;
            }
            opk[k] = op;
            ia[k] = offA + (is_unary(op) ? x + d : a);
            ib[k] = is_unary(op) ? ia[k] : offB + b;
          } else {
            opk[k] = build_refs(tv, m, r, ia[k], ib[k]);
          }
        }
        // every load, then the stores (the table being written is not read)
        word xv[kPer], yv[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) xv[k] = V[ia[k]], yv[k] = V[ib[k]];
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          if (opk[k] >= 0)
            out[rt + threadIdx.x + (uint64_t)k * kPairThreads] =
                apply_op<word>(opk[k], xv[k], yv[k]);
      } else if (J.kind == 1) {
        batch_pair<word, BLOCKS>(S, tv, J, blocks, lt, staged, bitops, bc, hits, hit_count,
                                 hit_cap);
      } else if (!BLOCKS) {  // leaves against the windows of base size 1
        const word *V = global_ptr<const word>(tv->base) + tv->off[1];
        if ((int)threadIdx.x < tv->L) {
          const word c = V[threadIdx.x];
          if (in_window(c, (word)J.lo_e, (word)J.lo_o, (word)J.span))
            record_hit<word>(hits, hit_count, hit_cap, threadIdx.x, 0, c, 7 | (J.inst << 8),
                             1u << 16);
        }
      }
    }
  }
}

// V[1] of every instance of a batch: the leaf values of its filter sample
// (workgroup i, lane l < L: tables_i[l] = leafvals[leaf_off_i + fs_i*L + l]).
template <typename word>
__global__ void k_batch_leaves(const TableView *__restrict__ tvs,
                               const uint64_t *__restrict__ leafvals,
                               const BatchInstData *__restrict__ idat, uint32_t count) {
  const uint32_t i = blockIdx.x;
  if (i >= count) return;
  const TableView &tv = tvs[i];
  const uint32_t l = threadIdx.x;
  if (tv.Kt > 0 && l < (uint32_t)tv.L) {  // Kt 0: an instance without tables
    const BatchInstData &d = idat[i];
    ((word *)tv.base)[tv.off[1] + l] = (word)leafvals[d.leaf_off + (uint64_t)d.fs * tv.L + l];
  }
}

__global__ void __launch_bounds__(256)
k_tv_update(TableView *__restrict__ tvs, const TvUpdate *__restrict__ up, uint32_t n) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const TvUpdate u = up[k];
    tvs[u.inst].off[u.m] = u.off;
    tvs[u.inst].Kt = u.kt;
  }
}

hipError_t launch_tv_update(TableView *d_tvs, const TvUpdate *d_up, uint32_t n,
                            hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_tv_update, dim3(blocks < 1024 ? blocks : 1024), dim3(256), 0, stream,
                     d_tvs, d_up, n);
  return hipGetLastError();
}

__global__ void k_batch_reset(unsigned long long *head, BatchCounters *cnt) {
  if (threadIdx.x == 0) {
    head[0] = 0;
    head[1] = 0;
    cnt->hits = 0;
    cnt->surv = 0;
    cnt->touched = 0;
  }
}

__global__ void __launch_bounds__(256)
k_batch_settle(const BatchStep *__restrict__ step, const unsigned long long *__restrict__ found,
               int32_t *__restrict__ settled, uint32_t count, int min_size,
               BatchCounters *cnt, uint32_t hit_cap) {
  const int done = step->size_done;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t hi = step->inst_hi < count ? step->inst_hi : count;
  for (uint32_t i = step->inst_lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi;
       i += stride) {
    const int st = settled[i];
    if (st > 0) continue;
    // st = -z: the instance has swept the sizes below z only
    const int lim = st < 0 && -st - 1 < done ? -st - 1 : done;
    const unsigned long long *f = found + (size_t)i * (kMaxSize + 1);
    for (int z = min_size < 1 ? 1 : min_size; z <= lim; ++z) {
      if (f[z] != ~0ull) {
        settled[i] = z;
        break;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && cnt->hits > hit_cap) {
    cnt->overflow = cnt->hits > cnt->overflow ? cnt->hits : cnt->overflow;
    if (step->index < cnt->first_overflow) cnt->first_overflow = step->index;
  }
}


// ---------------------------------------------------------------------------
// Wavefront verification: one wave per candidate, one lane per sample.  Lane
// 0 unranks the candidate into RPN in LDS (the reference's numberToFormula
// order, nocache.cu:23-83); every lane then evaluates the RPN on its sample
// with an LDS value stack (mba.cuh:11-66 semantics), the samples' leaf values
// staged in LDS, and a wave ballot ends the check at the first chunk of
// samples with a mismatch.
// ---------------------------------------------------------------------------

template <typename word>
__device__ void verify_wave(const TableView *__restrict__ tv, const uint64_t *__restrict__ lv,
                            const uint64_t *__restrict__ outv, int n_samples, int size,
                            uint64_t rank, uint8_t *__restrict__ ok_out) {
  __shared__ word s_leaf[kMaxSamples * kMaxLeaves];
  __shared__ word s_out[kMaxSamples];
  __shared__ word s_stack[kVerifyDepth][64];
  __shared__ uint8_t s_rpn[kMaxSize];
  __shared__ uint64_t s_wr[kMaxSize + 1];
  __shared__ int s_ws[kMaxSize + 1], s_wp[kMaxSize + 1];
  const int lane = threadIdx.x;
  const int L = tv->L;
  for (int i = lane; i < n_samples * L; i += 64) s_leaf[i] = (word)lv[i];
  for (int i = lane; i < n_samples; i += 64) s_out[i] = (word)outv[i];
  if (lane == 0) {
    int sp = 0;
    s_wr[0] = rank;
    s_ws[0] = size;
    s_wp[0] = 0;
    sp = 1;
    while (sp > 0) {
      --sp;
      uint64_t n = s_wr[sp];
      const int s = s_ws[sp], pos = s_wp[sp];
      if (s == 1) {
        s_rpn[pos] = (uint8_t)n;
        continue;
      }
      int op = 0;
      while (n >= tv->T[s][op]) n -= tv->T[s][op++];
      const int cs = s - 1;
      s_rpn[pos + cs] = (uint8_t)(kTokOp + op);
      if (is_unary(op)) {
        s_wr[sp] = n, s_ws[sp] = cs, s_wp[sp] = pos, ++sp;
        continue;
      }
      int j = 1;
      for (;; ++j) {
        const uint64_t blk = tv->T[j][8] * tv->T[cs - j][8];
        if (n < blk) break;
        n -= blk;
      }
      const uint64_t rs = tv->T[cs - j][8];
      const uint64_t a = n / rs;
      s_wr[sp] = a, s_ws[sp] = j, s_wp[sp] = pos, ++sp;
      s_wr[sp] = n - a * rs, s_ws[sp] = cs - j, s_wp[sp] = pos + j, ++sp;
    }
  }
  __syncthreads();
  bool good = true;
  for (int base = 0; base < n_samples; base += 64) {
    const int smp = base + lane;
    bool mismatch = false;
    if (smp < n_samples) {
      int top = -1;
      for (int i = 0; i < size; ++i) {
        const uint8_t t = s_rpn[i];
        if (t < kTokOp) {
          s_stack[++top][lane] = s_leaf[smp * L + t];
        } else {
          const int op = t - kTokOp;
          if (is_unary(op)) {
            s_stack[top][lane] = apply_op<word>(op, s_stack[top][lane], word(0));
          } else {
            const word rv = s_stack[top--][lane];
            s_stack[top][lane] = apply_op<word>(op, s_stack[top][lane], rv);
          }
        }
      }
      mismatch = s_stack[0][lane] != s_out[smp];
    }
    if (__any(mismatch)) {
      good = false;
      break;
    }
  }
  if (lane == 0) *ok_out = good ? 1 : 0;
}


template <typename word>
__global__ void __launch_bounds__(64)
k_verify_batch_wave(const TableView *__restrict__ tvs, const uint64_t *__restrict__ leafvals,
                    const uint64_t *__restrict__ outputs,
                    const BatchInstData *__restrict__ idat,
                    const VerifyCand *__restrict__ cand, const uint32_t *__restrict__ n_dev,
                    uint8_t *__restrict__ ok) {
  if (n_dev && blockIdx.x >= *n_dev) return;  // device-side candidate count
  const VerifyCand vc = cand[blockIdx.x];
  const BatchInstData d = idat[vc.inst];
  verify_wave<word>(tvs + vc.inst, leafvals + d.leaf_off, outputs + d.out_off,
                    (int)d.n_samples, (int)vc.size, vc.rank, ok + blockIdx.x);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
template <typename word>
static void table_blocks(word *base, const uint64_t *off, const CountTable &T, int m,
                         int max_grid, hipStream_t stream) {
  uint64_t r = 0;
  for (int op = 0; op < 8; ++op) {
    const bool un = is_unary(op);
    const int last = un ? 1 : (op == SUB ? m - 2 : (m - 1) / 2);
    for (int j = 1; j <= last; ++j) {
      const int jl = un ? m - 1 : j, jr = un ? 0 : m - 1 - j;
      const uint64_t nA = T.t[jl][8], nB = un ? 1 : T.t[jr][8];
      if (nA * nB == 0) continue;
      int logTB = 0;
      while ((1ull << logTB) < nB && logTB < 10) ++logTB;
      const uint64_t nBT = (nB + (1ull << logTB) - 1) >> logTB;
      const uint64_t nAT = (nA + (1024 >> logTB) - 1) / (1024 >> logTB);
      const uint64_t ntiles = nAT * nBT;
      const int grid = (int)(ntiles < (uint64_t)max_grid ? ntiles : (uint64_t)max_grid);
      hipLaunchKernelGGL(k_table_block<word>, dim3(grid), dim3(256), 0, stream,
                         base + off[jl], un ? base : base + off[jr], base + off[m] + r, op, nA,
                         nB, logTB, nBT, ntiles);
      r += nA * nB;
    }
  }
}

hipError_t launch_table_small(int width, const TableView *d_tv, int m_lo, int m_hi,
                              hipStream_t stream) {
  if (width == 32)
    hipLaunchKernelGGL(k_table_small<uint32_t>, dim3(1), dim3(1024), 0, stream, d_tv, m_lo, m_hi);
  else
    hipLaunchKernelGGL(k_table_small<uint64_t>, dim3(1), dim3(1024), 0, stream, d_tv, m_lo, m_hi);
  return hipGetLastError();
}

hipError_t launch_table_build(int width, void *d_base, const uint64_t *off, const CountTable &T,
                              int m, hipStream_t stream) {
  if (width == 32)
    table_blocks<uint32_t>((uint32_t *)d_base, off, T, m, 8192, stream);
  else
    table_blocks<uint64_t>((uint64_t *)d_base, off, T, m, 8192, stream);
  return hipGetLastError();
}

template <typename word>
static void pair_dispatch(const TableView *d_tv, const word *A, const word *B,
                          const PairLaunch &p, const CheckTargets &tg, HitRec *h,
                          uint32_t *hc, uint32_t cap, int grid, hipStream_t s) {
  const bool same = p.j == p.m;
  // the folded window layout (stage_mode): every instance, a span above 0
  const uint32_t full = same && !p.tri ? 0x3fu : 0x7fu;
  const bool win = (p.mask & full) == full && tg.span != 0;
#define PAIR_LAUNCH_F(S, BS, F)                                                          \
  do {                                                                                   \
    if (p.rb == 8)                                                                       \
      hipLaunchKernelGGL((k_pair_check<word, S, BS, 8, F>), dim3(grid), dim3(kPairThreads), \
                         0, s, d_tv, A, B, p, tg.lo_even, tg.lo_odd, tg.span, h, hc, cap);  \
    else                                                                                 \
      hipLaunchKernelGGL((k_pair_check<word, S, BS, 4, F>), dim3(grid), dim3(kPairThreads), \
                         0, s, d_tv, A, B, p, tg.lo_even, tg.lo_odd, tg.span, h, hc, cap);  \
  } while (0)
#define PAIR_LAUNCH(S, BS)                   \
  do {                                       \
    if (win) PAIR_LAUNCH_F(S, BS, kStageWindow); \
    else PAIR_LAUNCH_F(S, BS, -1);           \
  } while (0)
  // the B side: its table, a block formed from tables, or a generic rank range
  const int bs = !p.gen_b ? kBTable : p.bb.generic ? kBGeneric
                 : (1u << p.bb.logTB) == (uint32_t)(kPairThreads * p.rb) ? kBBand : kBGrid;
  if (same && p.tri) {
    PAIR_LAUNCH(false, kBTable);
  } else if (same) {
    if (bs == kBTable) PAIR_LAUNCH(true, kBTable);
    else if (bs == kBBand) PAIR_LAUNCH(true, kBBand);
    else if (bs == kBGrid) PAIR_LAUNCH(true, kBGrid);
    else PAIR_LAUNCH(true, kBGeneric);
  } else {
    if (bs == kBTable) PAIR_LAUNCH(false, kBTable);
    else if (bs == kBBand) PAIR_LAUNCH(false, kBBand);
    else if (bs == kBGrid) PAIR_LAUNCH(false, kBGrid);
    else PAIR_LAUNCH(false, kBGeneric);
  }
#undef PAIR_LAUNCH
#undef PAIR_LAUNCH_F
}

hipError_t launch_pair_check(int width, const TableView *d_tv, const void *d_base,
                             const uint64_t *off, const PairLaunch &p,
                             const CheckTargets &tg, HitRec *d_hits,
                             uint32_t *d_hit_count, uint32_t hit_cap, int grid,
                             hipStream_t stream) {
  // the A tiles must cover [0, nA) as the kernel maps them
  if (p.ta == 0 || p.ta > (uint32_t)kPairTA || p.nTA != (p.nA + p.ta - 1) / p.ta)
    return hipErrorInvalidValue;
  // sides above Kt have no table: their pointer is unused by the kernel
  if (width == 32) {
    const uint32_t *base = (const uint32_t *)d_base;
    pair_dispatch<uint32_t>(d_tv, base + (p.gen_a ? 0 : off[p.j]), base + (p.gen_b ? 0 : off[p.m]),
                            p, tg, d_hits, d_hit_count, hit_cap, grid, stream);
  } else {
    const uint64_t *base = (const uint64_t *)d_base;
    pair_dispatch<uint64_t>(d_tv, base + (p.gen_a ? 0 : off[p.j]), base + (p.gen_b ? 0 : off[p.m]),
                            p, tg, d_hits, d_hit_count, hit_cap, grid, stream);
  }
  return hipGetLastError();
}

hipError_t launch_rtid_scan(int width, const TableView *d_tv, const uint64_t *d_leafvals,
                            const uint64_t *d_outputs, int n_samples, int size,
                            uint64_t offset, uint64_t N, uint64_t g0, uint64_t g1,
                            uint64_t t0, unsigned long long *d_best, int grid,
                            hipStream_t stream) {
  if (width == 32)
    hipLaunchKernelGGL(k_rtid_scan<uint32_t>, dim3(grid), dim3(256), 0, stream, d_tv,
                       d_leafvals, d_outputs, n_samples, size, offset, N, g0, g1, t0, d_best);
  else
    hipLaunchKernelGGL(k_rtid_scan<uint64_t>, dim3(grid), dim3(256), 0, stream, d_tv,
                       d_leafvals, d_outputs, n_samples, size, offset, N, g0, g1, t0, d_best);
  return hipGetLastError();
}

hipError_t launch_hit_resolve(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                              const uint64_t *d_outputs, const BatchInstData *d_idat,
                              const HitRec *d_hits, const uint32_t *d_n_hits, uint32_t cap,
                              uint32_t screen_blocks, uint32_t confirm_blocks, int min_size,
                              int max_size, unsigned long long *d_found, SurvRec *d_surv,
                              uint32_t *d_n_surv, uint32_t *d_flags, uint32_t *d_touched,
                              uint32_t *d_n_touched, hipStream_t stream) {
  const dim3 gs(screen_blocks ? screen_blocks : 1), gc(confirm_blocks ? confirm_blocks : 1);
  if (width == 32) {
    hipLaunchKernelGGL(k_hit_screen<uint32_t>, gs, dim3(64), 0, stream, d_tvs, d_leafvals,
                       d_outputs, d_idat, d_hits, d_n_hits, cap, min_size, max_size, d_found,
                       d_surv, d_n_surv);
    hipLaunchKernelGGL(k_hit_confirm<uint32_t>, gc, dim3(64), 0, stream, d_tvs, d_leafvals,
                       d_outputs, d_idat, d_hits, d_surv, d_n_surv, d_found, d_flags, d_touched,
                       d_n_touched);
  } else {
    hipLaunchKernelGGL(k_hit_screen<uint64_t>, gs, dim3(64), 0, stream, d_tvs, d_leafvals,
                       d_outputs, d_idat, d_hits, d_n_hits, cap, min_size, max_size, d_found,
                       d_surv, d_n_surv);
    hipLaunchKernelGGL(k_hit_confirm<uint64_t>, gc, dim3(64), 0, stream, d_tvs, d_leafvals,
                       d_outputs, d_idat, d_hits, d_surv, d_n_surv, d_found, d_flags, d_touched,
                       d_n_touched);
  }
  return hipGetLastError();
}

hipError_t launch_found_gather(const uint32_t *d_touched, const uint32_t *d_n, uint32_t max_n,
                               const unsigned long long *d_found, uint32_t *d_flags,
                               VerifyCand *d_out, hipStream_t stream) {
  if (max_n == 0) return hipSuccess;
  const uint32_t blocks = (max_n + 255) / 256;
  hipLaunchKernelGGL(k_found_gather, dim3(blocks < 1024u ? blocks : 1024u), dim3(256), 0, stream,
                     d_touched, d_n, d_found, d_flags, d_out);
  return hipGetLastError();
}

hipError_t launch_leaf_scan(int width, const TableView *d_tv, int L,
                            const CheckTargets &tg, HitRec *d_hits,
                            uint32_t *d_hit_count, uint32_t hit_cap,
                            hipStream_t stream) {
  if (width == 32)
    hipLaunchKernelGGL(k_leaf_scan<uint32_t>, dim3(1), dim3(64), 0, stream, d_tv, L,
                       tg.lo_even, tg.lo_odd, tg.span, d_hits, d_hit_count, hit_cap);
  else
    hipLaunchKernelGGL(k_leaf_scan<uint64_t>, dim3(1), dim3(64), 0, stream, d_tv, L,
                       tg.lo_even, tg.lo_odd, tg.span, d_hits, d_hit_count, hit_cap);
  return hipGetLastError();
}

hipError_t launch_batch_leaves(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                               const BatchInstData *d_idat, uint32_t count,
                               hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (width == 32)
    hipLaunchKernelGGL(k_batch_leaves<uint32_t>, dim3(count), dim3(64), 0, stream, d_tvs,
                       d_leafvals, d_idat, count);
  else
    hipLaunchKernelGGL(k_batch_leaves<uint64_t>, dim3(count), dim3(64), 0, stream, d_tvs,
                       d_leafvals, d_idat, count);
  return hipGetLastError();
}

hipError_t launch_batch_reset(unsigned long long *d_head, BatchCounters *d_cnt,
                              hipStream_t stream) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, d_head, d_cnt);
  return hipGetLastError();
}

hipError_t launch_batch_settle(const BatchStep *d_step, const unsigned long long *d_found,
                               int32_t *d_settled, uint32_t count, int min_size,
                               BatchCounters *d_cnt, uint32_t hit_cap, hipStream_t stream) {
  const uint32_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_batch_settle, dim3(blocks ? blocks : 1), dim3(256), 0, stream, d_step,
                     d_found, d_settled, count, min_size, d_cnt, hit_cap);
  return hipGetLastError();
}

hipError_t launch_batch_step(int width, const BatchStep *d_step, const BatchJob *d_jobs_base,
                             const BBlock *d_blocks, const TableView *d_tvs,
                             unsigned long long *d_head, HitRec *d_hits,
                             uint32_t *d_hit_count, uint32_t hit_cap,
                             const int32_t *d_settled, int grid, hipStream_t stream) {
  // the tabulated part, then the part formed from blocks (an empty part's
  // kernel returns at once)
  if (width == 32) {
    hipLaunchKernelGGL((k_batch_step<uint32_t, false>), dim3(grid), dim3(kPairThreads), 0,
                       stream, d_step, d_jobs_base, d_blocks, d_tvs, d_head, d_hits,
                       d_hit_count, hit_cap, d_settled);
    hipLaunchKernelGGL((k_batch_step<uint32_t, true>), dim3(grid), dim3(kPairThreads), 0,
                       stream, d_step, d_jobs_base, d_blocks, d_tvs, d_head, d_hits,
                       d_hit_count, hit_cap, d_settled);
  } else {
    hipLaunchKernelGGL((k_batch_step<uint64_t, false>), dim3(grid), dim3(kPairThreads), 0,
                       stream, d_step, d_jobs_base, d_blocks, d_tvs, d_head, d_hits,
                       d_hit_count, hit_cap, d_settled);
    hipLaunchKernelGGL((k_batch_step<uint64_t, true>), dim3(grid), dim3(kPairThreads), 0,
                       stream, d_step, d_jobs_base, d_blocks, d_tvs, d_head, d_hits,
                       d_hit_count, hit_cap, d_settled);
  }
  return hipGetLastError();
}

hipError_t launch_verify_batch(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                               const uint64_t *d_outputs, const BatchInstData *d_idat,
                               const VerifyCand *d_cand, uint32_t n_cand, uint8_t *d_ok,
                               hipStream_t stream, const uint32_t *d_n) {
  if (n_cand == 0) return hipSuccess;
  if (width == 32)
    hipLaunchKernelGGL(k_verify_batch_wave<uint32_t>, dim3(n_cand), dim3(64), 0, stream, d_tvs,
                       d_leafvals, d_outputs, d_idat, d_cand, d_n, d_ok);
  else
    hipLaunchKernelGGL(k_verify_batch_wave<uint64_t>, dim3(n_cand), dim3(64), 0, stream, d_tvs,
                       d_leafvals, d_outputs, d_idat, d_cand, d_n, d_ok);
  return hipGetLastError();
}


}  // namespace mba
