// simba_kernels.hip -- gfx950 kernels of the SIMBA (cache-free, unranking)
// search.
//
// The reference (nocache.cu:85-107, mba.cuh:6-69) gives one thread one rank:
// unrank to RPN in local memory, then a divergent stack evaluation.  Here a
// size level is instead swept as outer products.  Every binary-rooted tree of
// size s is op(X, Y) with X of size j and Y of size m = s-1-j; for an
// unordered split {j, m} the seven operator instances & | ^ + * (X-Y) (Y-X)
// are fused and evaluated on sample 0 from the value tables V[j], V[m]
// (value of every tree of that size on sample 0, in rank order).  A lane
// holds RB (4 or 8) values of the B side in VGPRs; the A side streams through
// wave-uniform scalar loads, so the inner loop is pure integer VALU work with
// no divergence.  Unary-rooted trees are the composition of a chain of ~/neg
// with a binary-rooted tree or a leaf; such a chain maps x to +-x + k with
// |k| <= chain length, so two windowed compares per value cover every unary
// tree of every larger size.  Hits on sample 0 are rare; they are appended to
// a hit buffer, resolved to (size, rank) by the host and verified on all
// samples by launch_verify.
#include "simba_kernels.h"

namespace mba {

template <typename word>
struct WordOf;

// ---------------------------------------------------------------------------
// Table-assisted value of the tree (m, r) on sample 0: unrank top-down in the
// reference's order, but stop at any subtree of size <= Kt and read its value
// from V[size]; then evaluate the remaining skeleton in postfix order.
// ---------------------------------------------------------------------------
template <typename word>
__device__ __forceinline__ word build_value(const TableView *__restrict__ tv, int m,
                                            uint64_t r);

template <typename word>
__device__ word value_of(const TableView *__restrict__ tv, int m, uint64_t r) {
  const word *V = (const word *)tv->base;
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
    const uint64_t a = n / rs;
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
// Table build: V[m][r] for r in [0, T[m]).  One lane per rank; the op block
// and split are located by prefix sums over the count table (the reference's
// unrank order, nocache.cu:52-79), the children are read from V[<m].
// ---------------------------------------------------------------------------
template <typename word>
__device__ __forceinline__ word build_value(const TableView *__restrict__ tv, int m,
                                            uint64_t r) {
  const word *V = (const word *)tv->base;
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
// Inner loop (per A value, per B value held in VGPRs): the seven operator
// results are folded into one running minimum of their distances to the two
// windows, so each value costs its operator plus two subtractions and one
// 3-way minimum; one compare per tile decides whether a lane saw a hit.
// For 64-bit words only the high halves are formed and compared: a value in
// a window of width span < 2^32 has hi(c) - hi(lo) in {0, 1} (mod 2^32), so
// this is a necessary condition, and the rare lanes that pass it re-walk the
// tile with the exact 64-bit window test.
// ---------------------------------------------------------------------------
// Per-B-value precomputation, hoisted out of the A loop: the window offsets
// are folded into B so that the additive instances need one add per window.
//   64-bit words: offsets are multiples of 2^32 (He = hi(lo_even) << 32),
//   which leave the low halves and carries unchanged, so hi(a + be) equals
//   hi(a + b) - hi(lo_even) exactly.
template <typename word>
struct BPre {
  word b, be, bo, nbe, nbo;  // b, b-Oe, b-Oo, -b-Oe, -b-Oo
};

template <typename word>
__device__ __forceinline__ word win_off(word lo);
template <>
__device__ __forceinline__ uint32_t win_off<uint32_t>(uint32_t lo) { return lo; }
template <>
__device__ __forceinline__ uint64_t win_off<uint64_t>(uint64_t lo) {
  return lo & 0xffffffff00000000ull;
}

template <typename word>
__device__ __forceinline__ BPre<word> bpre(word b, word oe, word oo) {
  BPre<word> p;
  p.b = b;
  p.be = b - oe;
  p.bo = b - oo;
  p.nbe = word(0) - b - oe;
  p.nbo = word(0) - b - oo;
  return p;
}

// The distance that is compared: the whole word for 32-bit words, the high
// half for 64-bit words.
__device__ __forceinline__ uint32_t dpart(uint32_t x) { return x; }
__device__ __forceinline__ uint32_t dpart(uint64_t x) { return (uint32_t)(x >> 32); }

__device__ __forceinline__ uint32_t min3u(uint32_t a, uint32_t b, uint32_t c) {
  return min(min(a, b), c);
}

template <typename word, bool SAME>
__device__ __forceinline__ uint32_t fold7(uint32_t acc, word a, word na, const BPre<word> &p,
                                          uint32_t de, uint32_t dof) {
  const uint32_t x_and = dpart(word(a & p.b));
  const uint32_t x_or = dpart(word(a | p.b));
  const uint32_t x_xor = dpart(word(a ^ p.b));
  const uint32_t x_mul = dpart(word(a * p.b));
  acc = min3u(acc, x_and - de, x_and - dof);
  acc = min3u(acc, x_or - de, x_or - dof);
  acc = min3u(acc, x_xor - de, x_xor - dof);
  acc = min3u(acc, x_mul - de, x_mul - dof);
  acc = min3u(acc, dpart(word(a + p.be)), dpart(word(a + p.bo)));    // a + b
  acc = min3u(acc, dpart(word(a + p.nbe)), dpart(word(a + p.nbo)));  // a - b
  if (!SAME) acc = min3u(acc, dpart(word(p.be + na)), dpart(word(p.bo + na)));  // b - a
  return acc;
}

// Moves a wave-uniform word into scalar registers, so that expressions such
// as b + (-a) keep their single-add form.
__device__ __forceinline__ uint32_t uniform(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}
__device__ __forceinline__ uint64_t uniform(uint64_t x) {
  return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
         __builtin_amdgcn_readfirstlane((uint32_t)x);
}

// Values of V[m] for m = Kt + 1 (children tabulated) at the B indices of a
// tile: the tile's first index `first` is decoded once (op block, split,
// left/right rank -- the same for every lane), and each lane steps from it:
// inside one (op, split) block whose right side has at least a tile's worth
// of trees, the lane's right rank is first's plus its offset with at most
// one carry into the left rank.  Lanes past that block decode on their own.
template <typename word, int RB>
__device__ __forceinline__ void gen_b_tile(const TableView *__restrict__ tv, int m,
                                           uint64_t first, uint64_t lane_b0, word *bv,
                                           const bool *ok) {
  const word *V = (const word *)tv->base;
  uint64_t x = first;
  int op = 0;
  while (x >= tv->T[m][op]) x -= tv->T[m][op++];
  const bool un = is_unary(op);
  int j = 0;
  uint64_t rs = 1, a0 = 0, r0 = 0, rem;
  if (un) {
    rem = tv->T[m][op] - x;
  } else {
    const int cs = m - 1;
    j = 1;
    for (;; ++j) {
      const uint64_t blk = tv->T[j][8] * tv->T[cs - j][8];
      if (x < blk) {
        rem = blk - x;
        break;
      }
      x -= blk;
    }
    rs = tv->T[cs - j][8];
    a0 = x / rs;
    r0 = x - a0 * rs;
  }
  const bool step_ok = un || rs >= (uint64_t)(kPairThreads * RB);
  const word *VA = V + tv->off[un ? m - 1 : j];
  const word *VB = V + tv->off[un ? 0 : m - 1 - j];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const uint64_t b = lane_b0 + (uint64_t)k * kPairThreads;
    if (!ok[k]) {
      bv[k] = word(0);
      continue;
    }
    const uint64_t off = b - first;
    if (step_ok && off < rem) {
      if (un) {
        bv[k] = apply_op<word>(op, VA[x + off], word(0));
      } else {
        uint64_t rb = r0 + off, ra = a0;
        if (rb >= rs) {
          rb -= rs;
          ra += 1;
        }
        bv[k] = apply_op<word>(op, VA[ra], VB[rb]);
      }
    } else {
      bv[k] = build_value<word>(tv, m, b);
    }
  }
}

// One operator instance over the A tile: its distances folded into acc.
template <typename word, int INST, bool GEN_A, int RB>
__device__ __forceinline__ uint32_t fold_inst(uint32_t acc, const TableView *__restrict__ tv,
                                              const word *__restrict__ A, int j, uint64_t a0,
                                              uint64_t a1, const BPre<word> *pre, uint32_t de,
                                              uint32_t dof) {
  for (uint64_t a = a0; a < a1; ++a) {
    const word av = GEN_A ? value_of<word>(tv, j, a) : A[a];
    const word na = uniform(word(word(0) - av));
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const BPre<word> &q = pre[k];
      if (INST <= 2 || INST == 4) {
        const word c = INST == 0 ? word(av & q.b) : INST == 1 ? word(av | q.b)
                     : INST == 2 ? word(av ^ q.b) : word(av * q.b);
        const uint32_t x = dpart(c);
        acc = min3u(acc, x - de, x - dof);
      } else if (INST == 3) {
        acc = min3u(acc, dpart(word(av + q.be)), dpart(word(av + q.bo)));
      } else if (INST == 5) {
        acc = min3u(acc, dpart(word(av + q.nbe)), dpart(word(av + q.nbo)));
      } else {
        acc = min3u(acc, dpart(word(q.be + na)), dpart(word(q.bo + na)));
      }
    }
  }
  return acc;
}

// One tile of a pair check (shared by the per-instance and batched kernels).
// p.mask selects the operator instances (bit i: instance i, see HitRec).
// `tag` is or-ed into HitRec::inst above bit 8 (the batch's instance index).
template <typename word, bool SAME, bool GEN_A, bool GEN_B, int RB>
__device__ __forceinline__ void pair_tile(const TableView *__restrict__ tv,
                                          const word *__restrict__ A,
                                          const word *__restrict__ B, const PairLaunch &p,
                                          uint64_t t, word lo_e, word lo_o, word span,
                                          HitRec *__restrict__ hits,
                                          uint32_t *__restrict__ hit_count, uint32_t hit_cap,
                                          uint32_t tag) {
  // prefilter threshold on the folded distance
  const uint32_t thr = sizeof(word) == 8 ? 1u : (uint32_t)span;
  const uint32_t pair = ((uint32_t)p.sb << 16) | (uint32_t)p.j;
  {
    const uint64_t at = p.nTA == 1 ? 0 : t % p.nTA, bt = p.nTA == 1 ? t : t / p.nTA;
    const uint64_t a0 = at * kPairTA;
    const uint64_t a1 = a0 + kPairTA < p.nA ? a0 + kPairTA : p.nA;
    const uint64_t b0 = bt * (kPairThreads * RB) + threadIdx.x;
    word bv[RB];
    bool ok[RB];
    bool any_ok = false;
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const uint64_t b = b0 + (uint64_t)k * kPairThreads;
      ok[k] = b < p.nB;
      any_ok |= ok[k];
      if (!GEN_B) bv[k] = ok[k] ? B[b] : word(0);
    }
    if (GEN_B) {
      if (p.m == tv->Kt + 1) {
        gen_b_tile<word, RB>(tv, p.m, bt * (kPairThreads * RB), b0, bv, ok);
      } else {
#pragma unroll
        for (int k = 0; k < RB; ++k)
          bv[k] = ok[k] ? value_of<word>(tv, p.m, b0 + (uint64_t)k * kPairThreads) : word(0);
      }
    }
    // lanes past the end of B reuse a valid value of their own (recorded
    // hits are re-derived with the exact masks below)
#pragma unroll
    for (int k = 1; k < RB; ++k)
      if (!ok[k]) bv[k] = bv[0];
    const word oe = win_off<word>(lo_e), oo = win_off<word>(lo_o);
    const uint32_t de = dpart(oe), dof = dpart(oo);
    BPre<word> pre[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) pre[k] = bpre<word>(bv[k], oe, oo);
    uint32_t acc = 0xffffffffu;
    const uint32_t full = SAME ? 0x3fu : 0x7fu;
    if ((p.mask & full) == full) {
      uint64_t a = a0;
      // A values in groups of kAGroup: their scalar loads are issued together
      // and waited for once
      for (; a + kAGroup <= a1; a += kAGroup) {
        word av[kAGroup];
#pragma unroll
        for (int i = 0; i < kAGroup; ++i)
          av[i] = GEN_A ? value_of<word>(tv, p.j, a + i) : A[a + i];
#pragma unroll
        for (int i = 0; i < kAGroup; ++i) {
          const word na = uniform(word(word(0) - av[i]));
#pragma unroll
          for (int k = 0; k < RB; ++k)
            acc = fold7<word, SAME>(acc, av[i], na, pre[k], de, dof);
        }
      }
      for (; a < a1; ++a) {
        const word av = GEN_A ? value_of<word>(tv, p.j, a) : A[a];
        const word na = uniform(word(word(0) - av));
#pragma unroll
        for (int k = 0; k < RB; ++k) acc = fold7<word, SAME>(acc, av, na, pre[k], de, dof);
      }
    } else {
      // a subset of the operator instances (one block at a time): one pass
      // over the A tile per selected instance
      for (int inst = 0; inst < 7; ++inst) {
        if (!((p.mask & full) >> inst & 1u)) continue;
        switch (inst) {
          case 0: acc = fold_inst<word, 0, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
          case 1: acc = fold_inst<word, 1, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
          case 2: acc = fold_inst<word, 2, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
          case 3: acc = fold_inst<word, 3, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
          case 4: acc = fold_inst<word, 4, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
          case 5: acc = fold_inst<word, 5, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
          default: acc = fold_inst<word, 6, GEN_A, RB>(acc, tv, A, p.j, a0, a1, pre, de, dof); break;
        }
      }
    }
    const bool hit = any_ok && acc <= thr;
    // Rare: some lane passed the prefilter.  Re-walk the tile, exact test.
    if (__any(hit)) {
      if (hit) {
        for (uint64_t a = a0; a < a1; ++a) {
          const word av = GEN_A ? value_of<word>(tv, p.j, a) : A[a];
          for (int k = 0; k < RB; ++k) {
            if (!ok[k]) continue;
            const uint64_t b = b0 + (uint64_t)k * kPairThreads;
            for (int inst = 0; inst < (SAME ? 6 : 7); ++inst) {
              if (!(p.mask >> inst & 1u)) continue;
              const word c = inst_value<word>(inst, av, bv[k]);
              if (in_window(c, lo_e, lo_o, span))
                record_hit<word>(hits, hit_count, hit_cap, a, b, c, inst | tag, pair);
            }
          }
        }
      }
    }
  }
}

template <typename word, bool SAME, bool GEN_A, bool GEN_B, int RB>
__global__ void __launch_bounds__(kPairThreads)
k_pair_check(const TableView *__restrict__ tv, const word *__restrict__ A,
             const word *__restrict__ B, PairLaunch p, uint64_t lo_e64, uint64_t lo_o64,
             uint64_t span64, HitRec *__restrict__ hits, uint32_t *__restrict__ hit_count,
             uint32_t hit_cap) {
  if (!GEN_A && p.nTA == 1) {
    // the whole A side fits one tile: stage it in LDS once per workgroup
    // and reuse it for every B tile this workgroup takes
    __shared__ word sA[kPairTA];
    for (int i = threadIdx.x; i < (int)p.nA; i += blockDim.x) sA[i] = A[i];
    __syncthreads();
    for (uint64_t t = p.tile_begin + blockIdx.x; t < p.tile_end; t += gridDim.x)
      pair_tile<word, SAME, GEN_A, GEN_B, RB>(tv, sA, B, p, t, (word)lo_e64, (word)lo_o64,
                                          (word)span64, hits, hit_count, hit_cap, 0u);
    return;
  }
  for (uint64_t t = p.tile_begin + blockIdx.x; t < p.tile_end; t += gridDim.x)
    pair_tile<word, SAME, GEN_A, GEN_B, RB>(tv, A, B, p, t, (word)lo_e64, (word)lo_o64,
                                        (word)span64, hits, hit_count, hit_cap, 0u);
}

// Leaves (size-1 trees) against the same windows: exact leaf solutions and
// every unary chain over a leaf.
template <typename word>
__global__ void k_leaf_scan(const TableView *__restrict__ tv, int L, uint64_t lo_e64,
                            uint64_t lo_o64, uint64_t span64, HitRec *hits,
                            uint32_t *hit_count, uint32_t hit_cap) {
  const word *V = (const word *)tv->base + tv->off[1];
  const int i = threadIdx.x;
  if (i < L) {
    const word c = V[i];
    if (in_window(c, (word)lo_e64, (word)lo_o64, (word)span64))
      record_hit<word>(hits, hit_count, hit_cap, (uint64_t)i, 0, c, 7, 1u << 16);
  }
}

// ---------------------------------------------------------------------------
// Verification: the reference's per-candidate path (numberToFormula +
// is_solution_rpn, nocache.cu:23-83 / mba.cuh:11-66): unrank the full tree
// and evaluate it sample by sample, leaving at the first mismatch.  Leaf
// values of all samples are staged in LDS.
// ---------------------------------------------------------------------------
template <typename word>
__global__ void __launch_bounds__(256)
k_verify(const TableView *__restrict__ tv, const uint64_t *__restrict__ leafvals,
         const uint64_t *__restrict__ outputs, int n_samples,
         const uint32_t *__restrict__ sizes, const uint64_t *__restrict__ ranks,
         uint32_t n_cand, uint8_t *__restrict__ ok) {
  __shared__ word s_leaf[kMaxSamples * kMaxLeaves];
  __shared__ word s_out[kMaxSamples];
  const int L = tv->L;
  for (int i = threadIdx.x; i < n_samples * L; i += blockDim.x)
    s_leaf[i] = (word)leafvals[i];
  for (int i = threadIdx.x; i < n_samples; i += blockDim.x) s_out[i] = (word)outputs[i];
  __syncthreads();
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  const int size = (int)sizes[c];
  uint8_t rpn[kMaxSize];
  struct Item { uint64_t r; int s, pos; };
  Item st[kMaxSize];
  int sp = 0;
  st[sp++] = {ranks[c], size, 0};
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
    const uint64_t a = n / rs;
    st[sp++] = {a, j, it.pos};
    st[sp++] = {n - a * rs, cs - j, it.pos + j};
  }
  uint8_t good = 1;
  for (int s = 0; s < n_samples && good; ++s) {
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
    good = vs[0] == s_out[s];
  }
  ok[c] = good;
}

// ---------------------------------------------------------------------------
// Batched search: one persistent kernel per size step over the jobs of every
// active instance.  Workgroups pull tile indices from a device counter; a
// binary search over the jobs' first tiles locates the job.  Every workgroup
// leaves once the counter passes the step's total, so the grid drains.
// ---------------------------------------------------------------------------
template <typename word>
__device__ void batch_pair(const TableView *tv, const BatchJob &J, uint64_t lt,
                           HitRec *hits, uint32_t *hc, uint32_t cap) {
  const word *base = (const word *)tv->base;
  const PairLaunch &p = J.p;
  const word *A = base + (p.gen_a ? 0 : tv->off[p.j]);
  const word *B = base + (p.gen_b ? 0 : tv->off[p.m]);
  const word lo_e = (word)J.lo_e, lo_o = (word)J.lo_o, span = (word)J.span;
  const uint32_t tag = J.inst << 8;
  if (p.j == p.m) {
    if (p.gen_a)
      pair_tile<word, true, true, true, kPairRB>(tv, A, B, p, lt, lo_e, lo_o, span, hits, hc, cap, tag);
    else
      pair_tile<word, true, false, false, kPairRB>(tv, A, B, p, lt, lo_e, lo_o, span, hits, hc, cap, tag);
  } else if (p.gen_a) {
    pair_tile<word, false, true, true, kPairRB>(tv, A, B, p, lt, lo_e, lo_o, span, hits, hc, cap, tag);
  } else if (p.gen_b) {
    pair_tile<word, false, false, true, kPairRB>(tv, A, B, p, lt, lo_e, lo_o, span, hits, hc, cap, tag);
  } else {
    pair_tile<word, false, false, false, kPairRB>(tv, A, B, p, lt, lo_e, lo_o, span, hits, hc, cap, tag);
  }
}

template <typename word>
__global__ void __launch_bounds__(kPairThreads)
k_batch_step(const BatchStep *__restrict__ step, const BatchJob *__restrict__ jobs,
             const TableView *__restrict__ tvs, unsigned long long *head, HitRec *hits,
             uint32_t *hit_count, uint32_t hit_cap) {
  __shared__ unsigned long long s_t;
  __shared__ int s_job;
  const uint64_t total = step->total_tiles;
  const int njobs = (int)step->njobs;
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned long long t = atomicAdd(head, 1ull);
      int lo = 0, hi = njobs - 1;
      if (t < total) {
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (jobs[mid].tile0 <= t) lo = mid;
          else hi = mid - 1;
        }
      }
      s_t = t;
      s_job = lo;
    }
    __syncthreads();
    const uint64_t t = s_t;
    const int jb = s_job;
    __syncthreads();
    if (t >= total) break;
    const BatchJob J = jobs[jb];
    const TableView *tv = tvs + J.inst;
    const uint64_t lt = t - J.tile0;
    if (J.kind == 0) {  // table build: V[m] ranks [lt*kBuildTile, +kBuildTile)
      word *out = (word *)tv->base + tv->off[J.p.m];
      const uint64_t r0 = lt * kBuildTile;
      const uint64_t r1 = r0 + kBuildTile < J.p.nB ? r0 + kBuildTile : J.p.nB;
      for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x)
        out[r] = build_value<word>(tv, J.p.m, r);
    } else if (J.kind == 1) {
      batch_pair<word>(tv, J, lt, hits, hit_count, hit_cap);
    } else {  // leaves against the windows of base size 1
      const word *V = (const word *)tv->base + tv->off[1];
      if ((int)threadIdx.x < tv->L) {
        const word c = V[threadIdx.x];
        if (in_window(c, (word)J.lo_e, (word)J.lo_o, (word)J.span))
          record_hit<word>(hits, hit_count, hit_cap, threadIdx.x, 0, c, 7 | (J.inst << 8),
                           1u << 16);
      }
    }
  }
}

__global__ void k_batch_reset(unsigned long long *head, uint32_t *hit_count) {
  if (threadIdx.x == 0) {
    *head = 0;
    *hit_count = 0;
  }
}

// Verification of candidates of many instances (global-memory samples).
template <typename word>
__global__ void __launch_bounds__(256)
k_verify_batch(const TableView *__restrict__ tvs, const uint64_t *__restrict__ leafvals,
               const uint64_t *__restrict__ outputs, const BatchInstData *__restrict__ idat,
               const VerifyCand *__restrict__ cand, uint32_t n_cand, uint8_t *__restrict__ ok) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  const VerifyCand vc = cand[c];
  const TableView *tv = tvs + vc.inst;
  const BatchInstData d = idat[vc.inst];
  const int L = tv->L;
  const int size = (int)vc.size;
  uint8_t rpn[kMaxSize];
  struct Item { uint64_t r; int s, pos; };
  Item st[kMaxSize];
  int sp = 0;
  st[sp++] = {vc.rank, size, 0};
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
    const uint64_t a = n / rs;
    st[sp++] = {a, j, it.pos};
    st[sp++] = {n - a * rs, cs - j, it.pos + j};
  }
  uint8_t good = 1;
  for (int s = 0; s < (int)d.n_samples && good; ++s) {
    word vs[kMaxSize];
    int top = -1;
    for (int i = 0; i < size; ++i) {
      const uint8_t t = rpn[i];
      if (t < kTokOp) {
        vs[++top] = (word)leafvals[d.leaf_off + (uint64_t)s * L + t];
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
    good = vs[0] == (word)outputs[d.out_off + s];
  }
  ok[c] = good;
}

// ---------------------------------------------------------------------------
// Wavefront verification: one wave per candidate, one lane per sample.  Lane
// 0 unranks the candidate into RPN in LDS (the reference's numberToFormula
// order, nocache.cu:23-83); every lane then evaluates the RPN on its sample
// with an LDS value stack (mba.cuh:11-66 semantics), the samples' leaf values
// staged in LDS, and a wave ballot ends the check at the first chunk of
// samples with a mismatch.
// ---------------------------------------------------------------------------
constexpr int kVerifyDepth = kMaxSize / 2 + 2;  // max RPN stack depth

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
k_verify_wave(const TableView *__restrict__ tv, const uint64_t *__restrict__ leafvals,
              const uint64_t *__restrict__ outputs, int n_samples,
              const uint32_t *__restrict__ sizes, const uint64_t *__restrict__ ranks,
              uint8_t *__restrict__ ok) {
  const uint32_t c = blockIdx.x;
  verify_wave<word>(tv, leafvals, outputs, n_samples, (int)sizes[c], ranks[c], ok + c);
}

template <typename word>
__global__ void __launch_bounds__(64)
k_verify_batch_wave(const TableView *__restrict__ tvs, const uint64_t *__restrict__ leafvals,
                    const uint64_t *__restrict__ outputs,
                    const BatchInstData *__restrict__ idat,
                    const VerifyCand *__restrict__ cand, uint8_t *__restrict__ ok) {
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
#define PAIR_LAUNCH(S, GA, GB)                                                       \
  do {                                                                               \
    if (p.rb == 8)                                                                   \
      hipLaunchKernelGGL((k_pair_check<word, S, GA, GB, 8>), dim3(grid), dim3(kPairThreads), \
                         0, s, d_tv, A, B, p, tg.lo_even, tg.lo_odd, tg.span, h, hc, cap); \
    else                                                                             \
      hipLaunchKernelGGL((k_pair_check<word, S, GA, GB, 4>), dim3(grid), dim3(kPairThreads), \
                         0, s, d_tv, A, B, p, tg.lo_even, tg.lo_odd, tg.span, h, hc, cap); \
  } while (0)
  if (same) {
    if (p.gen_a) PAIR_LAUNCH(true, true, true);
    else PAIR_LAUNCH(true, false, false);
  } else {
    if (p.gen_a) PAIR_LAUNCH(false, true, true);
    else if (p.gen_b) PAIR_LAUNCH(false, false, true);
    else PAIR_LAUNCH(false, false, false);
  }
#undef PAIR_LAUNCH
}

hipError_t launch_pair_check(int width, const TableView *d_tv, const void *d_base,
                             const uint64_t *off, const PairLaunch &p,
                             const CheckTargets &tg, HitRec *d_hits,
                             uint32_t *d_hit_count, uint32_t hit_cap, int grid,
                             hipStream_t stream) {
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

hipError_t launch_batch_reset(unsigned long long *d_head, uint32_t *d_hit_count,
                              hipStream_t stream) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, d_head, d_hit_count);
  return hipGetLastError();
}

hipError_t launch_batch_step(int width, const BatchStep *d_step, const BatchJob *d_jobs,
                             const TableView *d_tvs, unsigned long long *d_head,
                             HitRec *d_hits, uint32_t *d_hit_count, uint32_t hit_cap,
                             int grid, hipStream_t stream) {
  if (width == 32)
    hipLaunchKernelGGL(k_batch_step<uint32_t>, dim3(grid), dim3(kPairThreads), 0, stream,
                       d_step, d_jobs, d_tvs, d_head, d_hits, d_hit_count, hit_cap);
  else
    hipLaunchKernelGGL(k_batch_step<uint64_t>, dim3(grid), dim3(kPairThreads), 0, stream,
                       d_step, d_jobs, d_tvs, d_head, d_hits, d_hit_count, hit_cap);
  return hipGetLastError();
}

hipError_t launch_verify_batch(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                               const uint64_t *d_outputs, const BatchInstData *d_idat,
                               const VerifyCand *d_cand, uint32_t n_cand, uint8_t *d_ok,
                               hipStream_t stream) {
  if (n_cand == 0) return hipSuccess;
  if (width == 32)
    hipLaunchKernelGGL(k_verify_batch_wave<uint32_t>, dim3(n_cand), dim3(64), 0, stream, d_tvs,
                       d_leafvals, d_outputs, d_idat, d_cand, d_ok);
  else
    hipLaunchKernelGGL(k_verify_batch_wave<uint64_t>, dim3(n_cand), dim3(64), 0, stream, d_tvs,
                       d_leafvals, d_outputs, d_idat, d_cand, d_ok);
  return hipGetLastError();
}

hipError_t launch_verify(int width, const TableView *d_tv, const uint64_t *d_leafvals,
                         const uint64_t *d_outputs, int n_samples,
                         const uint32_t *d_sizes, const uint64_t *d_ranks,
                         uint32_t n_cand, uint8_t *d_ok, hipStream_t stream) {
  if (n_cand == 0) return hipSuccess;
  if (width == 32)
    hipLaunchKernelGGL(k_verify_wave<uint32_t>, dim3(n_cand), dim3(64), 0, stream, d_tv,
                       d_leafvals, d_outputs, n_samples, d_sizes, d_ranks, d_ok);
  else
    hipLaunchKernelGGL(k_verify_wave<uint64_t>, dim3(n_cand), dim3(64), 0, stream, d_tv,
                       d_leafvals, d_outputs, n_samples, d_sizes, d_ranks, d_ok);
  return hipGetLastError();
}

}  // namespace mba
