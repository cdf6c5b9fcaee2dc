// simba_host.cpp -- host driver of the SIMBA search on one gfx950 device.
//
// Replaces NoCache::solve (simba/src/nocache.cu:197-289).  Sizes are visited
// in increasing order; the first solution is the minimum rank, in the
// reference's enumeration order, within the smallest size that has one (the
// reference keeps whichever thread wins atomicCAS inside the first
// (size, op) launch with a match, nocache.cu:102-105, which is one of the
// trees of that size and op block).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "context.h"
#include "simba_host.h"

namespace mba {

constexpr uint32_t kHitCap = 1u << 20;
constexpr uint64_t kDefaultTableBytes = 40ull << 30;
// Work per launch, in A x B value pairs, so that budgets are honoured.
constexpr uint64_t kPairsPerLaunch = 1ull << 36;

// Tile geometry of a B block of nA x nB values (see BBlock): columns run
// over the larger dimension, 2^logTB of them, and the tile's remaining
// (256*rb) >> logTB lanes take successive rows.
static BBlock shape_block(BBlock b, int rb) {
  const uint64_t tile = (uint64_t)kPairThreads * rb;
  b.trans = b.nB >= b.nA ? 0 : 1;
  const uint64_t nX = b.trans ? b.nA : b.nB;
  int lg = 0;
  while ((1ull << lg) < nX && (1ull << lg) < tile) ++lg;
  b.logTB = lg;
  b.ntc = (nX + (1ull << lg) - 1) >> lg;
  return b;
}

static uint64_t block_tiles(const BBlock &b, int rb) {
  const uint64_t tile = (uint64_t)kPairThreads * rb;
  if (b.generic) return (b.nA + tile - 1) / tile;
  const uint64_t nY = b.trans ? b.nB : b.nA;
  const uint64_t rows = tile >> b.logTB;
  return ((nY + rows - 1) / rows) * b.ntc;
}

// The B-side trees of size m > Kt in rank order, as blocks (see BBlock).
// `start` and `stride` map this subtree's rank r to the B side's rank
// start + stride * r; `steps` is the path above it (innermost first).  A
// binary node with both children tabulated is a 2-D block; one whose child
// above Kt has a sibling of at most kFixedSiblings trees is split into one
// path step per sibling tree (that tree's value fixed); anything else is a
// generic block.
constexpr uint64_t kFixedSiblings = 64;

static void b_blocks(const CountTable &T, int Kt, const uint64_t *off, int m, uint64_t start,
                     uint64_t stride, const std::vector<BStep> &steps, int rb,
                     std::vector<BBlock> &out) {
  auto emit = [&](BBlock b, const std::vector<BStep> &st) {
    b.stride = stride;
    if ((int)st.size() > kMaxBSteps && !b.generic) {
      // too deep a path: fall back to the generic evaluation of the range
      BBlock g{};
      g.start = b.start;
      g.stride = stride;
      g.nA = b.nA * b.nB;
      g.generic = 1;
      out.push_back(g);
      return;
    }
    b.nsteps = b.generic ? 0 : (int)st.size();
    for (int i = 0; i < b.nsteps; ++i) b.steps[i] = st[i];
    out.push_back(b.generic ? b : shape_block(b, rb));
  };
  auto with = [&](BStep first) {
    std::vector<BStep> v;
    v.reserve(steps.size() + 1);
    // merge consecutive unary nodes into one affine map: s*(s'*x + k') + k
    if (first.kind == 0 && !steps.empty() && steps[0].kind == 0) {
      const BStep &o = steps[0];
      BStep mrg;
      mrg.kind = 0;
      mrg.op = first.op ^ o.op;
      mrg.c = (o.op ? 0 - first.c : first.c) + o.c;
      v.push_back(mrg);
      v.insert(v.end(), steps.begin() + 1, steps.end());
    } else {
      v.push_back(first);
      v.insert(v.end(), steps.begin(), steps.end());
    }
    return v;
  };
  uint64_t r = 0;  // rank offset within size m
  for (int op = 0; op < 8; ++op) {
    const uint64_t cnt = T.t[m][op];
    if (cnt == 0) continue;
    if (is_unary(op)) {
      // ~x = -x - 1, neg x = -x
      const BStep u{0, 1, op == NOT ? ~0ull : 0ull};
      const std::vector<BStep> st = with(u);
      if (m - 1 <= Kt) {
        BBlock b{};
        b.start = start + stride * r;
        b.offA = off[m - 1];
        b.nA = T.t[m - 1][8];
        b.nB = 1;
        b.op = -1;
        emit(b, st);
      } else {
        b_blocks(T, Kt, off, m - 1, start + stride * r, stride, st, rb, out);
      }
    } else {
      const int cs = m - 1;
      const int last = op == SUB ? cs - 1 : cs / 2;
      uint64_t q = r;
      for (int j = 1; j <= last; ++j) {
        const int l = j, rr = cs - j;
        const uint64_t nl = T.t[l][8], nr = T.t[rr][8];
        const uint64_t base = start + stride * q;
        if (l <= Kt && rr <= Kt) {
          BBlock b{};
          b.start = base;
          b.offA = off[l];
          b.offB = off[rr];
          b.nA = nl;
          b.nB = nr;
          b.op = op;
          emit(b, steps);
        } else if (l > Kt && rr <= Kt && nr <= kFixedSiblings) {
          // right child fixed to tree b: rank base + stride*(ra*nr + b)
          for (uint64_t bi = 0; bi < nr; ++bi)
            b_blocks(T, Kt, off, l, base + stride * bi, stride * nr,
                     with(BStep{2, op, off[rr] + bi}), rb, out);
        } else if (rr > Kt && l <= Kt && nl <= kFixedSiblings) {
          // left child fixed to tree a: rank base + stride*(a*nr + rb)
          for (uint64_t ai = 0; ai < nl; ++ai)
            b_blocks(T, Kt, off, rr, base + stride * ai * nr, stride,
                     with(BStep{1, op, off[l] + ai}), rb, out);
        } else {
          BBlock b{};
          b.start = base;
          b.nA = nl * nr;
          b.generic = 1;
          emit(b, steps);
        }
        q += nl * nr;
      }
    }
    r += cnt;
  }
}

static inline int64_t as_signed(uint64_t x, int W) {
  return W == 64 ? (int64_t)x : (int64_t)(int32_t)(uint32_t)x;
}

// Rank at size sb of the binary-rooted tree X = inst(A[a], B[b]) of the
// split {j, m} (see HitRec), or of leaf `a` when inst == 7.
uint64_t base_rank(const CountTable &T, int sb, int j, int m, const HitRec &h) {
  static const int kInstOp[7] = {AND, OR, XOR, ADD, MUL, SUB, SUB};
  if (h.inst == 7) return h.a;
  const int op = kInstOp[h.inst];
  if (h.inst == 6)  // (B - A): '-' with left child of size m
    return T.op_offset(sb, SUB) + T.split_offset(sb, m) + h.b * T.t[j][8] + h.a;
  return T.op_offset(sb, op) + T.split_offset(sb, j) + h.a * T.t[m][8] + h.b;
}

// Turns one window hit into candidate trees.  X has size sb and value v on
// sample 0.  A chain of u unary operators over X computes (-1)^u * x + k; the
// target is met by the even map x + (t0 - v) or the odd map -x + (t0 + v).
// For each map the smallest chain length whose size lies in [min, max] is
// found, and among chains of that length the one with the smallest rank:
// a unary node at size s adds op_offset(s, op) to the rank of its child
// (nocache.cu:52-63), 0 for '~' and T[s][0..3] for neg.
void resolve_hit(const CountTable &T, int W, uint64_t t0, int sb, uint64_t rX,
                 uint64_t v, int min_size, int max_size, std::vector<Candidate> &out) {
  const uint64_t M = wmask(W);
  const int U = max_size - sb;
  const int64_t ke = as_signed((t0 - v) & M, W);
  const int64_t ko = as_signed((t0 + v) & M, W);
  const bool want_e = ke >= -U && ke <= U, want_o = ko >= -U && ko <= U;
  if (!want_e && !want_o) return;
  // cost[k + off] = min rank increment of a chain of length i reaching k
  // (fixed arrays: this runs once per window hit)
  const int off = U + 2, width = 2 * U + 5;
  const uint64_t INF = ~0ull;
  uint64_t cost_a[2 * kMaxSize + 8], next_a[2 * kMaxSize + 8];
  uint64_t *cost = cost_a, *next = next_a;
  std::fill(cost, cost + width, INF);
  cost[off] = 0;
  bool done_e = !want_e, done_o = !want_o;
  for (int i = 0; sb + i <= max_size && !(done_e && done_o); ++i) {
    const int size = sb + i;
    if (size >= min_size) {
      if (!done_e && (i % 2 == 0) && cost[ke + off] != INF) {
        out.push_back({size, rX + cost[ke + off]});
        done_e = true;
      }
      if (!done_o && (i % 2 == 1) && cost[ko + off] != INF) {
        out.push_back({size, rX + cost[ko + off]});
        done_o = true;
      }
    }
    if (size + 1 > max_size) break;
    std::fill(next, next + width, INF);
    const uint64_t neg_off = T.op_offset(size + 1, NEG);
    for (int k = -U - 1; k <= U + 1; ++k) {
      const uint64_t c = cost[k + off];
      if (c == INF) continue;
      const int kn = -k - 1, kg = -k;  // '~': -x-1 ; neg: -x
      if (kn >= -U - 1 && kn <= U + 1) next[kn + off] = std::min(next[kn + off], c);
      if (kg >= -U - 1 && kg <= U + 1)
        next[kg + off] = std::min(next[kg + off], c + neg_off);
    }
    std::swap(cost, next);
  }
}

void set_second_sample(TableView &tv, const std::vector<uint64_t> &lv,
                       const std::vector<uint64_t> &outv, int L, int fs, int s2) {
  tv.t1 = outv[fs];
  tv.has2 = s2 >= 0;
  if (!tv.has2) return;
  for (int l = 0; l < L; ++l) tv.leaf2[l] = lv[(size_t)s2 * L + l];
  tv.t2 = outv[s2];
}

int pick_filter_sample(const std::vector<uint64_t> &lv, const std::vector<uint64_t> &outv,
                       int n, int L, int W, int skip) {
  const uint64_t M = wmask(W);
  auto weak = [&](uint64_t x) {
    const uint64_t d = std::min(x & M, (0 - x) & M);
    return d <= 2 || x == (1ull << (W - 1));
  };
  // an output with few or many set bits is the value of many trees (the
  // bitwise operators drift toward 0 and ~0): prefer balanced ones
  auto balanced = [&](uint64_t x) {
    const int pc = __builtin_popcountll(x & M);
    return pc >= W / 4 && pc <= 3 * W / 4;
  };
  int best = -1, best_score = -1;
  for (int s = 0; s < n; ++s) {
    if (s == skip) continue;
    int score = weak(outv[s]) ? 0 : 2 * L;
    if (balanced(outv[s])) score += 2 * L;
    for (int a = 0; a < L; ++a) {
      const uint64_t x = lv[(size_t)s * L + a];
      if (weak(x)) continue;
      bool distinct = true;
      for (int b = 0; b < a; ++b)
        if (lv[(size_t)s * L + b] == x || lv[(size_t)s * L + b] == ((0 - x) & M) ||
            lv[(size_t)s * L + b] == (~x & M))
          distinct = false;
      score += distinct ? 2 : 1;
    }
    if (score > best_score) best = s, best_score = score;
  }
  return best;
}

int choose_kt(const CountTable &T, int max_size, int W, uint64_t budget) {
  const uint64_t wb = W / 8;
  int kt = 1;
  uint64_t used = T.t[1][8] * wb;
  for (int m = 2; m <= std::max(1, max_size - 2); ++m) {
    const uint64_t add = T.t[m][8] * wb;
    if (used + add > budget) break;
    used += add;
    kt = m;
  }
  return kt;
}

std::vector<BlockRef> binary_blocks(const CountTable &T, int s) {
  static const int kOpInst[8] = {-1, 0, 1, 2, -1, 3, -1, 4};
  std::vector<BlockRef> out;
  for (int op = 1; op < 8; ++op) {
    if (op == NEG) continue;
    const int last = op == SUB ? s - 2 : (s - 1) / 2;
    uint64_t r = T.op_offset(s, op);
    for (int j = 1; j <= last; ++j) {
      const int m = s - 1 - j;
      BlockRef b;
      b.op = op;
      b.split = j;
      b.begin = r;
      r += T.t[j][8] * T.t[m][8];
      b.end = r;
      if (op != SUB) {
        b.pair_j = j;
        b.inst = kOpInst[op];
      } else if (j <= m) {
        b.pair_j = j;
        b.inst = 5;  // A - B with A of size j
      } else {
        b.pair_j = m;
        b.inst = 6;  // B - A with B of size j
      }
      out.push_back(b);
    }
  }
  return out;
}

CheckTargets targets_for(uint64_t t0, int sb, int max_size, int W) {
  const uint64_t M = wmask(W);
  const uint64_t U = (uint64_t)(max_size - sb);
  CheckTargets tg;
  tg.lo_even = (t0 - U) & M;
  tg.lo_odd = ((0 - t0) - U) & M;
  tg.span = 2 * U;
  return tg;
}

int root_split(const uint8_t *rpn, int size) {
  if (size <= 1 || is_unary(rpn[size - 1] - kTokOp)) return 0;
  // walk the right subtree back from the root
  int need = 1, i = size - 2;
  while (need > 0) {
    const uint8_t t = rpn[i--];
    if (t < kTokOp) need -= 1;
    else if (!is_unary(t - kTokOp)) need += 1;
  }
  return i + 1;
}

void fill_result(const CountTable &T, int size, uint64_t rank, int n_vars,
                 const uint64_t *consts, int W, mba_result &res) {
  uint8_t rpn[kMaxSize];
  unrank_host(T, size, rank, rpn);
  res.status = MBA_FOUND;
  res.size = size;
  res.rank = rank;
  res.op = size > 1 ? rpn[size - 1] - kTokOp : -1;
  res.split = root_split(rpn, size);
  const std::string e = infix_simba(rpn, size, n_vars, consts, W);
  std::snprintf(res.expr, sizeof res.expr, "%s", e.c_str());
  const std::string r = rpn_chars(rpn, size);
  std::snprintf(res.rpn, sizeof res.rpn, "%s", r.c_str());
}

int simba_solve(Ctx &ctx, const mba_instance &inst, const mba_opts &o, mba_result &res,
                const float *block_scores, bool rtid) {
  const auto t_start = Clock::now();
  std::memset(&res, 0, sizeof res);
  res.status = MBA_NOT_FOUND;
  res.op = -1;
  const int W = (int)inst.width;
  const int nv = (int)inst.n_vars, nc = (int)inst.n_consts, n = (int)inst.n_samples;
  const int L = nv + nc;
  if ((W != 32 && W != 64) || L < 1 || L > kMaxLeaves || n < 1 || n > kMaxSamples ||
      !inst.inputs || !inst.outputs || (nc > 0 && !inst.consts)) {
    ctx.last_error = "instance outside supported limits";
    return MBA_ERR_LIMIT;
  }
  int min_size = o.min_size < 1 ? 1 : o.min_size;
  int max_size = o.max_size;
  if (o.compat) {  // reference: sizes 1..maxLen-1, size 1 holds no candidates
    min_size = 2;
    max_size = o.max_size - 1;
  }
  if (max_size > kMaxSize) {
    ctx.last_error = "max size above 32";
    return MBA_ERR_LIMIT;
  }
  if (max_size < min_size) {
    std::snprintf(res.expr, sizeof res.expr, "Not found !");
    res.seconds = seconds_since(t_start);
    return MBA_NOT_FOUND;
  }
  CountTable T;
  if (!T.build(L, max_size)) {
    ctx.last_error = "count table overflows uint64 at this max size";
    return MBA_ERR_COUNT;
  }
  const uint64_t M = wmask(W);
  const size_t wb = (size_t)W / 8;
  std::vector<uint64_t> lv((size_t)n * L);
  for (int s = 0; s < n; ++s) {
    for (int v = 0; v < nv; ++v) lv[(size_t)s * L + v] = inst.inputs[(size_t)s * nv + v] & M;
    for (int c = 0; c < nc; ++c) lv[(size_t)s * L + nv + c] = inst.consts[c] & M;
  }
  std::vector<uint64_t> outv(n);
  for (int s = 0; s < n; ++s) outv[s] = inst.outputs[s] & M;
  // Filter sample: the tables and fused checks work on one sample; every
  // hit is verified on all samples, so any choice gives the same answer.  A
  // sample whose leaf values are distinct and far from 0/+-1 rejects the most
  // candidates, so it is preferred.
  const int fs = pick_filter_sample(lv, outv, n, L, W);
  const uint64_t t0 = outv[fs];

  // ---- value tables V[1..Kt] -------------------------------------------
  const uint64_t budget = o.table_bytes ? o.table_bytes : kDefaultTableBytes;
  const int Kt = std::max(1, choose_kt(T, max_size, W, budget) - std::max(0, o.table_delta));
  TableView tv;
  std::memset(&tv, 0, sizeof tv);
  uint64_t elems = 0;
  for (int m = 1; m <= Kt; ++m) {
    tv.off[m] = elems;
    elems += T.t[m][8];
  }
  std::memcpy(tv.T, T.t, sizeof tv.T);
  tv.Kt = Kt;
  tv.L = L;
  set_second_sample(tv, lv, outv, L, fs, pick_filter_sample(lv, outv, n, L, W, fs));
  MBA_HIP(ctx.tables.ensure(std::max<size_t>(elems * wb, 256)));
  tv.base = ctx.tables.p;
  MBA_HIP(ctx.tview.ensure(sizeof(TableView)));
  MBA_HIP(ctx.hits.ensure(sizeof(HitRec) * kHitCap));
  MBA_HIP(ctx.hit_count.ensure(256));
  MBA_HIP(ctx.leafvals.ensure(sizeof(uint64_t) * lv.size()));
  MBA_HIP(ctx.outputs.ensure(sizeof(uint64_t) * n));
  hipStream_t st = ctx.stream;
  TableView *d_tv = ctx.tview.as<TableView>();
  MBA_HIP(hipMemcpyAsync(d_tv, &tv, sizeof tv, hipMemcpyHostToDevice, st));
  MBA_HIP(hipMemcpyAsync(ctx.leafvals.p, lv.data(), sizeof(uint64_t) * lv.size(),
                         hipMemcpyHostToDevice, st));
  MBA_HIP(hipMemcpyAsync(ctx.outputs.p, outv.data(), sizeof(uint64_t) * n,
                         hipMemcpyHostToDevice, st));
  {  // V[1] = leaf values on sample 0
    std::vector<uint64_t> v1(L);
    std::vector<uint32_t> v1w(L);
    for (int l = 0; l < L; ++l)
      v1[l] = lv[(size_t)fs * L + l], v1w[l] = (uint32_t)lv[(size_t)fs * L + l];
    if (W == 64)
      MBA_HIP(hipMemcpyAsync(ctx.tables.p, v1.data(), 8 * L, hipMemcpyHostToDevice, st));
    else
      MBA_HIP(hipMemcpyAsync(ctx.tables.p, v1w.data(), 4 * L, hipMemcpyHostToDevice, st));
  }
  for (int m = 2; m <= Kt; ++m)
    MBA_HIP(launch_table_build(W, ctx.tables.p, tv.off, T, m, st));

  if (rtid) {
    // SIMBA-RTid: every (size, op) block in order, thread g of the block
    // checking rank ((g * C) mod 2^64) mod N; the first block with a match
    // decides, and within it the smallest such g (size 1 has no op block).
    constexpr uint64_t kC = 2246822507ull;
    constexpr uint64_t kChunk = 1ull << 32;
    MBA_HIP(ctx.hit_count.ensure(256));
    auto *d_best = (unsigned long long *)ctx.hit_count.p;
    const int grid_r = ctx.cu_count * 16;
    uint64_t cand = 0;
    int reached = 0;
    for (int s = std::max(min_size, 2); s <= max_size; ++s) {
      uint64_t offset = 0;
      for (int op = 0; op < 8; ++op) {
        const uint64_t N = T.t[s][op];
        if (N == 0) continue;
        const uint64_t G = (N + 1023) / 1024 * 1024;
        MBA_HIP(hipMemsetAsync(d_best, 0xff, 8, st));
        for (uint64_t g0 = 0; g0 < G; g0 += kChunk) {
          const uint64_t g1 = std::min(G, g0 + kChunk);
          MBA_HIP(launch_rtid_scan(W, d_tv, ctx.leafvals.as<uint64_t>(),
                                   ctx.outputs.as<uint64_t>(), n, s, offset, N, g0, g1, t0,
                                   d_best, grid_r, st));
          unsigned long long best = ~0ull;
          MBA_HIP(hipMemcpyAsync(&best, d_best, 8, hipMemcpyDeviceToHost, st));
          MBA_HIP(hipStreamSynchronize(st));
          cand += g1 - g0;
          if (best != ~0ull) {
            fill_result(T, s, offset + (best * kC) % N, nv, inst.consts, W, res);
            res.candidates = cand;
            res.size_reached = reached;
            res.seconds = seconds_since(t_start);
            return MBA_FOUND;
          }
          const bool over = (o.cancel && *o.cancel) ||
                            (o.time_budget_s > 0 && seconds_since(t_start) > o.time_budget_s);
          if (over) {
            std::snprintf(res.expr, sizeof res.expr, "Not found !");
            res.candidates = cand;
            res.size_reached = reached;
            res.seconds = seconds_since(t_start);
            return MBA_BUDGET;
          }
        }
        offset += N;
      }
      reached = s;
    }
    std::snprintf(res.expr, sizeof res.expr, "Not found !");
    res.candidates = cand;
    res.size_reached = reached;
    res.seconds = seconds_since(t_start);
    return MBA_NOT_FOUND;
  }

  // ---- size sweep --------------------------------------------------------
  std::vector<uint64_t> found(kMaxSize + 2, ~0ull);
  std::vector<HitRec> hrec;
  std::vector<Candidate> cands;
  uint32_t *d_hc = ctx.hit_count.as<uint32_t>();
  HitRec *d_hits = ctx.hits.as<HitRec>();
  // 32 workgroups per CU by default: more than stay resident, so tiles are
  // handed out in small grid-stride shares and no CU waits on a long tail
  const int grid = ctx.cu_count * (o.wg_per_cu > 0 ? o.wg_per_cu : 32);
  double ktime = 0.0;
  uint32_t klaunch = 0;
  int answer = -1;
  int reached = 0;
  int status = MBA_NOT_FOUND;

  // Resolve the hits of one step, verify the candidates on all samples.
  auto process_hits = [&](int sb_step) -> int {
    uint32_t cnt = 0;
    MBA_HIP(hipMemcpyAsync(&cnt, d_hc, 4, hipMemcpyDeviceToHost, st));
    MBA_HIP(hipStreamSynchronize(st));
    if (cnt == 0) return 0;
    if (cnt > kHitCap) {
      ctx.last_error = "sample-0 hit buffer overflow";
      return MBA_ERR_LIMIT;
    }
    const HitRec *src = d_hits;
    if (tv.has2) {  // keep the hits that also meet the second sample
      uint32_t *d_fc = (uint32_t *)(ctx.hit_count.as<char>() + 128);
      MBA_HIP(ctx.hits2.ensure(sizeof(HitRec) * cnt));
      MBA_HIP(hipMemsetAsync(d_fc, 0, 4, st));
      MBA_HIP(launch_hit_filter(W, d_tv, d_hits, cnt, max_size, ctx.hits2.as<HitRec>(), d_fc,
                                st));
      MBA_HIP(hipMemcpyAsync(&cnt, d_fc, 4, hipMemcpyDeviceToHost, st));
      MBA_HIP(hipStreamSynchronize(st));
      if (cnt == 0) return 0;
      src = ctx.hits2.as<HitRec>();
    }
    hrec.resize(cnt);
    MBA_HIP(hipMemcpy(hrec.data(), src, sizeof(HitRec) * cnt, hipMemcpyDeviceToHost));
    cands.clear();
    for (const HitRec &h : hrec) {
      const int sb = (int)(h.pair >> 16), j = (int)(h.pair & 0xffff);
      const int m = sb - 1 - j;
      const uint64_t rX = base_rank(T, sb, j, m, h);
      resolve_hit(T, W, t0, sb, rX, h.v & M, min_size, max_size, cands);
    }
    (void)sb_step;
    if (cands.empty()) return 0;
    const uint32_t nc_ = (uint32_t)cands.size();
    MBA_HIP(ctx.vsizes.ensure(4ull * nc_));
    MBA_HIP(ctx.vranks.ensure(8ull * nc_));
    MBA_HIP(ctx.vok.ensure(nc_));
    std::vector<uint32_t> sz(nc_);
    std::vector<uint64_t> rk(nc_);
    for (uint32_t i = 0; i < nc_; ++i) sz[i] = (uint32_t)cands[i].size, rk[i] = cands[i].rank;
    MBA_HIP(hipMemcpyAsync(ctx.vsizes.p, sz.data(), 4ull * nc_, hipMemcpyHostToDevice, st));
    MBA_HIP(hipMemcpyAsync(ctx.vranks.p, rk.data(), 8ull * nc_, hipMemcpyHostToDevice, st));
    MBA_HIP(launch_verify(W, d_tv, ctx.leafvals.as<uint64_t>(), ctx.outputs.as<uint64_t>(),
                          n, ctx.vsizes.as<uint32_t>(), ctx.vranks.as<uint64_t>(), nc_,
                          ctx.vok.as<uint8_t>(), st));
    std::vector<uint8_t> okv(nc_);
    MBA_HIP(hipMemcpyAsync(okv.data(), ctx.vok.p, nc_, hipMemcpyDeviceToHost, st));
    MBA_HIP(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < nc_; ++i)
      if (okv[i]) found[sz[i]] = std::min(found[sz[i]], rk[i]);
    return 0;
  };

  auto over_budget = [&]() {
    if (o.cancel && *o.cancel) return true;
    return o.time_budget_s > 0 && seconds_since(t_start) > o.time_budget_s;
  };

  // Step 1: leaves (size 1) and, through unary chains, every unary tree over a
  // leaf (sizes 2..max).
  MBA_HIP(hipMemsetAsync(d_hc, 0, 4, st));
  MBA_HIP(launch_leaf_scan(W, d_tv, L, targets_for(t0, 1, max_size, W), d_hits, d_hc, kHitCap, st));
  if (int rc = process_hits(1)) return rc;
  for (int s = 1; s <= std::min(2, max_size); ++s) {
    reached = s;
    if (s >= min_size && found[s] != ~0ull) {
      answer = s;
      break;
    }
  }

  // Runs the pair {j, m} (j <= m) of size s for the operator instances in
  // `mask`, in launches of bounded work (B sides above Kt one block at a
  // time).  Returns 1 when the budget stopped it.
  std::map<int, std::vector<BBlock>> bcache;  // B blocks per size above Kt
  auto run_pair = [&](int s, int j, uint32_t mask, const CheckTargets &tg) -> int {
    const int m = s - 1 - j;
    PairLaunch p{};
    p.sb = s;
    p.j = j;
    p.m = m;
    p.nA = T.t[j][8];
    p.nB = T.t[m][8];
    p.nTA = (p.nA + kPairTA - 1) / kPairTA;
    p.rb = o.lanes_b == 4 ? 4 : 8;
    const uint64_t tb = (uint64_t)kPairThreads * p.rb;
    p.gen_a = j > Kt;
    p.gen_b = m > Kt;
    p.mask = mask;
    std::vector<BBlock> one(1);
    const std::vector<BBlock> *blocks = &one;
    if (p.gen_b) {
      auto it = bcache.find(m * 16 + p.rb);
      if (it == bcache.end()) {
        std::vector<BBlock> v;
        b_blocks(T, Kt, tv.off, m, 0, 1, {}, p.rb, v);
        it = bcache.emplace(m * 16 + p.rb, std::move(v)).first;
      }
      blocks = &it->second;
    }
    const uint64_t per_tile = (uint64_t)kPairTA * tb;
    const uint64_t tiles_per_launch = std::max<uint64_t>(1, kPairsPerLaunch / per_tile);
    // launches go out back to back; the stream is synchronised (timing and
    // budget check) once a group holds about kPairsPerLaunch pairs
    uint64_t group_tiles = 0;
    bool group_open = false;
    auto close_group = [&]() -> int {
      if (!group_open) return 0;
      hipError_t e = hipEventRecord(ctx.ev1, st);
      if (e == hipSuccess) e = hipEventSynchronize(ctx.ev1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1);
      if (e != hipSuccess) {
        ctx.last_error = std::string("pair launch: ") + hipGetErrorString(e);
        return -1;
      }
      ktime += ms * 1e-3;
      group_open = false;
      group_tiles = 0;
      if (answer < 0) {
        // drain a filling hit buffer between groups (rare: many window
        // hits that other samples reject)
        uint32_t cnt = 0;
        e = hipMemcpy(&cnt, d_hc, 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
          ctx.last_error = std::string("hit count: ") + hipGetErrorString(e);
          return -1;
        }
        if (cnt > kHitCap / 2) {
          if (process_hits(s) != 0) return -1;
          if (hipMemsetAsync(d_hc, 0, 4, st) != hipSuccess) return -1;
        }
      }
      return over_budget() ? 1 : 0;
    };
    for (const BBlock &bb : *blocks) {
      p.bb = bb;
      p.nTB = p.gen_b ? block_tiles(bb, p.rb) : (p.nB + tb - 1) / tb;
      const uint64_t tiles = p.nTA * p.nTB;
      for (uint64_t t0 = 0; t0 < tiles; t0 += tiles_per_launch) {
        p.tile_begin = t0;
        p.tile_end = std::min(tiles, t0 + tiles_per_launch);
        const uint64_t ntl = p.tile_end - p.tile_begin;
        const int g = (int)std::min<uint64_t>((uint64_t)grid, ntl);
        hipError_t e = hipSuccess;
        if (!group_open) e = hipEventRecord(ctx.ev0, st);
        group_open = true;
        if (e == hipSuccess)
          e = launch_pair_check(W, d_tv, ctx.tables.p, tv.off, p, tg, d_hits, d_hc, kHitCap, g,
                                st);
        if (e != hipSuccess) {
          ctx.last_error = std::string("pair launch: ") + hipGetErrorString(e);
          return -1;
        }
        ++klaunch;
        group_tiles += ntl;
        if (group_tiles >= tiles_per_launch) {
          if (int rc = close_group()) return rc;
        }
      }
    }
    return close_group();
  };

  for (int s = 3; (answer < 0 || o.exhaustive) && s <= max_size; ++s) {
    if (over_budget()) {
      status = MBA_BUDGET;
      break;
    }
    MBA_HIP(hipMemsetAsync(d_hc, 0, 4, st));
    const CheckTargets tg = targets_for(t0, s, max_size, W);
    bool stopped = false;
    // block-at-a-time only pays where a size's sweep outweighs a launch and
    // a sync per block; below that the fused sweep (minimum rank, also a
    // valid answer for the learned order) is used
    const uint64_t block_min = o.block_log2 > 0 ? (1ull << std::min(o.block_log2, 63))
                               : block_scores ? kLearnedBlockMin : kBlockModeMin;
    const bool by_block = !o.exhaustive && T.t[s][8] >= block_min;
    if (!o.exhaustive && s >= min_size && found[s] != ~0ull &&
        (block_scores || found[s] < T.t[s][NOT])) {
      // a unary-rooted solution of this size is already verified: with the
      // learned order any solution of the minimal size is accepted; in rank
      // order one in the '~' block precedes every binary-rooted tree
      answer = s;
      reached = s;
      break;
    }
    if (!by_block) {
      // every split of this size, all seven operator instances fused
      for (int j = 1; j <= (s - 1) / 2 && !stopped; ++j) {
        const int rc = run_pair(s, j, kAllInst, tg);
        if (rc < 0) return MBA_ERR_HIP;
        stopped = rc == 1;
      }
      // once the answer is settled (an exhaustive sweep going on to max_size)
      // the windows' hits no longer matter: their count may exceed the buffer
      if (!stopped && answer < 0) {
        if (int rc = process_hits(s)) return rc;
      }
    } else {
      // one (op, split) block at a time, in rank order (plain) or in the
      // order of the learned scores; stop once the answer is settled
      std::vector<BlockRef> blocks = binary_blocks(T, s);
      if (block_scores) {
        std::stable_sort(blocks.begin(), blocks.end(), [&](const BlockRef &x, const BlockRef &y) {
          return block_scores[x.op * 32 + x.split] > block_scores[y.op * 32 + y.split];
        });
      }
      std::vector<char> done(blocks.size(), 0);
      for (size_t bi = 0; bi < blocks.size() && !stopped; ++bi) {
        const BlockRef &b = blocks[bi];
        const int rc = run_pair(s, b.pair_j, 1u << b.inst, tg);
        if (rc < 0) return MBA_ERR_HIP;
        if (rc == 1) {
          stopped = true;
          break;
        }
        if (int rc2 = process_hits(s)) return rc2;
        MBA_HIP(hipMemsetAsync(d_hc, 0, 4, st));
        done[bi] = 1;
        if (found[s] == ~0ull || s < min_size) continue;
        if (block_scores) break;  // any verified tree of the minimal size
        // rank order: settled when every block below the best rank is done;
        // the neg block after '^' is complete already (unary chains)
        const uint64_t prefix = b.end == T.op_offset(s, NEG) ? T.op_offset(s, ADD) : b.end;
        if (found[s] < prefix) break;
      }
    }
    if (stopped) {
      status = MBA_BUDGET;
      break;
    }
    reached = s;
    if (o.verbose)
      std::fprintf(stderr, "Length %-2d | candidates %-20llu | %.3f s\n", s,
                   (unsigned long long)T.t[s][8], seconds_since(t_start));
    if (answer < 0 && s >= min_size && found[s] != ~0ull) answer = s;
  }

  uint64_t cand = 0;
  for (int s = min_size; s <= reached; ++s) cand += T.t[s][8];
  res.candidates = cand;
  res.size_reached = reached;
  res.kernel_seconds = ktime;
  res.kernel_launches = klaunch;
  if (answer > 0) {
    fill_result(T, answer, found[answer], nv, inst.consts, W, res);
    status = MBA_FOUND;
  } else {
    std::snprintf(res.expr, sizeof res.expr, "Not found !");
  }
  res.status = status;
  res.seconds = seconds_since(t_start);
  return status;
}

}  // namespace mba
