// simba_host.cpp -- host driver of the SIMBA search on one gfx950 device.
//
// Replaces NoCache::solve (simba/src/nocache.cu:197-289).  Sizes are visited
// in increasing order; the first solution is the minimum rank, in the
// reference's enumeration order, within the smallest size that has one (the
// reference keeps whichever thread wins atomicCAS inside the first
// (size, op) launch with a match, nocache.cu:102-105, which is one of the
// trees of that size and op block).
#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "context.h"
#include "simba_host.h"

namespace mba {

// default window-hit buffer: more than the 256 x 2048 x 7 hits one tile can
// record, so replaying an overflowing launch group in parts always succeeds
constexpr uint32_t kDefaultHitCap = 1u << 22;
// value tables up to this many entries in all are built by one workgroup
constexpr uint64_t kSmallTableValues = 1u << 16;
constexpr uint64_t kDefaultTableBytes = 40ull << 30;
// Work per launch, in A x B value pairs, so that budgets are honoured.
constexpr uint64_t kPairsPerLaunch = 1ull << 36;
// pair launches of fewer tiles than kSmallLaunchTiles use A tiles down to
// kMinTA values
constexpr uint64_t kSmallLaunchTiles = 128;
constexpr uint32_t kMinTA = 16;
// trees of the small sizes swept in one fused step (launched back to back,
// their hits resolved in one sync)
constexpr uint64_t kFuseTrees = 1ull << 22;

// Tile geometry of a B block of nA x nB values (see BBlock): columns run
// over the larger dimension, 2^logTB of them, and the tile's remaining
// (256*rb) >> logTB lanes take successive rows.  When the columns span a
// whole tile, a tile takes up to kBandRows rows of its columns (the column
// values are read once for all of them).

static BBlock shape_block(BBlock b, int rb) {
  const uint64_t tile = (uint64_t)kPairThreads * rb;
  b.trans = b.nB >= b.nA ? 0 : 1;
  const uint64_t nX = b.trans ? b.nA : b.nB, nY = b.trans ? b.nB : b.nA;
  int lg = 0;
  while ((1ull << lg) < nX && (1ull << lg) < tile) ++lg;
  b.logTB = lg;
  b.ntc = (nX + (1ull << lg) - 1) >> lg;
  b.ntc_div = div_magic(b.ntc <= 0xffffffffull ? (uint32_t)b.ntc : 1u);
  b.rows_pt = (1ull << lg) == tile ? (int32_t)std::min<uint64_t>(nY, kBandRows) : 1;
  return b;
}

uint64_t block_tiles(const BBlock &b, int rb) {
  const uint64_t tile = (uint64_t)kPairThreads * rb;
  if (b.generic) return (b.nA + tile - 1) / tile;
  const uint64_t nY = b.trans ? b.nB : b.nA;
  const uint64_t rows = (1ull << b.logTB) == tile ? (uint64_t)b.rows_pt : tile >> b.logTB;
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

void b_blocks(const CountTable &T, int Kt, int m, uint64_t start, uint64_t stride,
              const std::vector<BStep> &steps, int rb, std::vector<BBlock> &out) {
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
        b.offA = tab_ref(m - 1, 0);
        b.nA = T.t[m - 1][8];
        b.nB = 1;
        b.op = -1;
        emit(b, st);
      } else {
        b_blocks(T, Kt, m - 1, start + stride * r, stride, st, rb, out);
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
          b.offA = tab_ref(l, 0);
          b.offB = tab_ref(rr, 0);
          b.nA = nl;
          b.nB = nr;
          b.op = op;
          emit(b, steps);
        } else if (l > Kt && rr <= Kt && nr <= kFixedSiblings) {
          // right child fixed to tree b: rank base + stride*(ra*nr + b)
          for (uint64_t bi = 0; bi < nr; ++bi)
            b_blocks(T, Kt, l, base + stride * bi, stride * nr,
                     with(BStep{2, op, tab_ref(rr, bi)}), rb, out);
        } else if (rr > Kt && l <= Kt && nl <= kFixedSiblings) {
          // left child fixed to tree a: rank base + stride*(a*nr + rb)
          for (uint64_t ai = 0; ai < nl; ++ai)
            b_blocks(T, Kt, rr, base + stride * ai * nr, stride,
                     with(BStep{1, op, tab_ref(l, ai)}), rb, out);
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

std::vector<int> learned_split_order(const CountTable &T, int s, const float *scores) {
  const std::vector<BlockRef> blocks = binary_blocks(T, s);
  float top = -3.4e38f;
  for (const BlockRef &b : blocks) top = std::max(top, scores[b.op * 32 + b.split]);
  const int nj = (s - 1) / 2;
  std::vector<double> mass(nj + 1, 0.0);
  for (const BlockRef &b : blocks) mass[b.pair_j] += std::exp((double)(scores[b.op * 32 + b.split] - top));
  std::vector<int> js;
  for (int j = 1; j <= nj; ++j) js.push_back(j);
  auto key = [&](int j) { return mass[j] / (double)(T.t[j][8] * T.t[s - 1 - j][8]); };
  std::stable_sort(js.begin(), js.end(), [&](int x, int y) { return key(x) > key(y); });
  return js;
}

int pick_filter_sample(const uint64_t *lv, const uint64_t *outv, int n, int L, int W,
                       int skip) {
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
  // A chain of u <= U unary nodes over a tree of value v computes
  // (-1)^u v + k; even u reach |k| <= u/2, odd u reach
  // -(u+1)/2 <= k <= (u-1)/2 (each '~' adds -1 to the sign-adjusted
  // constant, each neg 0).  So v = t0 - k lies in [t0 - U/2, t0 + U/2] and,
  // for odd chains, v = k - t0 in [-t0 - (U+1)/2, -t0 + (U-1)/2]: both
  // inside windows of span U (one value wider than needed for one parity).
  const uint64_t M = wmask(W);
  const uint64_t U = (uint64_t)(max_size - sb);
  CheckTargets tg;
  tg.lo_even = (t0 - U / 2) & M;
  tg.lo_odd = ((0 - t0) - (U + 1) / 2) & M;
  tg.span = U;
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
                const float *block_scores, bool rtid, const std::function<int(int)> *score_size) {
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
  const int fs = pick_filter_sample(lv.data(), outv.data(), n, L, W);
  const uint64_t t0 = outv[fs];

  // ---- value tables V[1..Kt] -------------------------------------------
  const uint64_t budget = o.table_bytes ? o.table_bytes : kDefaultTableBytes;
  // V[max_size - 2] is only ever the B side of leaves (one read per value),
  // cheaper to form per tile than to build; RTid reads every size directly
  const int kt_cap = rtid ? max_size : max_size - 3;
  const int Kt = std::max(1, std::min(choose_kt(T, max_size, W, budget), kt_cap) -
                                 std::max(0, o.table_delta));
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
  MBA_HIP(ctx.tables.ensure(std::max<size_t>(elems * wb, 256)));
  tv.base = ctx.tables.p;
  MBA_HIP(ctx.tview.ensure(sizeof(TableView)));
  // hit buffer capacity (opts.hit_cap: smaller buffers exercise the drain
  // and replay paths in tests)
  const uint32_t kHitCap = o.hit_cap > 0 ? (uint32_t)o.hit_cap : kDefaultHitCap;
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
  {  // V[1] = leaf values on the filter sample
    std::vector<uint64_t> v1(L);
    std::vector<uint32_t> v1w(L);
    for (int l = 0; l < L; ++l)
      v1[l] = lv[(size_t)fs * L + l], v1w[l] = (uint32_t)lv[(size_t)fs * L + l];
    if (W == 64)
      MBA_HIP(hipMemcpyAsync(ctx.tables.p, v1.data(), 8 * L, hipMemcpyHostToDevice, st));
    else
      MBA_HIP(hipMemcpyAsync(ctx.tables.p, v1w.data(), 4 * L, hipMemcpyHostToDevice, st));
  }
  // Value tables are built as the sweep reaches the sizes that read them
  // (size s reads V[m] for m <= s - 2), so a search settled at a small size
  // builds only small tables: the small sizes in one workgroup (one
  // launch), the rest per block.
  int m_small = 1;
  for (uint64_t acc = 0; m_small + 1 <= Kt && acc + T.t[m_small + 1][8] <= kSmallTableValues;)
    acc += T.t[++m_small][8];
  int built = 1;  // V[1] holds the leaf values already
  auto ensure_tables = [&](int upto) -> int {
    upto = std::min(upto, Kt);
    if (upto <= built) return 0;
    if (built < m_small) {
      MBA_HIP(launch_table_small(W, d_tv, 2, m_small, st));
      built = m_small;
    }
    for (int m = built + 1; m <= upto; ++m)
      MBA_HIP(launch_table_build(W, ctx.tables.p, tv.off, T, m, st));
    built = std::max(built, upto);
    return 0;
  };
  if (rtid) {  // RTid reads every tabulated size directly
    if (int rc = ensure_tables(Kt)) return rc;
  }

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
  // sizes min..sweep_last are swept; the windows are those of max_size
  const int sweep_last = o.stop_size > 0 ? std::min(o.stop_size, max_size) : max_size;
  std::vector<uint64_t> found(kMaxSize + 2, ~0ull);
  HitRec *d_hits = ctx.hits.as<HitRec>();
  // 32 workgroups per CU by default: more than stay resident, so tiles are
  // handed out in small grid-stride shares and no CU waits on a long tail
  const int grid = ctx.cu_count * (o.wg_per_cu > 0 ? o.wg_per_cu : 32);
  double ktime = 0.0;
  uint32_t klaunch = 0;
  int answer = -1;
  int reached = 0;
  int status = MBA_NOT_FOUND;

  // Device hit resolution (k_hit_resolve): the per-size minimum chain rank
  // of the hits that meet every sample, flags and the list of lowered slots;
  // the instance's sample layout for it at byte 128 of the header.
  // The read-back area (the pinned layout: hit count, lowered slot count,
  // candidate count, candidates, their check results) follows, so one copy returns it.
  constexpr size_t kSlots = kMaxSize + 1;
  constexpr size_t kReadOff = (256 + kSlots * 16 + 255) / 256 * 256;
  constexpr size_t kReadBytes = 16 + kSlots * (sizeof(VerifyCand) + 1);
  MBA_HIP(ctx.b_misc.ensure(kReadOff + kReadBytes));
  char *d_read = ctx.b_misc.as<char>() + kReadOff;
  uint32_t *d_hc = (uint32_t *)d_read;
  uint32_t *d_nt = d_hc + 1;  // lowered slots, candidates
  // work-queue head of the fused steps' persistent launches
  unsigned long long *d_head = (unsigned long long *)ctx.b_misc.p;
  VerifyCand *d_vc = (VerifyCand *)(d_read + 16);
  uint8_t *d_ok = (uint8_t *)(d_vc + kSlots);
  BatchInstData *d_idat = (BatchInstData *)(ctx.b_misc.as<char>() + 128);
  unsigned long long *d_found = (unsigned long long *)(ctx.b_misc.as<char>() + 256);
  uint32_t *d_flags = (uint32_t *)(d_found + kSlots);
  uint32_t *d_touched = d_flags + kSlots;
  const BatchInstData idat1{0, 0, (uint32_t)n, (uint32_t)fs};
  MBA_HIP(hipMemcpyAsync(d_idat, &idat1, sizeof idat1, hipMemcpyHostToDevice, st));
  MBA_HIP(hipMemsetAsync(d_found, 0xff, kSlots * 8, st));
  MBA_HIP(hipMemsetAsync(d_flags, 0, kSlots * 4, st));
  MBA_HIP(ctx.surv.ensure(sizeof(SurvRec) * 2 * (size_t)kHitCap));
  static_assert(kReadBytes <= Ctx::kPinnedBytes, "pinned area");
  uint32_t *h_counts = (uint32_t *)ctx.pinned;
  VerifyCand *h_vc = (VerifyCand *)((char *)ctx.pinned + 16);
  uint8_t *h_ok = (uint8_t *)(h_vc + kSlots);

  // Resolve the hits of one step on the device and verify the lowered slots'
  // trees; the kernels read the counts on the device, so one sync suffices.
  // The hit buffer is empty afterwards (its count back to zero).
  // allow_overflow: a count above the buffer is the caller's to handle (the
  // first kHitCap hits are resolved).
  auto process_hits = [&](bool allow_overflow = false) -> int {
    MBA_HIP(launch_hit_resolve(W, d_tv, ctx.leafvals.as<uint64_t>(), ctx.outputs.as<uint64_t>(),
                               d_idat, d_hits, d_hc, kHitCap, 512, 512, min_size, max_size,
                               d_found, ctx.surv.as<SurvRec>(), d_nt + 1, d_flags, d_touched,
                               d_nt, st));
    MBA_HIP(launch_found_gather(d_touched, d_nt, kSlots, d_found, d_flags, d_vc, st));
    MBA_HIP(launch_verify_batch(W, d_tv, ctx.leafvals.as<uint64_t>(), ctx.outputs.as<uint64_t>(),
                                d_idat, d_vc, kSlots, d_ok, st, d_nt));
    MBA_HIP(hipMemcpyAsync(ctx.pinned, d_read, kReadBytes, hipMemcpyDeviceToHost, st));
    // the hit, lowered slot and candidate counts start again from zero
    MBA_HIP(hipMemsetAsync(d_read, 0, 12, st));
    MBA_HIP(hipStreamSynchronize(st));
    if (h_counts[0] > kHitCap && !allow_overflow) {
      ctx.last_error = "window-hit buffer overflow";
      return MBA_ERR_LIMIT;
    }
    const uint32_t nt = h_counts[1];
    if (nt > kSlots) {
      ctx.last_error = "hit resolution: slot list overflow";
      return MBA_ERR_LIMIT;
    }
    for (uint32_t i = 0; i < nt; ++i)
      if (h_ok[i]) found[h_vc[i].size] = std::min(found[h_vc[i].size], h_vc[i].rank);
    return 0;
  };

  int replays = 0;  // launch groups replayed after an overflow

  auto over_budget = [&]() {
    if (o.cancel && *o.cancel) return true;
    return o.time_budget_s > 0 && seconds_since(t_start) > o.time_budget_s;
  };

  // block-at-a-time only pays where a size's sweep outweighs a launch and a
  // sync per block; below that the fused sweep (minimum rank, also a valid
  // answer for the learned order) is used
  const uint64_t block_min = o.block_log2 > 0 ? (1ull << std::min(o.block_log2, 63))
                             : block_scores ? kLearnedBlockMin : kBlockModeMin;
  // Small sizes are swept in fused steps: sizes s0..last (below kFuseTrees
  // trees in all, each below the block threshold) are launched back to back
  // and their hits resolved in one sync; the answer is then the smallest size
  // of the step with a verified tree, as size by size.  Returns s0 - 1 when
  // s0 is not fused (exhaustive sweeps go size by size).
  auto fused_last = [&](int s0) -> int {
    int e = s0 - 1;
    if (o.exhaustive) return e;
    for (uint64_t acc = 0; e < sweep_last && T.t[e + 1][8] < block_min &&
                           acc + T.t[e + 1][8] <= kFuseTrees;)
      acc += T.t[++e][8];
    return e;
  };

  // A fused step as one persistent launch (k_batch_step over the step's
  // splits as jobs, narrow A tiles for the small ones; a B side above Kt --
  // V[max_size - 2], the B side of the leaves -- as one job per B block).
  // Returns 1 when its hits were resolved; otherwise the caller runs the
  // splits one launch at a time: 0 when nothing was launched (more jobs and
  // blocks than the staging area holds; the buffer keeps its hits), 2 when
  // the hits overflowed the buffer (those resolved from the truncated buffer
  // stand, as they only lower slots to verified ranks; the buffer is empty);
  // < 0 on an error.
  std::vector<BatchJob> f_jobs;
  std::vector<BBlock> f_blocks, f_tmp;
  auto fused_persistent = [&](int s0, int s1) -> int {
    if (int rc = ensure_tables(s1 - 2)) return rc;
    f_jobs.clear();
    f_blocks.assign(1, BBlock{});  // block 0: named by jobs with a tabulated B side
    uint64_t tiles = 0;
    auto push = [&](BatchJob J, uint64_t ntb) {
      J.p.nTB = ntb;
      J.p.ta = kPairTA;
      while (J.p.ta > kMinTA && (J.p.nA + J.p.ta - 1) / J.p.ta * ntb < kSmallLaunchTiles)
        J.p.ta /= 2;
      J.p.nTA = (J.p.nA + J.p.ta - 1) / J.p.ta;
      J.ntiles = J.p.nTA * ntb;
      J.tile0 = tiles;
      tiles += J.ntiles;
      f_jobs.push_back(J);
    };
    for (int z = s0; z <= s1; ++z) {
      const CheckTargets tgz = targets_for(t0, z, max_size, W);
      for (int j = 1; j <= (z - 1) / 2; ++j) {
        BatchJob J;
        std::memset(&J, 0, sizeof J);
        J.p.sb = z;
        J.p.j = j;
        J.p.m = z - 1 - j;
        J.p.nA = T.t[j][8];
        J.p.nB = T.t[J.p.m][8];
        J.p.gen_a = j > Kt;
        J.p.gen_b = J.p.m > Kt;
        J.p.mask = kAllInst;
        J.p.tri = j == J.p.m && !J.p.gen_a && !J.p.gen_b;
        J.p.rb = kPairRB;
        J.lo_e = tgz.lo_even;
        J.lo_o = tgz.lo_odd;
        J.span = tgz.span;
        J.kind = 1;
        if (!J.p.gen_b) {
          push(J, (J.p.nB + kPairTB - 1) / kPairTB);
          continue;
        }
        f_tmp.clear();
        b_blocks(T, Kt, J.p.m, 0, 1, {}, kPairRB, f_tmp);
        for (const BBlock &b : f_tmp) {
          J.bb = (uint32_t)f_blocks.size();
          f_blocks.push_back(b);
          push(J, block_tiles(b, kPairRB));
        }
      }
    }
    const size_t nj = f_jobs.size();
    if (nj == 0) return 0;
    // staging (pinned, then one copy to the device): step header, jobs,
    // blocks
    constexpr size_t kStageOff = 4096;
    const size_t off_blk = (256 + sizeof(BatchJob) * nj + 255) / 256 * 256;
    const size_t bytes = off_blk + sizeof(BBlock) * f_blocks.size();
    if (kStageOff + bytes > Ctx::kPinnedBytes) return 0;
    char *stage = (char *)ctx.pinned + kStageOff;
    auto *h_step = (BatchStep *)stage;
    std::memset(h_step, 0, sizeof *h_step);
    h_step->total_tiles = tiles;
    // tabulated B sides first, then those formed from blocks (BatchStep)
    {
      auto tabulated = [](const BatchJob &J) { return !J.p.gen_b; };
      std::stable_partition(f_jobs.begin(), f_jobs.end(), tabulated);
      uint64_t t = 0;
      h_step->tiles_table = tiles;
      for (BatchJob &J : f_jobs) {
        if (h_step->tiles_table == tiles && !tabulated(J)) h_step->tiles_table = t;
        J.tile0 = t;
        t += J.ntiles;
      }
    }
    h_step->njobs = (uint32_t)nj;
    std::memcpy(stage + 256, f_jobs.data(), sizeof(BatchJob) * nj);
    std::memcpy(stage + off_blk, f_blocks.data(), sizeof(BBlock) * f_blocks.size());
    MBA_HIP(ctx.b_queue.ensure(Ctx::kPinnedBytes - kStageOff));
    MBA_HIP(hipMemcpyAsync(ctx.b_queue.p, stage, bytes, hipMemcpyHostToDevice, st));
    MBA_HIP(hipMemsetAsync(d_head, 0, 16, st));  // the two parts' tile counters
    MBA_HIP(hipEventRecord(ctx.ev0, st));
    const int g = (int)std::min<uint64_t>((uint64_t)ctx.cu_count * 8, tiles);
    MBA_HIP(launch_batch_step(W, ctx.b_queue.as<BatchStep>(),
                              (const BatchJob *)(ctx.b_queue.as<char>() + 256),
                              (const BBlock *)(ctx.b_queue.as<char>() + off_blk), d_tv,
                              d_head, d_hits, d_hc, kHitCap, nullptr, g, st));
    MBA_HIP(hipEventRecord(ctx.ev1, st));
    ++klaunch;
    if (int rc = process_hits(true)) return rc;  // syncs; rc < 0
    float ms = 0.f;
    MBA_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
    ktime += ms * 1e-3;
    return *(const uint32_t *)ctx.pinned > kHitCap ? 2 : 1;
  };

  // Step 1: leaves (size 1) and, through unary chains, every unary tree over a
  // leaf (sizes 2..max).  Their hits join the first fused step when there is
  // one.
  MBA_HIP(hipMemsetAsync(d_read, 0, 12, st));  // hit, lowered slot, candidate counts
  MBA_HIP(launch_leaf_scan(W, d_tv, L, targets_for(t0, 1, max_size, W), d_hits, d_hc, kHitCap, st));
  bool leaf_pending = sweep_last >= 3 && fused_last(3) >= 3;
  if (!leaf_pending) {
    if (int rc = process_hits()) return rc;
    for (int s = 1; s <= std::min(2, sweep_last); ++s) {
      reached = s;
      if (s >= min_size && found[s] != ~0ull) {
        answer = s;
        break;
      }
    }
  }

  // The open launch group: pair launches sent back to back since the last
  // sync (all of one size's targets), kept for a replay in parts when their
  // hits overflow the buffer.
  bool group_open = false;
  uint64_t group_tiles = 0;
  // pairs of the split being swept: of B blocks whose launches were all
  // issued (pending) and of those whose launches have completed (done)
  uint64_t split_pairs_pending = 0, split_pairs_done = 0;
  std::vector<std::pair<PairLaunch, CheckTargets>> group;
  // Replays a launch's tile range in `pieces` parts, resolving each part's
  // hits before the next; a part that still overflows the hit buffer is
  // split again (a single tile that overflows it is an error).
  std::function<int(const PairLaunch &, const CheckTargets &, int)> replay =
      [&](const PairLaunch &q, const CheckTargets &tg, int pieces) -> int {
    const uint64_t n = q.tile_end - q.tile_begin;
    const uint64_t step = std::max<uint64_t>(1, (n + pieces - 1) / pieces);
    for (uint64_t b = q.tile_begin; b < q.tile_end; b += step) {
      PairLaunch r = q;
      r.tile_begin = b;
      r.tile_end = std::min(q.tile_end, b + step);
      const int g = (int)std::min<uint64_t>((uint64_t)grid, r.tile_end - r.tile_begin);
      uint32_t cnt = 0;
      hipError_t e =
          launch_pair_check(W, d_tv, ctx.tables.p, tv.off, r, tg, d_hits, d_hc, kHitCap, g, st);
      if (e == hipSuccess) e = hipMemcpyAsync(&cnt, d_hc, 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);  // `st` does not wait on null
      if (e != hipSuccess) {
        ctx.last_error = std::string("pair replay: ") + hipGetErrorString(e);
        return -1;
      }
      ++klaunch;
      if (cnt > kHitCap) {
        if (r.tile_end - r.tile_begin == 1) {
          ctx.last_error = "hit buffer overflow within one tile";
          return -1;
        }
        if (hipMemsetAsync(d_hc, 0, 4, st) != hipSuccess) return -1;
        if (int rc = replay(r, tg, 8)) return rc;
        continue;
      }
      if (cnt > 0 && process_hits() != 0) return -1;
    }
    return 0;
  };
  // Closes the open group: waits for it (timing, budget check) and, while the
  // answer is open, watches the hit buffer.  resolve: resolve every recorded
  // hit in the same sync (end of a fused step); otherwise a filling buffer is
  // drained between groups.  A group that overflowed the buffer is replayed
  // in parts after the recorded hits (everything before the group and part
  // of it) are resolved: resolving any subset of the hits only lowers slots
  // to verified ranks.  Returns 1 when over the budget.
  auto close_group = [&](bool resolve) -> int {
    if (!group_open && !resolve) return 0;
    hipError_t e = group_open ? hipEventRecord(ctx.ev1, st) : hipSuccess;
    uint32_t cnt = 0;
    if (e == hipSuccess && resolve) {
      if (process_hits(true) != 0) return -1;
      cnt = *(const uint32_t *)ctx.pinned;
    } else if (e == hipSuccess && group_open) {
      e = hipEventSynchronize(ctx.ev1);
      // drain a filling hit buffer between groups (rare: many window hits
      // that other samples reject)
      if (e == hipSuccess && answer < 0) e = hipMemcpy(&cnt, d_hc, 4, hipMemcpyDeviceToHost);
    }
    float ms = 0.f;
    if (e == hipSuccess && group_open) e = hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1);
    if (e != hipSuccess) {
      ctx.last_error = std::string("pair launch: ") + hipGetErrorString(e);
      return -1;
    }
    ktime += ms * 1e-3;
    group_open = false;
    group_tiles = 0;
    split_pairs_done += split_pairs_pending;  // every launch issued so far is done
    split_pairs_pending = 0;
    if (resolve || answer < 0) {
      if (cnt > kHitCap) {
        if (group.empty()) {
          ctx.last_error = "window-hit buffer overflow";
          return -1;
        }
        if (!resolve && process_hits(true) != 0) return -1;
        ++replays;
        for (const auto &q : group)
          if (int rc = replay(q.first, q.second, 8)) return rc;
      } else if (!resolve && cnt > kHitCap / 2) {
        if (process_hits() != 0) return -1;
      }
    }
    group.clear();
    return over_budget() ? 1 : 0;
  };

  // Runs the pair {j, m} (j <= m) of size s for the operator instances in
  // `mask`, in launches of bounded work (B sides above Kt one block at a
  // time).  keep_open: the last launch group stays open (closed by the
  // caller).  Returns 1 when the budget stopped it.
  std::map<int, std::vector<BBlock>> bcache;  // B blocks per size above Kt
  auto run_pair = [&](int s, int j, uint32_t mask, const CheckTargets &tg,
                      bool keep_open = false) -> int {
    const int m = s - 1 - j;
    PairLaunch p{};
    p.sb = s;
    p.j = j;
    p.m = m;
    p.nA = T.t[j][8];
    p.nB = T.t[m][8];
    p.rb = o.lanes_b == 4 ? 4 : 8;
    const uint64_t tb = (uint64_t)kPairThreads * p.rb;
    p.gen_a = j > Kt;
    p.gen_b = m > Kt;
    p.mask = mask;
    // an equal-size split with both sides tabulated: only the tiles that
    // reach the diagonal, every operator instance (min-rank trees a <= b)
    p.tri = j == m && !p.gen_a && !p.gen_b && (mask & kAllInst) == kAllInst;
    std::vector<BBlock> one(1);
    const std::vector<BBlock> *blocks = &one;
    if (p.gen_b) {
      auto it = bcache.find(m * 16 + p.rb);
      if (it == bcache.end()) {
        std::vector<BBlock> v;
        b_blocks(T, Kt, m, 0, 1, {}, p.rb, v);
        it = bcache.emplace(m * 16 + p.rb, std::move(v)).first;
      }
      blocks = &it->second;
    }
    const uint64_t per_tile = (uint64_t)kPairTA * tb;
    const uint64_t tiles_per_launch = std::max<uint64_t>(1, kPairsPerLaunch / per_tile);
    for (const BBlock &bb : *blocks) {
      p.bb = bb;
      p.nTB = p.gen_b ? block_tiles(bb, p.rb) : (p.nB + tb - 1) / tb;
      // a launch of few tiles gets narrower A tiles, so that more
      // workgroups share it (latency of the small sizes)
      p.ta = kPairTA;
      while (p.ta > kMinTA && (p.nA + p.ta - 1) / p.ta * p.nTB < kSmallLaunchTiles) p.ta /= 2;
      p.nTA = (p.nA + p.ta - 1) / p.ta;
      const uint64_t tiles = p.nTA * p.nTB;
      const uint64_t block_pairs =
          p.nA * (!p.gen_b ? p.nB : bb.generic ? bb.nA : bb.nA * bb.nB);
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
        group.emplace_back(p, tg);
        group_tiles += ntl;
        if (p.tile_end == tiles) split_pairs_pending += block_pairs;  // block fully issued
        if (group_tiles >= tiles_per_launch) {
          if (int rc = close_group(false)) return rc;
        }
      }
    }
    return keep_open ? 0 : close_group(false);
  };

  for (int s = 3; (answer < 0 || o.exhaustive) && s <= sweep_last; ++s) {
    if (over_budget()) {
      status = MBA_BUDGET;
      break;
    }
    const int last = fused_last(s);
    if (int rc = ensure_tables(std::max(last, s) - 2)) return rc;
    if (last >= s) {
      // fused step: every split of sizes s..last, all seven operator
      // instances (the hit buffer is empty, or holds the leaves' hits).
      // The splits become the jobs of one persistent launch, so they run
      // side by side; a step whose hits overflow the buffer (or with more
      // jobs than the staging holds) runs as one launch per split instead.
      int rc = 0;
      int prc = fused_persistent(s, last);
      if (prc < 0) return prc;
      if (prc == 2) {
        // the persistent launch overflowed the hit buffer: the leaf scan's
        // hits (resolved only up to the buffer's capacity) are scanned
        // again on their own, then the splits run one launch at a time
        ++replays;
        if (leaf_pending) {
          MBA_HIP(launch_leaf_scan(W, d_tv, L, targets_for(t0, 1, max_size, W), d_hits, d_hc,
                                   kHitCap, st));
          if (int lrc = process_hits()) return lrc;  // leaf hits above the buffer: an error
        }
        prc = 0;
      }
      if (prc == 0) {
        for (int z = s; z <= last && rc == 0; ++z) {
          const CheckTargets tgz = targets_for(t0, z, max_size, W);
          for (int j = 1; j <= (z - 1) / 2 && rc == 0; ++j)
            rc = run_pair(z, j, kAllInst, tgz, true);
        }
      }
      if (rc < 0) return MBA_ERR_HIP;
      if (rc == 1) {
        status = MBA_BUDGET;  // stopped before the step's hits were resolved
        break;
      }
      // resolve the hits (done already by a persistent launch); 1 when over
      // the budget after it
      rc = prc == 1 ? (over_budget() ? 1 : 0) : close_group(true);
      if (rc < 0) return MBA_ERR_HIP;
      for (int z = leaf_pending ? 1 : s; z <= last; ++z) {
        reached = z;
        if (o.verbose && z >= 3)
          std::fprintf(stderr, "Length %-2d | candidates %-20llu | %.3f s\n", z,
                       (unsigned long long)T.t[z][8], seconds_since(t_start));
        if (z >= min_size && found[z] != ~0ull) {
          answer = z;
          break;
        }
      }
      leaf_pending = false;
      if (answer < 0 && rc == 1) {
        status = MBA_BUDGET;
        break;
      }
      s = last;
      continue;
    }
    MBA_HIP(hipMemsetAsync(d_hc, 0, 4, st));
    const CheckTargets tg = targets_for(t0, s, max_size, W);
    bool stopped = false;
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
      // every split of this size, all seven operator instances fused, in one
      // launch group (unless large) whose hits are resolved in its closing
      // sync; once the answer is settled (an exhaustive sweep going on to
      // max_size) the windows' hits no longer matter
      int splits_done = 0;
      for (int j = 1; j <= (s - 1) / 2 && !stopped; ++j) {
        split_pairs_pending = split_pairs_done = 0;
        const int rc = run_pair(s, j, kAllInst, tg, true);
        if (rc < 0) return MBA_ERR_HIP;
        stopped = rc == 1;
        if (rc == 0) splits_done = j;
      }
      if (stopped) {
        res.split_reached = splits_done;
        res.split_pairs = split_pairs_done;
      }
      if (!stopped) {
        const int rc = close_group(answer < 0);
        if (rc < 0) return MBA_ERR_HIP;
        stopped = rc == 1;
      }
    } else {
      // (op, split) blocks in rank order (plain), or splits in the order of
      // the learned scores; stop once the answer is settled
      std::vector<BlockRef> blocks = binary_blocks(T, s);
      if (block_scores && score_size) {
        if (int rc = (*score_size)(s)) return rc;
      }
      if (block_scores) {
        ++res.ordered_sizes;
        // learned order: the splits (pairs {j, s-1-j}) in the order of the
        // scores' probability of their blocks per pair checked, each swept
        // once with all operator instances fused (its B side is formed
        // once), in rounds: the first split alone, then one split per
        // round, small splits sharing a round until it holds kRoundPairs
        // pairs (a round costs a sync); a round that yields a verified
        // solution ends the size (any solution of the minimal size is
        // accepted)
        constexpr uint64_t kRoundPairs = 1ull << 26;
        const std::vector<int> js = learned_split_order(T, s, block_scores);
        uint64_t round_pairs = 0;
        for (size_t k = 0; k < js.size() && !stopped; ++k) {
          const int rc = run_pair(s, js[k], kAllInst, tg);
          if (rc < 0) return MBA_ERR_HIP;
          stopped = rc == 1;
          if (stopped) break;
          round_pairs += T.t[js[k]][8] * T.t[s - 1 - js[k]][8];
          if (k > 0 && k + 1 < js.size() && round_pairs < kRoundPairs) continue;
          round_pairs = 0;
          if (int rc2 = process_hits()) return rc2;
          if (s >= min_size && found[s] != ~0ull) break;
        }
      } else {
        // rank order: one block at a time; settled when every block below
        // the best rank is done (the neg block after '^' is complete
        // already: unary chains)
        for (size_t bi = 0; bi < blocks.size() && !stopped; ++bi) {
          const BlockRef &b = blocks[bi];
          const int rc = run_pair(s, b.pair_j, 1u << b.inst, tg);
          if (rc < 0) return MBA_ERR_HIP;
          if (rc == 1) {
            stopped = true;
            break;
          }
          if (int rc2 = process_hits()) return rc2;
          if (found[s] == ~0ull || s < min_size) continue;
          const uint64_t prefix = b.end == T.op_offset(s, NEG) ? T.op_offset(s, ADD) : b.end;
          if (found[s] < prefix) break;
        }
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
  res.hit_replays = replays;
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
