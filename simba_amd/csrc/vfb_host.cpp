// vfb_host.cpp -- host driver of the VFB bottom-up cached search.
//
// Replaces MBA() (vfb/src/full-cache.cu:665-774) and generateMBAs<op>
// (:568-663).  Level 1 holds the leaves and is checked on the host in leaf
// order (:694-707).  Level L >= 2 enumerates, in order, the ops
// ~ & | ^ neg + - * ; unary ops over the rows of level L-1, binary ops over
// splits i = 1..(L-1)/2 with left rows of level i (outer) and right rows of
// level L-1-i (inner), '-' also trying (r, l).  A candidate is kept iff its
// fingerprint has not been seen before (the intended reading of
// full-cache.cu:122-125, 275-276), the representative of a new fingerprint
// being the first candidate in that order; the solution is the first kept
// candidate equal to the target.  The reference's temp-chunking (:587-599)
// and thrust compaction (:382-409) are replaced by three passes per level.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "context.h"
#include "vfb_kernels.h"

namespace mba {

namespace {

constexpr uint64_t kDefaultCacheBytes = 8ull << 30;

}  // namespace

int vfb_solve(Ctx &ctx, const mba_instance &inst, const mba_opts &o, mba_result &res,
              mba_vfb_stats *stats) {
  const auto t_start = Clock::now();
  std::memset(&res, 0, sizeof res);
  mba_vfb_stats local_stats;
  mba_vfb_stats &S = stats ? *stats : local_stats;
  std::memset(&S, 0, sizeof S);
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
  const int max_len = std::min(o.max_size, kMaxSize);
  if (max_len < 1) {
    std::snprintf(res.expr, sizeof res.expr, "Not found !");
    return MBA_NOT_FOUND;
  }
  const uint64_t M = W == 64 ? ~0ull : 0xffffffffull;
  const size_t wb = (size_t)W / 8;
  std::vector<uint64_t> tgt(n);
  for (int i = 0; i < n; ++i) tgt[i] = inst.outputs[i] & M;

  // ---- level 1 on the host (full-cache.cu:694-707) ------------------------
  std::vector<std::vector<uint64_t>> leaf(L, std::vector<uint64_t>(n));
  for (int v = 0; v < L; ++v)
    for (int i = 0; i < n; ++i)
      leaf[v][i] = (v < nv ? inst.inputs[(size_t)i * nv + v] : inst.consts[v - nv]) & M;
  S.level_new[1] = (uint64_t)L;
  S.level_cand[1] = (uint64_t)L;
  int leaf_sol = -1;
  for (int v = 0; v < L && leaf_sol < 0; ++v) {
    if (leaf[v] == tgt) leaf_sol = v;
  }
  if (leaf_sol >= 0) {
    uint8_t rpn[1] = {(uint8_t)leaf_sol};
    std::snprintf(res.expr, sizeof res.expr, "%s",
                  infix_vfb(rpn, 1, nv, inst.consts, W).c_str());
    res.status = MBA_FOUND;
    res.size = 1;
    res.rank = (uint64_t)leaf_sol;
    res.candidates = (uint64_t)leaf_sol + 1;
    res.size_reached = 1;
    res.seconds = seconds_since(t_start);
    if (!o.exhaustive) return MBA_FOUND;
  }

  // ---- capacity -------------------------------------------------------------
  const uint64_t budget = o.cache_bytes ? o.cache_bytes : kDefaultCacheBytes;
  const uint64_t row_bytes = (uint64_t)n * wb + 9;
  uint64_t cap = std::max<uint64_t>(budget / row_bytes, (uint64_t)L + 1);
  cap = std::min<uint64_t>(cap, (1ull << 31) - 1);  // int32 child indices
  uint64_t slots = 1ull << 16;
  // the table also holds the new fingerprints of the last, hash-only level
  // (about 6x the previous level's rows): 8 slots per stored row
  // at most 2^31 slots: slot indices are stored as u32 with ~0 = none
  while (slots < 8 * cap && slots < (1ull << 31)) slots <<= 1;
  S.table_slots = slots;
  S.rows_capacity = cap;

  hipStream_t st = ctx.stream;
  MBA_HIP(ctx.vfb_rows.ensure(cap * n * wb));
  MBA_HIP(ctx.vfb_left.ensure(cap * 4));
  MBA_HIP(ctx.vfb_right.ensure(cap * 4));
  MBA_HIP(ctx.vfb_flags.ensure(cap));  // row ops
  // slot s = {key at [2s], companion at [2s+1]}: one line per probe
  MBA_HIP(ctx.vfb_keys.ensure(slots * 16));
  MBA_HIP(ctx.vfb_target.ensure(n * 8 + 64));
  MBA_HIP(ctx.vfb_misc.ensure(4096));
  MBA_HIP(vfb_fill_u64(ctx.vfb_keys.as<uint64_t>(), 2 * slots, kVfbEmpty, st));  // = ~0
  {  // leaves -> rows 0..L-1 (SoA), target words
    std::vector<uint64_t> col(L);
    std::vector<uint32_t> col32(L);
    for (int i = 0; i < n; ++i) {
      for (int v = 0; v < L; ++v) col[v] = leaf[v][i], col32[v] = (uint32_t)leaf[v][i];
      char *dst = ctx.vfb_rows.as<char>() + (size_t)i * cap * wb;
      MBA_HIP(hipMemcpyAsync(dst, W == 64 ? (void *)col.data() : (void *)col32.data(),
                             L * wb, hipMemcpyHostToDevice, st));
    }
    std::vector<uint32_t> t32(n);
    for (int i = 0; i < n; ++i) t32[i] = (uint32_t)tgt[i];
    MBA_HIP(hipMemcpyAsync(ctx.vfb_target.p, W == 64 ? (void *)tgt.data() : (void *)t32.data(),
                           n * wb, hipMemcpyHostToDevice, st));
  }
  // misc cells: [0] found (u64), [1] target minidx (u64), [2] fail (u32),
  // [3] probes (u64), [4] rpn length (i32), [64..] rpn tokens
  uint64_t *d_misc = ctx.vfb_misc.as<uint64_t>();
  unsigned long long *d_found = (unsigned long long *)(d_misc + 0);
  uint64_t *d_tmin = d_misc + 1;
  unsigned int *d_fail = (unsigned int *)(d_misc + 2);
  unsigned long long *d_probes = (unsigned long long *)(d_misc + 3);
  int32_t *d_rlen = (int32_t *)(d_misc + 4);
  uint8_t *d_rpn = (uint8_t *)(d_misc + 64);
  MBA_HIP(hipMemsetAsync(d_misc, 0, 64 * 8, st));

  VfbLevelArgs a{};
  std::memset(&a, 0, sizeof a);
  a.rows = ctx.vfb_rows.p;
  a.cap = cap;
  a.n = n;
  a.keys = ctx.vfb_keys.as<uint64_t>();
  a.minidx = ctx.vfb_keys.as<uint64_t>() + 1;  // companion words, stride 2
  a.mask = slots - 1;
  a.target = ctx.vfb_target.p;
  a.found = d_found;
  a.fail = d_fail;
  a.max_probes = std::max<uint64_t>(1024, slots / 4);
  a.out_rows = ctx.vfb_rows.p;
  a.left = ctx.vfb_left.as<int32_t>();
  a.right = ctx.vfb_right.as<int32_t>();
  a.row_op = ctx.vfb_flags.as<uint8_t>();
  a.probes = d_probes;
  unsigned long long *d_new = (unsigned long long *)(d_misc + 5);
  a.new_keys = d_new;
  unsigned long long *d_atomics = (unsigned long long *)(d_misc + 6);  // [6], [7]
  a.atomics = d_atomics;
  MBA_HIP(vfb_launch_seed(W, a, L, st));

  std::vector<uint64_t> lstart(kMaxSize + 3, 0);
  lstart[1] = 0;
  lstart[2] = (uint64_t)L;
  uint64_t rows = (uint64_t)L;
  const int grid = ctx.cu_count * 8;
  double ktime = 0.0;
  uint32_t klaunch = 0;
  int status = MBA_NOT_FOUND;
  int reached = 1;
  uint64_t cand_total = (uint64_t)L;
  int64_t sol_idx = leaf_sol >= 0 ? -2 : -1;  // -2: the solution is a leaf
  int sol_level = 0;
  std::vector<VfbBlock> blocks;

  auto over_budget = [&]() {
    if (o.cancel && *o.cancel) return true;
    return o.time_budget_s > 0 && seconds_since(t_start) > o.time_budget_s;
  };

  for (int lev = 2; lev <= max_len; ++lev) {
    if (over_budget()) {
      status = MBA_BUDGET;
      break;
    }
    const auto t_lev = Clock::now();
    blocks.clear();
    uint64_t total = 0, stream_bytes = 0;
    for (int op = 0; op < 8; ++op) {
      if (is_unary(op)) {
        const uint64_t nl = lstart[lev] - lstart[lev - 1];
        if (nl == 0) continue;
        blocks.push_back({total, nl, (uint32_t)lstart[lev - 1], (uint32_t)nl, 0, 0, op, 0});
        total += nl;
        stream_bytes += nl * n * wb;
        continue;
      }
      for (int i = 1; 2 * i <= lev - 1; ++i) {
        const uint64_t nl = lstart[i + 1] - lstart[i];
        const uint64_t nr = lstart[lev - i] - lstart[lev - 1 - i];
        const uint64_t cnt = nl * nr * (op == SUB ? 2 : 1);
        if (cnt == 0) continue;
        blocks.push_back({total, cnt, (uint32_t)lstart[i], (uint32_t)nl,
                          (uint32_t)lstart[lev - 1 - i], (uint32_t)nr, op, 0});
        total += cnt;
        stream_bytes += (nl + nr) * n * wb;
      }
    }
    if ((int)blocks.size() > kVfbMaxBlocks) {
      ctx.last_error = "too many enumeration blocks in one level";
      return MBA_ERR_LIMIT;
    }
    // row-pair view of the same blocks (one per split, plus the unary rows)
    std::vector<VfbPairBlock> pblocks;
    uint64_t npairs = 0;
    for (const VfbBlock &B : blocks) {
      const bool un = is_unary(B.op);
      VfbPairBlock *pb = nullptr;
      for (VfbPairBlock &x : pblocks)
        if (x.unary == (int)un && x.l0 == B.l0 && x.nl == B.nl && x.r0 == B.r0 && x.nr == B.nr)
          pb = &x;
      if (!pb) {
        VfbPairBlock nb;
        std::memset(&nb, 0, sizeof nb);
        nb.pbegin = npairs;
        nb.l0 = B.l0;
        nb.nl = B.nl;
        nb.r0 = B.r0;
        nb.nr = B.nr;
        nb.unary = un;
        npairs += un ? (uint64_t)B.nl : (uint64_t)B.nl * B.nr;
        pblocks.push_back(nb);
        pb = &pblocks.back();
      }
      pb->opbegin[B.op] = B.begin;
    }
    if ((int)pblocks.size() > kVfbMaxPairBlocks) {
      ctx.last_error = "too many row-pair blocks in one level";
      return MBA_ERR_LIMIT;
    }
    if (total == 0) {
      lstart[lev + 1] = rows;
      reached = lev;
      continue;
    }
    if (total >= (1ull << 48)) {
      ctx.last_error = "level candidate count above 2^48";
      return MBA_ERR_LIMIT;
    }
    const uint64_t nchunks = (total + kVfbChunk - 1) / kVfbChunk;
    MBA_HIP(ctx.vfb_counts.ensure(total * 4 + 64));            // slot_of
    MBA_HIP(ctx.b_misc.ensure(total + 64));                    // keep flags
    MBA_HIP(ctx.b_queue.ensure((nchunks + 1) * 12 + 64));      // chunk count/base
    MBA_HIP(ctx.b_inst.ensure(sizeof(VfbBlock) * blocks.size()));
    MBA_HIP(hipMemcpyAsync(ctx.b_inst.p, blocks.data(), sizeof(VfbBlock) * blocks.size(),
                           hipMemcpyHostToDevice, st));
    a.blocks = ctx.b_inst.as<VfbBlock>();
    MBA_HIP(ctx.b_results.ensure(sizeof(VfbPairBlock) * pblocks.size() + 64));
    MBA_HIP(hipMemcpyAsync(ctx.b_results.p, pblocks.data(),
                           sizeof(VfbPairBlock) * pblocks.size(), hipMemcpyHostToDevice, st));
    a.nblocks = (int)blocks.size();
    a.total = total;
    a.tag = (uint64_t)lev << 48;
    a.slot_of = ctx.vfb_counts.as<uint32_t>();
    a.keep = ctx.b_misc.as<uint8_t>();
    a.chunk_base = ctx.b_queue.as<uint64_t>();
    a.chunk_count = (uint32_t *)(ctx.b_queue.as<uint64_t>() + nchunks + 1);
    a.out_base = rows;
    const int g = (int)std::min<uint64_t>((uint64_t)grid, nchunks);
    const bool last = lev == max_len;
    a.last = last ? 1 : 0;
    MBA_HIP(hipMemsetAsync(d_found, 0xff, 8, st));
    MBA_HIP(hipMemsetAsync(d_new, 0, 24, st));  // new keys, CAS attempts, atomicMin
    MBA_HIP(hipEventRecord(ctx.ev0, st));
    {
      const uint64_t pgroups = (npairs + kVfbThreads - 1) / kVfbThreads;
      const int gp = (int)std::min<uint64_t>((uint64_t)grid, std::max<uint64_t>(pgroups, 1));
      MBA_HIP(vfb_launch_insert_pairs(W, a, ctx.b_results.as<VfbPairBlock>(),
                                      (int)pblocks.size(), npairs, gp, st));
    }
    MBA_HIP(vfb_launch_lookup_target(W, a, d_tmin, st));
    uint64_t cells[8];
    MBA_HIP(hipMemcpyAsync(cells, d_misc, 64, hipMemcpyDeviceToHost, st));
    MBA_HIP(hipStreamSynchronize(st));
    if ((uint32_t)cells[2]) {
      status = MBA_OVERFLOW;
      break;
    }
    cand_total += total;
    S.level_cand[lev] = total;
    S.level_cas[lev] = cells[6];
    S.level_atomics[lev] = cells[6] + cells[7];
    const uint64_t found = cells[0], tmin = cells[1];
    if (sol_idx == -1 && found != ~0ull && tmin == (a.tag | found)) {
      sol_idx = (int64_t)found;
      sol_level = lev;
      reached = lev;
      if (!o.exhaustive) {
        float ms = 0.f;
        MBA_HIP(hipEventRecord(ctx.ev1, st));
        MBA_HIP(hipEventSynchronize(ctx.ev1));
        MBA_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
        ktime += ms * 1e-3;
        ++klaunch;
        break;
      }
    }
    // kept rows: the flag pass marks each new fingerprint's first candidate
    // and the scan places them; the hash-only last level only counts the
    // fingerprints its insert pass added
    uint64_t kept = cells[5];
    if (!last) {
      MBA_HIP(vfb_launch_flag(W, a, g, st));
      MBA_HIP(vfb_launch_scan(a, nchunks, st));
      MBA_HIP(hipMemcpyAsync(&kept, a.chunk_base + nchunks, 8, hipMemcpyDeviceToHost, st));
      MBA_HIP(hipStreamSynchronize(st));
    }
    bool overflow = false;
    if (!last) {
      if (rows + kept > cap) {
        overflow = true;
      } else {
        MBA_HIP(vfb_launch_write(W, a, g, st));
      }
    }
    MBA_HIP(hipEventRecord(ctx.ev1, st));
    MBA_HIP(hipEventSynchronize(ctx.ev1));
    float ms = 0.f;
    MBA_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
    ktime += ms * 1e-3;
    ++klaunch;
    S.level_new[lev] = kept;
    S.level_seconds[lev] = seconds_since(t_lev);
    // algorithmic bytes: one slot read per candidate, a key write per new
    // fingerprint, row + child indices per stored row, child rows streamed
    // once per block (the last level stores no rows)
    S.level_bytes[lev] = total * 8 + kept * 8 + (last ? 0 : kept * (n * wb + 8)) + stream_bytes;
    if (o.verbose)
      std::fprintf(stderr, "Length %-2d | candidates %-14llu | new %-12llu | %.3f s\n", lev,
                   (unsigned long long)total, (unsigned long long)kept,
                   seconds_since(t_start));
    if (overflow) {
      status = MBA_OVERFLOW;
      break;
    }
    if (!last) rows += kept;
    lstart[lev + 1] = rows;
    reached = lev;
  }

  res.kernel_seconds = ktime;
  res.kernel_launches = klaunch;
  res.candidates = cand_total;
  res.size_reached = reached;
  if (sol_idx >= 0) {
    // decode the candidate on the host from the level's block table
    const uint64_t idx = (uint64_t)sol_idx;
    int b = 0;
    while (b + 1 < (int)blocks.size() && idx >= blocks[b + 1].begin) ++b;
    // blocks may have been rebuilt for later levels when exhaustive: rebuild
    // the solution level's table if needed
    if (o.exhaustive && sol_level != max_len) {
      blocks.clear();
      uint64_t total = 0;
      for (int op = 0; op < 8; ++op) {
        if (is_unary(op)) {
          const uint64_t nl = lstart[sol_level] - lstart[sol_level - 1];
          if (nl) blocks.push_back({total, nl, (uint32_t)lstart[sol_level - 1], (uint32_t)nl, 0, 0, op, 0});
          total += nl;
          continue;
        }
        for (int i = 1; 2 * i <= sol_level - 1; ++i) {
          const uint64_t nl = lstart[i + 1] - lstart[i];
          const uint64_t nr = lstart[sol_level - i] - lstart[sol_level - 1 - i];
          const uint64_t cnt = nl * nr * (op == SUB ? 2 : 1);
          if (cnt) blocks.push_back({total, cnt, (uint32_t)lstart[i], (uint32_t)nl,
                                     (uint32_t)lstart[sol_level - 1 - i], (uint32_t)nr, op, 0});
          total += cnt;
        }
      }
      b = 0;
      while (b + 1 < (int)blocks.size() && idx >= blocks[b + 1].begin) ++b;
    }
    const VfbBlock &B = blocks[b];
    uint64_t local = idx - B.begin;
    int32_t l, r = -1;
    if (is_unary(B.op)) {
      l = (int32_t)(B.l0 + local);
    } else {
      int sw = 0;
      if (B.op == SUB) {
        sw = (int)(local & 1);
        local >>= 1;
      }
      l = (int32_t)(B.l0 + local / B.nr);
      r = (int32_t)(B.r0 + local % B.nr);
      if (sw) std::swap(l, r);
    }
    const int cap_tok = 4096 - 64 * 8;
    MBA_HIP(vfb_launch_rpn(a.left, a.right, a.row_op, L, B.op, l, r, d_rpn, d_rlen, cap_tok,
                           st));
    int32_t len = 0;
    std::vector<uint8_t> tok(cap_tok);
    MBA_HIP(hipMemcpyAsync(&len, d_rlen, 4, hipMemcpyDeviceToHost, st));
    MBA_HIP(hipMemcpyAsync(tok.data(), d_rpn, cap_tok, hipMemcpyDeviceToHost, st));
    MBA_HIP(hipStreamSynchronize(st));
    res.status = MBA_FOUND;
    res.size = sol_level;
    res.rank = idx;
    res.op = B.op;
    // The reference collects the non-root internal nodes, 3 ints each, in a
    // 100-int buffer and gives up from 33 of them (full-cache.cu:420,
    // 544-551); trees of at most 32 nodes never reach that.
    int internal = 0;
    for (int i = 0; i < len && i < cap_tok; ++i) internal += tok[i] >= kTokOp;
    if (internal - 1 >= 33 || len > cap_tok) {
      std::snprintf(res.expr, sizeof res.expr, "Size of the output is too big !");
    } else {
      std::snprintf(res.expr, sizeof res.expr, "%s",
                    infix_vfb(tok.data(), len, nv, inst.consts, W).c_str());
    }
    status = MBA_FOUND;
  } else if (sol_idx == -2) {  // a leaf, found before the levels (exhaustive run)
    status = MBA_FOUND;
    res.size = 1;
    res.rank = (uint64_t)leaf_sol;
  } else {
    std::snprintf(res.expr, sizeof res.expr, "Not found !");
  }
  res.status = status;
  res.seconds = seconds_since(t_start);
  return status;
}

int vfb_fingerprints(Ctx &ctx, int W, const uint64_t *values, int n, size_t count,
                     uint64_t *out) {
  if ((W != 32 && W != 64) || n < 1 || n > kMaxSamples || (count && (!values || !out))) {
    ctx.last_error = "fingerprints: width 32 or 64, 1..128 words per signature";
    return MBA_ERR_ARG;
  }
  if (count == 0) return 0;
  const size_t in_bytes = count * (size_t)n * 8;
  MBA_HIP(ctx.f_in.ensure(in_bytes));
  MBA_HIP(ctx.f_out.ensure(count * 8));
  hipStream_t st = ctx.stream;
  MBA_HIP(hipMemcpyAsync(ctx.f_in.p, values, in_bytes, hipMemcpyHostToDevice, st));
  MBA_HIP(vfb_launch_fingerprints(W, ctx.f_in.as<uint64_t>(), n, count, ctx.f_out.as<uint64_t>(),
                                  st));
  MBA_HIP(hipMemcpyAsync(out, ctx.f_out.p, count * 8, hipMemcpyDeviceToHost, st));
  MBA_HIP(hipStreamSynchronize(st));
  return 0;
}

int vfb_probe_ceiling(Ctx &ctx, int slots_log2, uint64_t count, uint64_t distinct,
                      uint64_t prefill, double *probe_s, double *insert_s, uint64_t *fresh) {
  if (slots_log2 < 10 || slots_log2 > 31 || count == 0 || distinct == 0 ||
      count + prefill > (1ull << 40)) {
    ctx.last_error = "probe ceiling: 10 <= slots_log2 <= 31, count >= 1, distinct >= 1";
    return MBA_ERR_ARG;
  }
  const uint64_t slots = 1ull << slots_log2;
  const size_t need = (size_t)(16 * slots + 8 * (count + prefill) + 64);
  if (ctx.hbm_bytes && need > ctx.hbm_bytes / 10 * 8) {
    ctx.last_error = "probe ceiling: table and keys exceed 80% of device memory";
    return MBA_ERR_LIMIT;
  }
  hipStream_t st = ctx.stream;
  void *mem = nullptr;
  MBA_HIP(hipMalloc(&mem, need));
  uint64_t *table = (uint64_t *)mem;
  uint64_t *keys = table + 2 * slots, *pre = keys + count;
  unsigned long long *cnt = (unsigned long long *)(pre + prefill);  // new keys, fail, occupied
  VfbLevelArgs a;
  std::memset(&a, 0, sizeof a);
  a.keys = table;
  a.minidx = table + 1;
  a.mask = slots - 1;
  a.max_probes = slots;
  a.last = 1;
  a.new_keys = cnt;
  a.fail = (unsigned int *)(cnt + 1);
  const int grid = ctx.cu_count * 8;
  float ms_probe = 0.f, ms_insert = 0.f;
  unsigned long long host[3] = {0, 0, 0};
  hipError_t e = vfb_launch_ceiling(0, a, keys, count, distinct, 1, nullptr, grid, st);
  if (e == hipSuccess)  // prefill keys: distinct values of their own (another salt)
    e = vfb_launch_ceiling(0, a, pre, prefill, ~0ull >> 2, 1ull << 62, nullptr, grid, st);
  // the measured pass on a table holding the prefill keys only
  auto reset = [&]() -> hipError_t {
    hipError_t r = vfb_fill_u64(table, 2 * slots, kVfbEmpty, st);
    if (r == hipSuccess) r = hipMemsetAsync(cnt, 0, 24, st);
    if (r == hipSuccess) r = vfb_launch_ceiling(2, a, pre, prefill, 0, 0, nullptr, grid, st);
    if (r == hipSuccess) r = hipMemsetAsync(cnt, 0, 24, st);
    return r;
  };
  if (e == hipSuccess) e = reset();
  if (e == hipSuccess) e = hipEventRecord(ctx.ev0, st);
  if (e == hipSuccess) e = vfb_launch_ceiling(1, a, keys, count, 0, 0, cnt + 2, grid, st);
  if (e == hipSuccess) e = hipEventRecord(ctx.ev1, st);
  if (e == hipSuccess) e = hipEventSynchronize(ctx.ev1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms_probe, ctx.ev0, ctx.ev1);
  if (e == hipSuccess) e = reset();
  if (e == hipSuccess) e = hipEventRecord(ctx.ev0, st);
  if (e == hipSuccess) e = vfb_launch_ceiling(2, a, keys, count, 0, 0, nullptr, grid, st);
  if (e == hipSuccess) e = hipEventRecord(ctx.ev1, st);
  if (e == hipSuccess) e = hipEventSynchronize(ctx.ev1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms_insert, ctx.ev0, ctx.ev1);
  if (e == hipSuccess) e = hipMemcpy(host, cnt, 24, hipMemcpyDeviceToHost);
  const hipError_t ef = hipFree(mem);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) {
    ctx.last_error = std::string("probe ceiling: ") + hipGetErrorString(e);
    return MBA_ERR_HIP;
  }
  if ((uint32_t)host[1]) {
    ctx.last_error = "probe ceiling: table too full";
    return MBA_OVERFLOW;
  }
  if (probe_s) *probe_s = ms_probe * 1e-3;
  if (insert_s) *insert_s = ms_insert * 1e-3;
  if (fresh) *fresh = host[0];
  return 0;
}

}  // namespace mba
