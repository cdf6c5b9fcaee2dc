// batch_host.cpp -- batched SIMBA search: many instances of uneven shape
// (variables, samples, target size) solved together.
//
// Every size step of every active instance becomes a list of jobs (value
// table builds, fused pair checks, leaf scans) in a device work queue; one
// persistent kernel per step drains it (workgroups pull tiles from an atomic
// counter), so small and large instances share the GPU.  The per-step launch
// sequence -- step header upload, counter reset, persistent kernel, hit-count
// read back -- is captured once into a hipGraph and replayed per step (the
// step's jobs are copied in just before).  Window hits are resolved on the
// device (k_hit_resolve: every sample, chain DP, per-slot minimum rank) and
// the lowered slots' trees verified.  Each instance's answer is the one
// mba_simba_solve returns on it: the same tables, windows and candidate
// order per instance.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <vector>

#include "context.h"
#include "simba_host.h"

namespace mba {

namespace {

constexpr uint32_t kBatchHitCap = 1u << 26;  // 2 GiB of hit records
// Workgroups (one wave each) of the all-sample candidate check: the kernel
// is latency-bound (a serial unrank per candidate), so more waves than fit
// at once keep every SIMD busy while the survivors last
constexpr uint32_t kConfirmBlocks = 8192;
// default value-table budget of a batch, split evenly over its instances
// (at most 2/5 of the device memory): 4096 instances of up to 4 leaves keep
// V[8] in tables at a size budget of 11, so the last size's B sides are
// formed from tabulated children
constexpr uint64_t kBatchTableBytes = 96ull << 30;
// the table pool is packed again (see table_steps) from this size on
constexpr int kRepackMin = 8;

struct InstState {
  int L = 0, n = 0, fs = 0, Kt = 1;
  uint64_t t0 = 0;
  uint64_t elems = 0;            // value-table elements V[1..Kt]
  const CountTable *T = nullptr;
  uint64_t *found = nullptr;     // verified minimum rank per size
  bool active = true;
  int reached = 0;
};

// Runs f(lo, hi) over [0, n) in up to kSetupThreads host threads.
constexpr size_t kSetupThreads = 8;
constexpr uint64_t kMinAllowance = 256;  // budgeted batches: tiles an instance always gets
template <typename F>
void parallel_ranges(size_t n, F f) {
  const size_t nt = n >= 1024 ? kSetupThreads : 1;
  if (nt == 1) {
    f((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t) th.emplace_back(f, n * t / nt, n * (t + 1) / nt);
  for (auto &x : th) x.join();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

int batch_solve(Ctx &ctx, const mba_instance *insts, size_t count, const mba_opts &o,
                mba_result *results, const std::function<int(int, float *)> *score_size) {
  const auto t_start = Clock::now();
  if (count == 0) return 0;
  const int W = (int)insts[0].width;
  int min_size = o.min_size < 1 ? 1 : o.min_size;
  int max_size = o.max_size;
  if (o.compat) {
    min_size = 2;
    max_size = o.max_size - 1;
  }
  if (max_size > kMaxSize || (W != 32 && W != 64)) {
    ctx.last_error = "batch: unsupported width or size";
    return MBA_ERR_LIMIT;
  }
  const uint64_t M = wmask(W);
  const size_t wb = (size_t)W / 8;
  std::map<int, CountTable> tables_by_L;
  std::vector<InstState> S(count);
  std::vector<uint64_t> found_all(count * (kMaxSize + 2), ~0ull);
  std::vector<uint64_t> leaf_off(count), out_off(count);
  const uint64_t table_budget =
      o.table_bytes ? (uint64_t)o.table_bytes
                    : std::min<uint64_t>(kBatchTableBytes, ctx.hbm_bytes / 5 * 2);
  // tables up to this size (V[max_size - 2] is only ever the B side of
  // leaves: formed per tile)
  const int kt_cap = std::max(1, max_size - 3 - std::max(0, o.table_delta));
  // pass 1 (in order): limits, count tables per leaf count, sample offsets
  uint64_t nleaf = 0, nout = 0;
  for (size_t i = 0; i < count; ++i) {
    const mba_instance &in = insts[i];
    const int nv = (int)in.n_vars, nc = (int)in.n_consts, n = (int)in.n_samples;
    const int L = nv + nc;
    if ((int)in.width != W || L < 1 || L > kMaxLeaves || n < 1 || n > kMaxSamples ||
        !in.inputs || !in.outputs || (nc > 0 && !in.consts)) {
      ctx.last_error = "batch: instance outside supported limits or mixed widths";
      return MBA_ERR_LIMIT;
    }
    auto it = tables_by_L.find(L);
    if (it == tables_by_L.end()) {
      CountTable T;
      if (!T.build(L, std::max(max_size, 1))) {
        ctx.last_error = "count table overflows uint64 at this max size";
        return MBA_ERR_COUNT;
      }
      it = tables_by_L.emplace(L, T).first;
    }
    InstState &st = S[i];
    st.L = L;
    st.n = n;
    st.T = &it->second;
    st.found = found_all.data() + i * (kMaxSize + 2);
    st.active = max_size >= min_size;
    leaf_off[i] = nleaf;
    out_off[i] = nout;
    nleaf += (uint64_t)n * L;
    nout += (uint64_t)n;
  }
  // pinned staging (kept in the context, grown when needed): step headers
  // at 0, step counters at 1536, then the instances' table views, leaf
  // values, outputs and sample layouts
  // jobs of every size of a search (a group of sizes runs as one launch
  // sequence): the leaf scan, a table build per tabulated size, the pairs
  size_t jobs_per_inst = 1;
  for (int z = 2; z <= max_size; ++z) jobs_per_inst += 1 + (size_t)std::max(0, (z - 1) / 2);
  // (B sides above Kt add a job per B block: the host arrays grow with them)
  const size_t job_cap = count * jobs_per_inst;
  const size_t off_tvs = 2048;
  const size_t off_leaf = align256(off_tvs + sizeof(TableView) * count);
  const size_t off_out = align256(off_leaf + 8 * nleaf);
  const size_t off_idat = align256(off_out + 8 * nout);
  const size_t pin_bytes = off_idat + sizeof(BatchInstData) * count;
  hipStream_t stq = ctx.stream;
  if (ctx.b_pinned_bytes < pin_bytes) {
    MBA_HIP(hipStreamSynchronize(stq));
    if (ctx.b_pinned) MBA_HIP(hipHostFree(ctx.b_pinned));
    ctx.b_pinned = nullptr;
    ctx.b_pinned_bytes = 0;
    const size_t want = pin_bytes + pin_bytes / 4;
    MBA_HIP(hipHostMalloc(&ctx.b_pinned, want, hipHostMallocDefault));
    ctx.b_pinned_bytes = want;
  }
  char *pin = (char *)ctx.b_pinned;
  // step headers of a group (one per size, plus the scratch header of a
  // step run in parts) at 0, the step counters read back at 1536
  BatchStep *h_steps = (BatchStep *)pin;
  static_assert(sizeof(BatchStep) * (kMaxSize + 2) <= 1536, "step headers");
  BatchCounters *h_cnt = (BatchCounters *)(pin + 1536);
  // the jobs of a group and the B blocks they name (block 0: a placeholder
  // for jobs whose B side is a table).  A B side above the tables is a list
  // of blocks that depends on the leaf count, the tabulated depth and the B
  // size only (table references are resolved through each instance's view),
  // so each list is built once per call and named by every job of that shape.
  std::vector<BatchJob> jobv(job_cap);
  BatchJob *h_jobs = jobv.data();
  std::vector<BBlock> blkv(1), blk_tmp;
  struct BList {
    uint32_t first, n;  // blocks [first, first + n) of blkv
    uint64_t tiles;     // B tiles of the list
  };
  std::map<uint64_t, BList> blists;
  auto block_list = [&](const InstState &st, int m) -> BList {
    const uint64_t key = (uint64_t)st.L | (uint64_t)st.Kt << 8 | (uint64_t)m << 16;
    auto it = blists.find(key);
    if (it != blists.end()) return it->second;
    blk_tmp.clear();
    b_blocks(*st.T, st.Kt, m, 0, 1, {}, kPairRB, blk_tmp);
    BList bl{(uint32_t)blkv.size(), (uint32_t)blk_tmp.size(), 0};
    for (BBlock &b : blk_tmp) {
      b.tile_pre = bl.tiles;
      bl.tiles += block_tiles(b, kPairRB);
      blkv.push_back(b);
    }
    blists.emplace(key, bl);
    return bl;
  };
  TableView *tvs = (TableView *)(pin + off_tvs);
  uint64_t *leafvals = (uint64_t *)(pin + off_leaf);
  uint64_t *outputs = (uint64_t *)(pin + off_out);
  BatchInstData *idat = (BatchInstData *)(pin + off_idat);
  // pass 2 (host threads): samples, filter sample, tables, table view
  parallel_ranges(count, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      const mba_instance &in = insts[i];
      InstState &st = S[i];
      const int nv = (int)in.n_vars, nc = (int)in.n_consts, n = st.n, L = st.L;
      mba_result &r = results[i];
      std::memset(&r, 0, sizeof r);
      r.status = MBA_NOT_FOUND;
      r.op = -1;
      std::snprintf(r.expr, sizeof r.expr, "Not found !");
      uint64_t *lv = leafvals + leaf_off[i], *outv = outputs + out_off[i];
      for (int smp = 0; smp < n; ++smp) {
        for (int v = 0; v < nv; ++v)
          lv[(size_t)smp * L + v] = in.inputs[(size_t)smp * nv + v] & M;
        for (int c = 0; c < nc; ++c) lv[(size_t)smp * L + nv + c] = in.consts[c] & M;
        outv[smp] = in.outputs[smp] & M;
      }
      st.fs = pick_filter_sample(lv, outv, n, L, W);
      st.t0 = outv[st.fs];
      // V[1] only; larger tables are placed as the sweep reaches them
      st.Kt = 1;
      TableView &tv = tvs[i];
      std::memset(&tv, 0, sizeof tv);
      uint64_t e = 0;
      for (int m = 2; m <= kt_cap; ++m) e += st.T->t[m][8];
      st.elems = e;  // every table the search could place
      std::memcpy(tv.T, st.T->t, sizeof tv.T);
      tv.Kt = 1;
      tv.L = L;
      idat[i].leaf_off = leaf_off[i];
      idat[i].out_off = out_off[i];
      idat[i].n_samples = (uint32_t)n;
      idat[i].fs = (uint32_t)st.fs;
    }
  });
  // pass 3: the value-table pool: every instance's V[1] first, then tables
  // placed as the sweep reaches their size (place_table)
  uint64_t pool_used = 0, pool_need = 0;
  for (size_t i = 0; i < count; ++i) {
    tvs[i].off[1] = pool_used;
    pool_used += (uint64_t)S[i].L;
    pool_need += S[i].elems;
  }
  const uint64_t pool_elems = pool_used + std::min<uint64_t>(pool_need, table_budget / wb);
  const bool pool_all = pool_need <= table_budget / wb;  // room for every table
  MBA_HIP(ctx.b_data.ensure(std::max<size_t>(pool_elems * wb, 256)));
  for (size_t i = 0; i < count; ++i) tvs[i].base = ctx.b_data.p;
  MBA_HIP(ctx.b_inst.ensure(sizeof(TableView) * count));
  MBA_HIP(ctx.leafvals.ensure(8 * std::max<uint64_t>(nleaf, 1)));
  MBA_HIP(ctx.outputs.ensure(8 * std::max<uint64_t>(nout, 1)));
  MBA_HIP(ctx.b_results.ensure(sizeof(BatchInstData) * count));
  MBA_HIP(hipMemcpyAsync(ctx.b_inst.p, tvs, sizeof(TableView) * count, hipMemcpyHostToDevice,
                         stq));
  MBA_HIP(hipMemcpyAsync(ctx.leafvals.p, leafvals, 8 * nleaf, hipMemcpyHostToDevice, stq));
  MBA_HIP(hipMemcpyAsync(ctx.outputs.p, outputs, 8 * nout, hipMemcpyHostToDevice, stq));
  MBA_HIP(hipMemcpyAsync(ctx.b_results.p, idat, sizeof(BatchInstData) * count,
                         hipMemcpyHostToDevice, stq));
  // V[1] of every instance: leaf values of its filter sample (one kernel)
  MBA_HIP(launch_batch_leaves(W, ctx.b_inst.as<TableView>(), ctx.leafvals.as<uint64_t>(),
                              ctx.b_results.as<BatchInstData>(), (uint32_t)count, stq));

  // ---- device state ---------------------------------------------------------
  // hit buffer (opts.hit_cap; smaller buffers exercise the split replays) and
  // its screened candidates (at most two per hit)
  const uint32_t hit_cap = o.hit_cap > 0 ? (uint32_t)o.hit_cap : kBatchHitCap;
  MBA_HIP(ctx.hits.ensure(sizeof(HitRec) * hit_cap));
  MBA_HIP(ctx.surv.ensure(sizeof(SurvRec) * 2 * (size_t)hit_cap));
  MBA_HIP(ctx.hit_count.ensure(256));
  BatchCounters *d_cnt = ctx.hit_count.as<BatchCounters>();
  unsigned long long *d_head = (unsigned long long *)(ctx.hit_count.as<char>() + 64);
  HitRec *d_hits = ctx.hits.as<HitRec>();
  const TableView *d_tvs = ctx.b_inst.as<TableView>();
  // device copies of a group: step headers, jobs, blocks (grown with a
  // group that holds more)
  BatchStep *d_steps = nullptr;
  BatchJob *d_jobs = nullptr;
  BBlock *d_blocks = nullptr;
  size_t dev_jobs = 0, dev_blocks = 0;
  auto queue_room = [&](size_t nj, size_t nb) -> int {
    if (d_steps && nj <= dev_jobs && nb <= dev_blocks) return 0;
    dev_jobs = std::max(dev_jobs, std::max(nj, job_cap));
    dev_blocks = std::max(dev_blocks, std::max<size_t>(nb, 4096));
    const size_t off_j = align256(sizeof(BatchStep) * (kMaxSize + 2));
    const size_t off_b = align256(off_j + sizeof(BatchJob) * dev_jobs);
    MBA_HIP(hipStreamSynchronize(stq));
    MBA_HIP(ctx.b_queue.ensure(off_b + sizeof(BBlock) * dev_blocks));
    d_steps = ctx.b_queue.as<BatchStep>();
    d_jobs = (BatchJob *)(ctx.b_queue.as<char>() + off_j);
    d_blocks = (BBlock *)(ctx.b_queue.as<char>() + off_b);
    return 0;
  };
  if (int rc = queue_room(job_cap, 4096)) return rc;
  // Orders a step's jobs [j0, j1) as its kernels take them (see BatchStep):
  // tabulated B sides, table builds and leaf scans first, then the B sides
  // formed from blocks, each part in its former order; renumbers their
  // tiles from 0 and returns the tiles of the first part.
  auto order_step = [&](size_t j0, size_t j1) -> uint64_t {
    auto tabulated = [](const BatchJob &J) { return !(J.kind == 1 && J.p.gen_b); };
    std::stable_partition(h_jobs + j0, h_jobs + j1, tabulated);
    uint64_t t = 0, first = ~0ull;
    for (size_t k = j0; k < j1; ++k) {
      if (first == ~0ull && !tabulated(h_jobs[k])) first = t;
      h_jobs[k].tile0 = t;
      t += h_jobs[k].ntiles;
    }
    return first == ~0ull ? t : first;
  };
  // per-(instance, size) slots: the verified minimum rank, the slot flags and
  // the list of lowered slots of the resolution kernels; per instance, the
  // smallest size settled on the device
  const size_t nslots = count * (kMaxSize + 1);
  MBA_HIP(ctx.b_misc.ensure(nslots * 16 + count * 4 + 256));
  unsigned long long *d_found = (unsigned long long *)(ctx.b_misc.as<char>() + 256);
  uint32_t *d_flags = (uint32_t *)(d_found + nslots);
  uint32_t *d_touched = d_flags + nslots;
  int32_t *d_settled = (int32_t *)(d_touched + nslots);
  MBA_HIP(hipMemsetAsync(d_found, 0xff, nslots * 8, stq));
  MBA_HIP(hipMemsetAsync(d_flags, 0, nslots * 4, stq));
  MBA_HIP(hipMemsetAsync(d_settled, 0, count * 4, stq));
  MBA_HIP(hipMemsetAsync(d_cnt, 0, sizeof(BatchCounters), stq));
  const int mult = o.wg_per_cu > 0 ? o.wg_per_cu : 8;  // workgroups per CU
  const int grid = ctx.cu_count * mult;
  // learned order from sizes of this many trees (opts.block_log2 lowers it)
  const uint64_t learned_min =
      o.block_log2 > 0 ? (1ull << std::min(o.block_log2, 63)) : kLearnedBlockMin;

  double ktime = 0.0;
  uint32_t klaunch = 0;
  int replays = 0;  // steps run again in parts after a hit buffer overflow
  double host_build = 0.0, host_collect = 0.0;  // host time (verbose report)
  size_t blocks_sent = 0;
  bool budget_hit = false;
  std::vector<double> own(count, 0.0);  // each instance's share of device time
  std::vector<uint64_t> inst_tiles(count, 0);
  std::vector<int32_t> settled(count, 0);
  std::vector<uint64_t> found_h(nslots);
  auto finish = [&](size_t i, int z) {
    InstState &st = S[i];
    fill_result(*st.T, z, st.found[z], (int)insts[i].n_vars, insts[i].consts, W, results[i]);
    results[i].seconds = seconds_since(t_start);
    st.reached = z;
    st.active = false;
  };
  std::vector<std::vector<int>> order(count);  // learned: the splits in proposed order
  std::vector<float> scores(score_size ? count * 256 : 0);
  if (o.verbose)
    std::fprintf(stderr, "batch: %zu instances set up in %.3f ms\n", count,
                 1e3 * seconds_since(t_start));

  // table placements of the group being built, applied on the device before
  // it runs (every record of an instance carries its final Kt)
  std::vector<TvUpdate> tv_updates;
  size_t n_active_now = count;
  bool leaves_pending = false;  // V[1] of every view to rewrite (pool packed again)
  auto flush_views = [&]() -> int {
    if (tv_updates.empty()) return 0;
    for (TvUpdate &u : tv_updates) u.kt = S[u.inst].Kt;
    MBA_HIP(ctx.b_tvup.ensure(sizeof(TvUpdate) * tv_updates.size()));
    MBA_HIP(hipMemcpyAsync(ctx.b_tvup.p, tv_updates.data(), sizeof(TvUpdate) * tv_updates.size(),
                           hipMemcpyHostToDevice, stq));
    MBA_HIP(launch_tv_update(ctx.b_inst.as<TableView>(), ctx.b_tvup.as<TvUpdate>(),
                             (uint32_t)tv_updates.size(), stq));
    MBA_HIP(hipStreamSynchronize(stq));  // the record vector is reused
    tv_updates.clear();
    return 0;
  };
  bool pool_overrun = false;  // a table placed past the pool (never: reported as an error)
  auto flush_tables = [&]() -> int {
    if (pool_overrun) {
      ctx.last_error = "batch: a value table placed past the table pool";
      return MBA_ERR_LIMIT;
    }
    if (leaves_pending) {  // after the views: V[1] rewritten where the pool was packed again
      leaves_pending = false;
      if (int rc = flush_views()) return rc;
      MBA_HIP(launch_batch_leaves(W, ctx.b_inst.as<TableView>(), ctx.leafvals.as<uint64_t>(),
                                  ctx.b_results.as<BatchInstData>(), (uint32_t)count, stq));
      return 0;
    }
    return flush_views();
  };

  // The jobs of size s for instance i (learned rounds: the splits of
  // `splits` -- (split, instance mask) pairs -- else every split with every
  // operator instance; table build and leaf scan in the first
  // round), appended to h_jobs[nj..] with step-relative first tiles.  Only
  // the pair tiles [win_from, win_from + win_len) of the size (counted over
  // its splits in order) are taken, the table build and leaf scan with the
  // window at 0; returns the end of the window when tiles remain past it,
  // else 0 (the size is complete).
  auto size_jobs = [&](size_t i, int s, bool first_round,
                       const std::vector<std::pair<int, uint32_t>> *splits,
                       size_t &nj, uint64_t &tiles, uint64_t win_from = 0,
                       uint64_t win_len = ~0ull) -> uint64_t {
    InstState &st = S[i];
    const CountTable &T = *st.T;
    const uint64_t win_end = win_len > ~0ull - win_from ? ~0ull : win_from + win_len;
    uint64_t loc = 0;  // pair tiles of the size before the current job
    auto push = [&](BatchJob &J) -> bool {
      if (J.kind == 1) {
        const uint64_t a = loc, b = loc + J.ntiles;
        loc = b;
        const uint64_t lo = std::max(a, win_from), hi = std::min(b, win_end);
        if (lo >= hi) return false;
        J.p.tile_begin += lo - a;
        J.ntiles = hi - lo;
      } else if (win_from > 0) {
        return false;
      }
      J.tile0 = tiles;
      tiles += J.ntiles;
      inst_tiles[i] += J.ntiles;
      if (nj == jobv.size()) {
        jobv.resize(2 * nj);
        h_jobs = jobv.data();
      }
      h_jobs[nj++] = J;
      return true;
    };
    auto pair_job = [&](int j, uint32_t mask) {
      const CheckTargets tg = targets_for(st.t0, s, max_size, W);
      BatchJob J;
      std::memset(&J, 0, sizeof J);
      J.p.sb = s;
      J.p.j = j;
      J.p.m = s - 1 - j;
      J.p.nA = T.t[j][8];
      J.p.nB = T.t[J.p.m][8];
      J.p.ta = kPairTA;
      J.p.nTA = (J.p.nA + kPairTA - 1) / kPairTA;
      J.p.gen_a = j > st.Kt;
      J.p.gen_b = J.p.m > st.Kt;
      J.p.mask = mask;
      J.p.tri = j == J.p.m && !J.p.gen_a && !J.p.gen_b && (mask & kAllInst) == kAllInst;
      J.p.rb = kPairRB;
      J.lo_e = tg.lo_even;
      J.lo_o = tg.lo_odd;
      J.span = tg.span;
      J.inst = (uint32_t)i;
      J.kind = 1;
      if (!J.p.gen_b) {
        J.p.nTB = (J.p.nB + kPairTB - 1) / kPairTB;
        // few A values per tile: up to 8 B tiles per job tile, so a tile
        // holds about 64 A values' worth of pairs
        const uint64_t na = std::min<uint64_t>(J.p.nA, kPairTA);
        J.bsub = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(8, 64 / na));
        J.ntiles = J.p.nTA * ((J.p.nTB + J.bsub - 1) / J.bsub);
        push(J);
        return;
      }
      // a B side above Kt: the blocks formed from the tables, one job
      const BList bl = block_list(st, J.p.m);
      J.bb = bl.first;
      J.nbb = bl.n;
      J.p.nTB = bl.tiles;
      J.ntiles = J.p.nTA * bl.tiles;
      push(J);
    };
    if (first_round && s == 1) {  // leaves and every unary chain over them
      BatchJob J;
      std::memset(&J, 0, sizeof J);
      const CheckTargets tg = targets_for(st.t0, 1, max_size, W);
      J.ntiles = 1;
      J.lo_e = tg.lo_even;
      J.lo_o = tg.lo_odd;
      J.span = tg.span;
      J.inst = (uint32_t)i;
      J.kind = 2;
      push(J);
    }
    if (first_round && win_from == 0 && s >= 2 && s == st.Kt + 1 && s <= kt_cap) {
      // V[s] is tabulated when it fits this instance's share of the free
      // pool (the free elements over the instances still active)
      const uint64_t n = T.t[s][8];
      const uint64_t free_elems = pool_elems - pool_used;
      if (n <= free_elems &&
          (pool_all || n <= free_elems / std::max<size_t>(n_active_now, 1))) {
        tvs[i].off[s] = pool_used;
        tvs[i].Kt = s;
        st.Kt = s;
        tv_updates.push_back(TvUpdate{(uint32_t)i, (uint32_t)s, pool_used, s, 0});
        pool_used += n;
      }
    }
    if (first_round && s >= 2 && s <= st.Kt && st.Kt == s) {  // value table V[s]
      BatchJob J;
      std::memset(&J, 0, sizeof J);
      J.p.m = s;
      J.p.nB = T.t[s][8];
      J.ntiles = (J.p.nB + kBuildTile - 1) / kBuildTile;
      J.inst = (uint32_t)i;
      J.kind = 0;
      push(J);
    }
    if (s >= 3) {
      if (splits) {
        for (const auto &jm : *splits) pair_job(jm.first, jm.second);
      } else {
        for (int j = 1; j <= (s - 1) / 2; ++j) pair_job(j, kAllInst);
      }
    }
    return loc > win_end ? win_end : 0;
  };

  // Launch sequence of one size step k of a group: reset the counters, drain
  // the step's jobs, resolve its hits on the device (screen on two samples,
  // check on every sample, per-slot minimum rank), mark settled instances.
  auto step_launches = [&](const BatchStep *d_step) -> hipError_t {
    hipError_t e = launch_batch_reset(d_head, d_cnt, stq);
    if (e == hipSuccess)
      e = launch_batch_step(W, d_step, d_jobs, d_blocks, d_tvs, d_head, d_hits, &d_cnt->hits,
                            hit_cap, d_settled, grid, stq);
    if (e == hipSuccess)
      e = launch_hit_resolve(W, d_tvs, ctx.leafvals.as<uint64_t>(), ctx.outputs.as<uint64_t>(),
                             ctx.b_results.as<BatchInstData>(), d_hits, &d_cnt->hits, hit_cap,
                             1024, kConfirmBlocks, min_size, max_size, d_found, ctx.surv.as<SurvRec>(),
                             &d_cnt->surv, d_flags, d_touched, &d_cnt->touched, stq);
    if (e == hipSuccess)
      e = launch_batch_settle(d_step, d_found, d_settled, (uint32_t)count, min_size, d_cnt,
                              hit_cap, stq);
    return e;
  };

  // A group of nsz size steps (headers h_steps[0..nsz), jobs already in
  // h_jobs[0..njobs)) as one captured graph: the jobs are copied in, the
  // graph runs every step and copies the counters back; one sync.  The graph
  // reads the steps from device memory, so it is kept across groups and
  // calls of the same shape.
  auto run_group = [&](int nsz, size_t njobs) -> int {
    if (int rc = queue_room(njobs, blkv.size())) return rc;
    for (int k = 0; k < nsz; ++k) h_steps[k].index = (uint32_t)k;
    MBA_HIP(hipMemcpyAsync(d_steps, h_steps, sizeof(BatchStep) * nsz, hipMemcpyHostToDevice,
                           stq));
    MBA_HIP(hipMemsetAsync(&d_cnt->overflow, 0, 4, stq));
    MBA_HIP(hipMemsetAsync(&d_cnt->first_overflow, 0xff, 4, stq));
    MBA_HIP(hipMemcpyAsync(d_jobs, h_jobs, sizeof(BatchJob) * njobs, hipMemcpyHostToDevice,
                           stq));
    MBA_HIP(hipMemcpyAsync(d_blocks, blkv.data(), sizeof(BBlock) * blkv.size(),
                           hipMemcpyHostToDevice, stq));
    // every value and buffer the captured launches name
    const uint64_t key[Ctx::kGraphKey] = {(uint64_t)W,
                                          (uint64_t)grid,
                                          (uint64_t)hit_cap,
                                          (uint64_t)nsz | (uint64_t)count << 8,
                                          (uint64_t)min_size | (uint64_t)max_size << 8,
                                          (uint64_t)(uintptr_t)d_steps,
                                          (uint64_t)(uintptr_t)d_jobs,
                                          (uint64_t)(uintptr_t)d_blocks,
                                          (uint64_t)(uintptr_t)d_tvs,
                                          (uint64_t)(uintptr_t)d_hits,
                                          (uint64_t)(uintptr_t)d_cnt,
                                          (uint64_t)(uintptr_t)d_found,
                                          (uint64_t)(uintptr_t)ctx.surv.p,
                                          (uint64_t)(uintptr_t)ctx.leafvals.p,
                                          (uint64_t)(uintptr_t)ctx.outputs.p,
                                          (uint64_t)(uintptr_t)ctx.b_results.p,
                                          (uint64_t)(uintptr_t)h_cnt};
    if (!ctx.b_gexec || std::memcmp(key, ctx.b_graph_key, sizeof key) != 0) {
      MBA_HIP(hipStreamSynchronize(stq));
      if (ctx.b_gexec) MBA_HIP(hipGraphExecDestroy(ctx.b_gexec));
      if (ctx.b_graph) MBA_HIP(hipGraphDestroy(ctx.b_graph));
      ctx.b_gexec = nullptr;
      ctx.b_graph = nullptr;
      MBA_HIP(hipStreamBeginCapture(stq, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < nsz; ++k) MBA_HIP(step_launches(d_steps + k));
      MBA_HIP(hipMemcpyAsync(h_cnt, d_cnt, sizeof(BatchCounters), hipMemcpyDeviceToHost, stq));
      MBA_HIP(hipStreamEndCapture(stq, &ctx.b_graph));
      MBA_HIP(hipGraphInstantiate(&ctx.b_gexec, ctx.b_graph, nullptr, nullptr, 0));
      std::memcpy(ctx.b_graph_key, key, sizeof key);
    }
    MBA_HIP(hipEventRecord(ctx.ev0, stq));
    MBA_HIP(hipGraphLaunch(ctx.b_gexec, stq));
    MBA_HIP(hipEventRecord(ctx.ev1, stq));
    MBA_HIP(hipEventSynchronize(ctx.ev1));
    float ms = 0.f;
    MBA_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
    ktime += ms * 1e-3;
    ++klaunch;
    return 0;
  };

  // One size step run by direct launches on jobs [j0, j1) of h_jobs (a
  // step that overflowed the hit buffer inside a group): its hits are
  // resolved before returning; an overflowing part is run again as halves
  // of its jobs, a single job as halves of its tiles.  size_done is the
  // settle bound (the sizes complete once every part has run).
  uint32_t parts_lo = 0, parts_hi = 0;  // instances of the step run in parts
  std::function<int(size_t, size_t, int)> run_parts = [&](size_t j0, size_t j1,
                                                          int size_done) -> int {
    const uint64_t tbase = h_jobs[j0].tile0;
    const uint64_t tend = h_jobs[j1 - 1].tile0 + h_jobs[j1 - 1].ntiles;
    for (size_t k = j0; k < j1; ++k) h_jobs[k].tile0 -= tbase;
    BatchStep &hs = h_steps[kMaxSize + 1];
    hs.total_tiles = tend - tbase;
    hs.tiles_table = hs.total_tiles;
    for (size_t k = j0; k < j1; ++k)
      if (h_jobs[k].kind == 1 && h_jobs[k].p.gen_b) {
        hs.tiles_table = h_jobs[k].tile0;
        break;
      }
    hs.job_off = (uint32_t)j0;
    hs.njobs = (uint32_t)(j1 - j0);
    hs.size_done = size_done;
    hs.inst_lo = parts_lo;
    hs.inst_hi = parts_hi;
    BatchStep *d_hs = d_steps + kMaxSize + 1;
    MBA_HIP(hipMemcpyAsync(d_hs, &hs, sizeof hs, hipMemcpyHostToDevice, stq));
    MBA_HIP(hipMemcpyAsync(d_jobs + j0, h_jobs + j0, sizeof(BatchJob) * (j1 - j0),
                           hipMemcpyHostToDevice, stq));
    MBA_HIP(hipEventRecord(ctx.ev0, stq));
    MBA_HIP(step_launches(d_hs));
    MBA_HIP(hipMemcpyAsync(h_cnt, d_cnt, sizeof(BatchCounters), hipMemcpyDeviceToHost, stq));
    MBA_HIP(hipEventRecord(ctx.ev1, stq));
    MBA_HIP(hipEventSynchronize(ctx.ev1));
    for (size_t k = j0; k < j1; ++k) h_jobs[k].tile0 += tbase;
    float ms = 0.f;
    MBA_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
    ktime += ms * 1e-3;
    ++klaunch;
    if (h_cnt->hits <= hit_cap) return 0;
    ++replays;
    if (j1 - j0 > 1) {
      const size_t jm = j0 + (j1 - j0) / 2;
      if (int rc = run_parts(j0, jm, size_done)) return rc;
      return run_parts(jm, j1, size_done);
    }
    BatchJob &J = h_jobs[j0];
    const BatchJob keep = J;
    if (keep.ntiles < 2) {
      ctx.last_error = "batch: one tile's window hits overflow the hit buffer";
      return MBA_ERR_LIMIT;
    }
    const uint64_t half = keep.ntiles / 2;
    J.ntiles = half;
    int rc = run_parts(j0, j1, size_done);
    if (rc == 0) {
      J = keep;
      J.p.tile_begin += half;
      J.ntiles = keep.ntiles - half;
      rc = run_parts(j0, j1, size_done);
    }
    J = keep;
    return rc;
  };

  // Reads the verified slots and the settled sizes back: instances settled
  // at a size finish there (the smallest size with a verified tree, its
  // minimum rank; learned rounds: any verified tree of that size).
  auto collect = [&]() -> int {
    MBA_HIP(hipMemcpyAsync(found_h.data(), d_found, nslots * 8, hipMemcpyDeviceToHost, stq));
    MBA_HIP(hipMemcpyAsync(settled.data(), d_settled, count * 4, hipMemcpyDeviceToHost, stq));
    MBA_HIP(hipStreamSynchronize(stq));
    for (size_t i = 0; i < count; ++i) {
      InstState &st = S[i];
      std::memcpy(st.found, found_h.data() + i * (kMaxSize + 1), 8 * (kMaxSize + 1));
      if (st.active && settled[i] > 0) finish(i, settled[i]);
    }
    return 0;
  };

  // Charges each active instance its share of a group's device time (the
  // group's time split by the tiles of its jobs) and stops the instances
  // past the budget.
  auto charge = [&](double seconds, uint64_t total_tiles) {
    if (total_tiles == 0) return;
    for (size_t i = 0; i < count; ++i) {
      own[i] += seconds * (double)inst_tiles[i] / (double)total_tiles;
      inst_tiles[i] = 0;
    }
  };
  auto over_budget = [&](size_t i) {
    return o.time_budget_s > 0 && own[i] > o.time_budget_s;
  };
  // With a budget, an instance takes at most the tiles its remaining budget
  // covers at the tile rate of the last launch sequence (one sequence may
  // otherwise hold a whole large size of one instance, many budgets long);
  // the rest of the size follows in the next sequences while it has budget.
  double tile_rate = 2e7;  // tiles per second, updated after every sequence
  auto allowance = [&](size_t i) -> uint64_t {
    if (o.time_budget_s <= 0) return ~0ull;
    const double left = std::max(0.0, o.time_budget_s - own[i]);
    return std::max<uint64_t>(kMinAllowance, (uint64_t)(left * tile_rate));
  };

  // Tables for the B sides of size s (V[m], m <= s - 2).  An active
  // instance whose tables end below min(kt_cap, s - 2) gets the missing
  // sizes when the free pool holds its share of them.  Once the active
  // instances have halved since the pool was last packed, it is packed anew
  // from its start: every active instance's tables are placed again (V[1]
  // rewritten, V[2..] rebuilt), each up to the largest size its share of
  // the whole pool holds, and the finished instances' tables are dropped.
  // The builds run as steps of their own, one per table size and smaller
  // sizes first (a table is formed from the smaller ones), at the start of
  // a launch sequence: the views change before the sequence runs, so no
  // earlier step of it may read them.  Returns 1 when tables are due and
  // the group already holds steps (they open the next group).
  // Cohorts: from the size where the live instances' tables no longer fit
  // the pool whole, only a cohort of instances whose whole tables fit is
  // swept (the others wait at their next size, holding no tables), so every
  // B side of a large size comes from full tables rather than from deep
  // blocks over small shares; the next cohort starts once the current one
  // has finished.  Instances are independent, so the order changes no answer.
  std::vector<uint8_t> waiting(count, 0), swept(count, 0);  // swept: every size done
  auto live = [&](size_t i) { return S[i].active && !waiting[i] && !swept[i]; };
  bool force_repack = false;  // a new cohort: its tables placed whole
  size_t packed_active = count;
  int tables_for = -1;  // the size whose tables were last provided
  std::vector<std::pair<int, uint32_t>> builds;  // (table size, instance)
  int step_size[kMaxSize + 2];
  auto table_steps = [&](int s, size_t &nj, int &nsz, uint64_t &group_tiles) -> int {
    if (tables_for == s) return 0;
    const int want = std::min(kt_cap, s - 2);
    if (want < 2) {
      tables_for = s;
      return 0;
    }
    const size_t act = std::max<size_t>(n_active_now, 1);
    const bool repack = force_repack || (s >= kRepackMin && n_active_now * 2 <= packed_active);
    const uint64_t free_share =
        pool_all ? pool_elems - pool_used : (pool_elems - pool_used) / act;
    bool due = false;  // some table to place
    for (size_t i = 0; i < count && !due; ++i)
      due = live(i) && S[i].Kt < want && (repack || S[i].T->t[S[i].Kt + 1][8] <= free_share);
    if (!due) {
      tables_for = s;
      return 0;
    }
    if (nsz > 0) return 1;
    tables_for = s;
    builds.clear();
    auto place = [&](size_t i, int m) {
      if (pool_used + S[i].T->t[m][8] > pool_elems) {  // never: the callers leave room
        pool_overrun = true;
        return;
      }
      tvs[i].off[m] = pool_used;
      tv_updates.push_back(TvUpdate{(uint32_t)i, (uint32_t)m, pool_used, 0, 0});
      pool_used += S[i].T->t[m][8];
      if (m >= 2) builds.emplace_back(m, (uint32_t)i);
    };
    if (repack) {
      packed_active = n_active_now;
      pool_used = 0;
      // a cohort's tables whole (its first instance may alone need more
      // than the pool: it gets what fits); else an equal share each.  An
      // instance never takes the room the later live instances' V[1] need,
      // so every placement stays inside the pool.
      const uint64_t share = force_repack ? ~0ull : pool_elems / act;
      force_repack = false;
      uint64_t leaves_after = 0;  // V[1] elements of the live instances not placed yet
      for (size_t i = 0; i < count; ++i)
        if (live(i)) leaves_after += (uint64_t)S[i].L;
      for (size_t i = 0; i < count; ++i) {
        InstState &st = S[i];
        if (!live(i)) {  // tables dropped: the view keeps Kt 0
          if (st.Kt > 0) tv_updates.push_back(TvUpdate{(uint32_t)i, 1u, 0, 0, 0});
          st.Kt = 0;
          continue;
        }
        leaves_after -= (uint64_t)st.L;
        const uint64_t room = pool_elems - pool_used - leaves_after;  // >= st.L
        const uint64_t lim = std::min(share, room);
        const int top = std::min(kt_cap, std::max(want, st.Kt));
        uint64_t used = (uint64_t)st.L;
        int k = 1;
        while (k < top && used + st.T->t[k + 1][8] <= lim) used += st.T->t[++k][8];
        for (int m = 1; m <= k; ++m) place(i, m);
        st.Kt = k;
        tvs[i].Kt = k;
      }
      leaves_pending = true;
    } else {
      const uint64_t share = pool_all ? ~0ull : free_share;
      for (size_t i = 0; i < count; ++i) {
        InstState &st = S[i];
        if (!live(i) || st.Kt >= want) continue;
        uint64_t got = 0;
        while (st.Kt < want) {
          const uint64_t n = st.T->t[st.Kt + 1][8];
          if (n > pool_elems - pool_used || got + n > share) break;
          got += n;
          place(i, ++st.Kt);
          tvs[i].Kt = st.Kt;
        }
      }
    }
    std::sort(builds.begin(), builds.end());
    for (size_t b = 0; b < builds.size();) {
      const int m = builds[b].first;
      const size_t j0 = nj;
      uint64_t tiles = 0;
      for (; b < builds.size() && builds[b].first == m; ++b) {
        const uint32_t i = builds[b].second;
        BatchJob J;
        std::memset(&J, 0, sizeof J);
        J.p.m = m;
        J.p.nB = S[i].T->t[m][8];
        J.ntiles = (J.p.nB + kBuildTile - 1) / kBuildTile;
        J.inst = i;
        J.kind = 0;
        J.tile0 = tiles;
        tiles += J.ntiles;
        inst_tiles[i] += J.ntiles;
        if (nj == jobv.size()) {
          jobv.resize(2 * nj);
          h_jobs = jobv.data();
        }
        h_jobs[nj++] = J;
      }
      step_size[nsz] = 0;
      BatchStep &hs = h_steps[nsz++];
      std::memset(&hs, 0, sizeof hs);
      hs.total_tiles = tiles;
      hs.tiles_table = tiles;
      hs.job_off = (uint32_t)j0;
      hs.njobs = (uint32_t)(nj - j0);
      group_tiles += tiles;
    }
    return 0;
  };

  // sizes whose steps would use the learned order for some active instance
  auto learned_size = [&](int s) {
    if (!score_size || s < 3) return false;
    for (size_t i = 0; i < count; ++i)
      if (live(i) && S[i].T->t[s][8] >= learned_min) return true;
    return false;
  };
  // A group ends once it holds this many tiles: with a budget or a cancel
  // flag the host checks between groups (smaller groups); without, large
  // groups keep small searches in one launch sequence while the host still
  // drops settled instances from the jobs of large sizes.
  const bool checkpoints = o.time_budget_s > 0 || o.cancel != nullptr;
  const uint64_t kGroupTiles = checkpoints ? (1ull << 18) : (1ull << 20);

  int s = 0;
  size_t cursor = 0;  // the first instance of size s still to sweep
  // Per instance: the next size it sweeps, and the pair tiles of that size
  // swept so far (nonzero only for an instance whose budget window ended
  // inside the size).  Every active instance is at size s or s + 1.
  std::vector<int32_t> next_sz(count, 0);
  std::vector<uint64_t> part_t(count, 0);
  // elements of V[1..want] of instance i, and whether the live instances'
  // whole tables for size s exceed the pool
  // An active instance inside a size's tiles (a budget window) or waiting
  // for its cohort settles no size from its next one on: settled = -(next
  // size) bounds the device's settle scan below it (a window hit of a
  // smaller tree may already hold a larger size's slot).  Written to the
  // device before a launch sequence whenever some bound is or was set.
  bool bounds_set = false;
  auto bound_settle = [&]() -> int {
    bool any = o.time_budget_s > 0;
    for (size_t k = 0; k < count && !any; ++k) any = S[k].active && waiting[k];
    if (!any && !bounds_set) return 0;
    bounds_set = any;
    for (size_t k = 0; k < count; ++k)
      if (settled[k] <= 0)
        settled[k] = S[k].active && (waiting[k] || part_t[k] > 0) ? -next_sz[k] : 0;
    MBA_HIP(hipMemcpyAsync(d_settled, settled.data(), count * 4, hipMemcpyHostToDevice, stq));
    return 0;
  };
  auto table_need = [&](size_t i, int want) {
    uint64_t e = (uint64_t)S[i].L;
    for (int m = 2; m <= want; ++m) e += S[i].T->t[m][8];
    return e;
  };
  auto cohort_over = [&](int z) {
    const int want = std::min(kt_cap, z - 2);
    if (pool_all || want < 2) return false;
    uint64_t need = 0;
    for (size_t i = 0; i < count; ++i)
      if (live(i)) need += table_need(i, want);
    return need > pool_elems;
  };
  auto pending_at = [&](int z) {
    for (size_t i = 0; i < count; ++i)
      if (live(i) && next_sz[i] == z) return true;
    return false;
  };
  for (;;) {
    if (o.cancel && *o.cancel) {
      budget_hit = true;
      break;
    }
    if (s > max_size)  // the live instances have swept every size
      for (size_t i = 0; i < count; ++i)
        if (live(i)) swept[i] = 1;
    size_t n_active = 0;
    for (size_t i = 0; i < count; ++i) {
      if (S[i].active && over_budget(i)) {
        S[i].active = false;
        results[i].status = MBA_BUDGET;
        results[i].seconds = seconds_since(t_start);
      }
      n_active += live(i);
    }
    if (n_active == 0) {
      // the next cohort: the waiting instances at the smallest next size
      int z = kMaxSize + 2;
      for (size_t i = 0; i < count; ++i)
        if (S[i].active && waiting[i]) z = std::min(z, (int)next_sz[i]);
      if (z > max_size) break;
      for (size_t i = 0; i < count; ++i)
        if (S[i].active && waiting[i] && next_sz[i] == z) waiting[i] = 0, ++n_active;
      s = z;
      cursor = 0;
      tables_for = -1;
      force_repack = true;
    }
    if (cursor == 0 && cohort_over(s)) {
      // keep the instances (in order) whose whole tables fit the pool
      const int want = std::min(kt_cap, s - 2);
      uint64_t used = 0;
      n_active = 0;
      for (size_t i = 0; i < count; ++i) {
        if (!live(i)) continue;
        const uint64_t need = table_need(i, want);
        if (n_active > 0 && used + need > pool_elems) {
          waiting[i] = 1;
          continue;
        }
        used += need;
        ++n_active;
      }
      tables_for = -1;
      force_repack = true;
    }
    n_active_now = n_active;
    if (cursor > 0 || !learned_size(s)) {
      // ---- a group of plain size steps, one launch sequence -------------
      // Each step is one size over a range of instances: all of them, or as
      // many as keep the group near kGroupTiles tiles, the next group
      // resuming within the size.  With a budget, an instance takes the
      // window of the size's tiles its budget allows; one whose window ends
      // inside the size ends the group at the end of the size's pass (its
      // next window waits for its charge) and settles nothing at that size.
      const auto t_b = Clock::now();
      size_t nj = 0;
      uint64_t group_tiles = 0;
      int nsz = 0;
      const int s0 = s;
      while (s <= max_size && nsz < kMaxSize + 1) {
        if (nsz > 0 && cursor == 0 && (learned_size(s) || cohort_over(s))) break;
        if (table_steps(s, nj, nsz, group_tiles)) break;  // no room: the next group
        if (nsz == kMaxSize + 1) break;
        const size_t j0 = nj;
        uint64_t tiles = 0;
        size_t i = cursor;
        for (; i < count; ++i) {
          if (!live(i) || next_sz[i] != s) continue;
          const uint64_t next =
              size_jobs(i, s, true, nullptr, nj, tiles, part_t[i], allowance(i));
          part_t[i] = next;  // > 0: instance i resumes size s in a later pass
          if (next == 0) next_sz[i] = s + 1;
          if (group_tiles + tiles >= kGroupTiles) {
            ++i;
            break;
          }
        }
        step_size[nsz] = s;
        BatchStep &hs = h_steps[nsz++];
        hs.total_tiles = tiles;
        hs.tiles_table = order_step(j0, nj);
        hs.job_off = (uint32_t)j0;
        hs.njobs = (uint32_t)(nj - j0);
        // sizes 1..2 complete with the leaf scan (step 1); steps 0 and 2
        // settle nothing
        hs.size_done = s == 1 ? std::min(2, max_size) : (s == 0 || s == 2) ? 0 : s;
        hs.inst_lo = (uint32_t)cursor;
        hs.inst_hi = (uint32_t)i;
        group_tiles += tiles;
        bool skip_rest = true;  // no instance after i sweeps size s
        for (size_t k = i; k < count && skip_rest; ++k)
          skip_rest = !(live(k) && next_sz[k] == s);
        if (i < count && !skip_rest) {
          cursor = i;  // the group ends inside size s
          break;
        }
        cursor = 0;
        if (pending_at(s)) break;  // size s takes another pass
        ++s;
        if (group_tiles >= kGroupTiles) break;
      }
      if (int rc = flush_tables()) return rc;
      if (int rc = bound_settle()) return rc;
      const auto t_g = Clock::now();
      host_build += seconds_since(t_b);
      blocks_sent = blkv.size();
      if (nj > 0) {
        if (int rc = run_group(nsz, nj)) return rc;
        if (h_cnt->overflow > 0) {
          // step k0 was the first whose hits overflowed the buffer: the
          // settled sizes are rebuilt from before the group through step
          // k0 - 1 (their hits were all resolved), and steps k0.. run again
          // one at a time, in parts (resolving any subset of the hits only
          // lowers slots to verified ranks; a size settled on part of its
          // hits may not hold its minimum rank yet)
          const int k0 = (int)std::min<uint32_t>(h_cnt->first_overflow, (uint32_t)nsz);
          MBA_HIP(hipMemcpyAsync(d_settled, settled.data(), count * 4, hipMemcpyHostToDevice,
                                 stq));
          for (int k = 0; k < k0; ++k)
            MBA_HIP(launch_batch_settle(d_steps + k, d_found, d_settled, (uint32_t)count,
                                        min_size, d_cnt, hit_cap, stq));
          for (int k = k0; k < nsz; ++k) {
            const BatchStep hs = h_steps[k];
            if (hs.njobs == 0) continue;
            parts_lo = hs.inst_lo;
            parts_hi = hs.inst_hi;
            if (int rc = run_parts(hs.job_off, hs.job_off + hs.njobs,
                                   std::max(0, hs.size_done - 1)))
              return rc;
            // settle the complete size
            h_steps[kMaxSize + 1] = hs;
            h_steps[kMaxSize + 1].njobs = 0;
            h_steps[kMaxSize + 1].total_tiles = 0;
            MBA_HIP(hipMemcpyAsync(d_steps + kMaxSize + 1, &h_steps[kMaxSize + 1],
                                   sizeof(BatchStep), hipMemcpyHostToDevice, stq));
            MBA_HIP(launch_batch_settle(d_steps + kMaxSize + 1, d_found, d_settled,
                                        (uint32_t)count, min_size, d_cnt, hit_cap, stq));
          }
        }
      }
      if (o.verbose)
        std::fprintf(stderr, "batch: sizes %d..%d in one launch sequence (%d steps): %zu jobs, "
                     "%llu tiles, %.3f ms, last step %u hits, replays %d\n", s0,
                     cursor ? s : s - 1, nsz, nj, (unsigned long long)group_tiles,
                     1e3 * seconds_since(t_g), h_cnt->hits, replays);
      const double t_seq = seconds_since(t_g);
      if (group_tiles >= 1024 && t_seq > 0) tile_rate = (double)group_tiles / t_seq;
      charge(t_seq, group_tiles);
      const auto t_c = Clock::now();
      if (int rc = collect()) return rc;
      host_collect += seconds_since(t_c);
      for (int k = 0; k < nsz; ++k)  // sizes completed by the group's steps
        for (size_t i = h_steps[k].inst_lo; i < h_steps[k].inst_hi; ++i)
          if (S[i].active && h_steps[k].size_done > 0 && next_sz[i] > step_size[k])
            S[i].reached = std::max(S[i].reached, h_steps[k].size_done);
      continue;
    }
    // ---- a learned size: rounds of the best-scored split, the next split,
    // then the rest (each split swept once with every operator instance
    // fused); instances with fewer trees at this size take every split in
    // the first round; any verified solution of this (minimal) size settles
    {  // the size's tables first (a launch sequence of build steps, if any)
      size_t nj = 0;
      int nsz = 0;
      uint64_t tiles = 0;
      table_steps(s, nj, nsz, tiles);  // an empty group has room for them
      if (int rc = flush_tables()) return rc;
      if (nj > 0) {
        const auto t_g = Clock::now();
        if (int rc = bound_settle()) return rc;
        if (int rc = run_group(nsz, nj)) return rc;
        charge(seconds_since(t_g), tiles);
      }
    }
    if (int rc = (*score_size)(s, scores.data())) return rc;
    size_t max_splits = 0;
    for (size_t i = 0; i < count; ++i) {
      InstState &st = S[i];
      order[i].clear();
      if (!live(i) || st.T->t[s][8] < learned_min) continue;
      results[i].ordered_sizes++;
      order[i] = learned_split_order(*st.T, s, scores.data() + i * 256);
      max_splits = std::max(max_splits, order[i].size());
    }
    // rounds of splits (each with every operator instance fused): the first
    // proposed split, the second, the rest
    static const size_t kRoundEnd[3] = {1, 2, ~(size_t)0};
    std::vector<std::pair<int, uint32_t>> sub;
    for (int round = 0; round < 3; ++round) {
      const size_t rb = round == 0 ? 0 : kRoundEnd[round - 1];
      const size_t re = kRoundEnd[round];
      size_t nj = 0;
      uint64_t tiles = 0;
      for (size_t i = 0; i < count; ++i) {
        if (!live(i)) continue;
        if (order[i].empty()) {
          if (round == 0) size_jobs(i, s, true, nullptr, nj, tiles);
          continue;
        }
        sub.clear();
        for (size_t k = std::min(rb, order[i].size()); k < std::min(re, order[i].size()); ++k)
          sub.emplace_back(order[i][k], kAllInst);
        size_jobs(i, s, round == 0, &sub, nj, tiles);
      }
      if (nj > 0) {
        BatchStep &hs = h_steps[0];
        hs.total_tiles = tiles;
        hs.tiles_table = order_step(0, nj);
        hs.job_off = 0;
        hs.njobs = (uint32_t)nj;
        hs.size_done = s;  // any verified tree of this size settles
        hs.inst_lo = 0;
        hs.inst_hi = (uint32_t)count;
        if (int rc = flush_tables()) return rc;
        const auto t_g = Clock::now();
        if (int rc = bound_settle()) return rc;
        if (int rc = run_group(1, nj)) return rc;
        if (h_cnt->overflow > 0) {
          MBA_HIP(hipMemsetAsync(&d_cnt->overflow, 0, 4, stq));
          parts_lo = 0;
          parts_hi = (uint32_t)count;
          if (int rc = run_parts(0, nj, s)) return rc;
        }
        if (o.verbose)
          std::fprintf(stderr, "batch: size %d round %d: %zu jobs, %llu tiles, %.3f ms, "
                       "%u hits, replays %d\n", s, round, nj, (unsigned long long)tiles,
                       1e3 * seconds_since(t_g), h_cnt->hits, replays);
        charge(seconds_since(t_g), tiles);
        if (int rc = collect()) return rc;
      }
      if (re >= max_splits) break;
    }
    for (size_t i = 0; i < count; ++i)
      if (live(i)) {
        S[i].reached = s;
        next_sz[i] = s + 1;
      }
    ++s;
  }

  // every answer checked once more on all samples by the verify kernel
  std::vector<VerifyCand> vc;
  std::vector<size_t> vi;
  for (size_t i = 0; i < count; ++i)
    if (results[i].status == MBA_FOUND) {
      vc.push_back(VerifyCand{(uint32_t)i, (uint32_t)results[i].size, results[i].rank});
      vi.push_back(i);
    }
  if (!vc.empty()) {
    MBA_HIP(ctx.vranks.ensure(sizeof(VerifyCand) * vc.size()));
    MBA_HIP(ctx.vok.ensure(vc.size()));
    MBA_HIP(hipMemcpyAsync(ctx.vranks.p, vc.data(), sizeof(VerifyCand) * vc.size(),
                           hipMemcpyHostToDevice, stq));
    MBA_HIP(launch_verify_batch(W, d_tvs, ctx.leafvals.as<uint64_t>(), ctx.outputs.as<uint64_t>(),
                                ctx.b_results.as<BatchInstData>(), ctx.vranks.as<VerifyCand>(),
                                (uint32_t)vc.size(), ctx.vok.as<uint8_t>(), stq));
    std::vector<uint8_t> ok(vc.size());
    MBA_HIP(hipMemcpyAsync(ok.data(), ctx.vok.p, vc.size(), hipMemcpyDeviceToHost, stq));
    MBA_HIP(hipStreamSynchronize(stq));
    for (size_t k = 0; k < vc.size(); ++k)
      if (!ok[k]) {
        ctx.last_error = "batch: an answer failed verification on every sample";
        return MBA_ERR_LIMIT;
      }
  }
  const double total = seconds_since(t_start);
  if (o.verbose)
    std::fprintf(stderr, "batch: %u launch sequences, device %.3f ms, total %.3f ms (host: "
                 "job building %.3f ms, %zu blocks; read backs %.3f ms)\n", klaunch, 1e3 * ktime,
                 1e3 * total, 1e3 * host_build, blocks_sent, 1e3 * host_collect);
  for (size_t i = 0; i < count; ++i) {
    InstState &st = S[i];
    mba_result &r = results[i];
    uint64_t cand = 0;
    for (int z = min_size; z <= st.reached; ++z) cand += st.T->t[z][8];
    r.candidates = cand;
    r.size_reached = st.reached;
    r.kernel_seconds = ktime;
    r.kernel_launches = klaunch;
    r.hit_replays = replays;
    if (st.active) {
      r.status = budget_hit ? MBA_BUDGET : MBA_NOT_FOUND;
      r.seconds = total;
    }
  }
  return 0;
}

}  // namespace mba
