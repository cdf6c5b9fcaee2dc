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

constexpr uint32_t kBatchHitCap = 1u << 24;  // 512 MiB of hit records
// default value-table budget of a batch, split evenly over its instances
// (at most 2/5 of the device memory): 4096 instances of up to 4 leaves keep
// V[8] in tables at a size budget of 11, so the last size's B sides are
// formed from tabulated children
constexpr uint64_t kBatchTableBytes = 96ull << 30;

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
                mba_result *results, const float *scores) {
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
  const uint64_t per_budget = std::max<uint64_t>(table_budget / count, 1 << 16);
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
  // pinned staging (kept in the context, grown when needed): step header at
  // 0, hit and slot counts at 256, then the jobs of a step, the instances'
  // table views, leaf values, outputs and sample layouts
  size_t job_cap = 0;
  for (size_t i = 0; i < count; ++i) job_cap += 2 + (size_t)std::max(0, (max_size - 1) / 2);
  const size_t off_jobs = 512;
  const size_t off_tvs = align256(off_jobs + sizeof(BatchJob) * job_cap);
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
  BatchStep *h_step = (BatchStep *)pin;
  uint32_t *h_hc = (uint32_t *)(pin + 256);
  BatchJob *h_jobs = (BatchJob *)(pin + off_jobs);
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
      // V[max_size - 2] is only ever the B side of leaves: formed per tile
      st.Kt = std::max(1, std::min(choose_kt(*st.T, std::max(max_size, 1), W, per_budget),
                                   max_size - 3) -
                              std::max(0, o.table_delta));
      TableView &tv = tvs[i];
      std::memset(&tv, 0, sizeof tv);
      uint64_t e = 0;
      for (int m = 1; m <= st.Kt; ++m) {
        tv.off[m] = e;
        e += st.T->t[m][8];
      }
      st.elems = e;
      std::memcpy(tv.T, st.T->t, sizeof tv.T);
      tv.Kt = st.Kt;
      tv.L = L;
      idat[i].leaf_off = leaf_off[i];
      idat[i].out_off = out_off[i];
      idat[i].n_samples = (uint32_t)n;
      idat[i].fs = (uint32_t)st.fs;
    }
  });
  // pass 3: the value-table pool, one slice per instance
  uint64_t pool_elems = 0;
  for (size_t i = 0; i < count; ++i) pool_elems += S[i].elems;
  MBA_HIP(ctx.b_data.ensure(std::max<size_t>(pool_elems * wb, 256)));
  {
    uint64_t pe = 0;
    for (size_t i = 0; i < count; ++i) {
      tvs[i].base = ctx.b_data.as<char>() + pe * wb;
      pe += S[i].elems;
    }
  }
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

  // ---- per-step graph: upload jobs, reset counters, drain, read hit count --
  MBA_HIP(ctx.b_queue.ensure(sizeof(BatchJob) * job_cap + 256));
  // hit buffer (opts.hit_cap; smaller buffers exercise the split replays)
  const uint32_t hit_cap = o.hit_cap > 0 ? (uint32_t)o.hit_cap : kBatchHitCap;
  MBA_HIP(ctx.hits.ensure(sizeof(HitRec) * hit_cap));
  MBA_HIP(ctx.hit_count.ensure(256));
  BatchJob *d_jobs = (BatchJob *)(ctx.b_queue.as<char>() + 256);
  BatchStep *d_step = ctx.b_queue.as<BatchStep>();
  unsigned long long *d_head = (unsigned long long *)(ctx.hit_count.as<char>() + 64);
  uint32_t *d_hc = ctx.hit_count.as<uint32_t>();
  HitRec *d_hits = ctx.hits.as<HitRec>();
  const TableView *d_tvs = ctx.b_inst.as<TableView>();
  // device hit resolution: per-(instance, size) minimum rank, slot flags and
  // the list of slots a resolve launch lowered
  const size_t nslots = count * (kMaxSize + 1);
  MBA_HIP(ctx.b_misc.ensure(nslots * 16 + 256));
  unsigned long long *d_found = (unsigned long long *)(ctx.b_misc.as<char>() + 256);
  uint32_t *d_flags = (uint32_t *)(d_found + nslots);
  uint32_t *d_touched = d_flags + nslots;
  uint32_t *d_ntouched = ctx.b_misc.as<uint32_t>();
  uint32_t *h_nt = h_hc + 4;
  MBA_HIP(hipMemsetAsync(d_found, 0xff, nslots * 8, stq));
  MBA_HIP(hipMemsetAsync(d_flags, 0, nslots * 4, stq));
  const int mult = o.wg_per_cu > 0 ? o.wg_per_cu : 8;  // workgroups per CU
  // learned order from sizes of this many trees (opts.block_log2 lowers it)
  const uint64_t learned_min =
      o.block_log2 > 0 ? (1ull << std::min(o.block_log2, 63)) : kLearnedBlockMin;
  const int grid = ctx.cu_count * mult;

  // the step graph names these buffers and the launch shape: it is captured
  // again only when one of them changed since the previous call
  const uint64_t key[11] = {(uint64_t)W, (uint64_t)grid, (uint64_t)hit_cap,
                            (uint64_t)(uintptr_t)d_step,
                            (uint64_t)(uintptr_t)d_jobs, (uint64_t)(uintptr_t)d_tvs,
                            (uint64_t)(uintptr_t)d_head, (uint64_t)(uintptr_t)d_hits,
                            (uint64_t)(uintptr_t)d_hc, (uint64_t)(uintptr_t)h_step,
                            (uint64_t)(uintptr_t)h_hc};
  MBA_HIP(hipStreamSynchronize(stq));
  if (!ctx.b_gexec || std::memcmp(key, ctx.b_graph_key, sizeof key) != 0) {
    if (ctx.b_gexec) MBA_HIP(hipGraphExecDestroy(ctx.b_gexec));
    if (ctx.b_graph) MBA_HIP(hipGraphDestroy(ctx.b_graph));
    ctx.b_gexec = nullptr;
    ctx.b_graph = nullptr;
    MBA_HIP(hipStreamBeginCapture(stq, hipStreamCaptureModeThreadLocal));
    MBA_HIP(hipMemcpyAsync(d_step, h_step, sizeof(BatchStep), hipMemcpyHostToDevice, stq));
    MBA_HIP(launch_batch_reset(d_head, d_hc, stq));
    MBA_HIP(launch_batch_step(W, d_step, d_jobs, d_tvs, d_head, d_hits, d_hc, hit_cap, grid,
                              stq));
    MBA_HIP(hipMemcpyAsync(h_hc, d_hc, 4, hipMemcpyDeviceToHost, stq));
    MBA_HIP(hipStreamEndCapture(stq, &ctx.b_graph));
    MBA_HIP(hipGraphInstantiate(&ctx.b_gexec, ctx.b_graph, nullptr, nullptr, 0));
    std::memcpy(ctx.b_graph_key, key, sizeof key);
  }
  hipGraphExec_t gexec = ctx.b_gexec;

  std::vector<HitRec> hrec;
  std::vector<VerifyCand> vc;
  std::vector<uint8_t> okv;
  double ktime = 0.0;
  uint32_t klaunch = 0;
  bool budget_hit = false;
  auto finish = [&](size_t i, int z) {
    InstState &st = S[i];
    fill_result(*st.T, z, st.found[z], (int)insts[i].n_vars, insts[i].consts, W, results[i]);
    results[i].seconds = seconds_since(t_start);
    st.reached = z;
    st.active = false;
  };
  std::vector<std::vector<int>> order(count);  // learned: splits in score order
  if (o.verbose)
    std::fprintf(stderr, "batch: %zu instances set up in %.3f ms\n", count,
                 1e3 * seconds_since(t_start));
  double t_build = 0, t_gpu = 0, t_resolve = 0;
  uint64_t n_hits = 0, n_kept = 0, n_cands = 0;
  int replays = 0;  // steps run again in parts after a hit buffer overflow
  std::vector<uint32_t> hits_of(o.verbose ? count : 0, 0);
  for (int s = 0; s <= max_size; ++s) {
    if ((o.cancel && *o.cancel) ||
        (o.time_budget_s > 0 && seconds_since(t_start) > o.time_budget_s)) {
      budget_hit = true;
      break;
    }
    const bool learned = scores && s >= 3;
    size_t max_blocks = 0;
    if (learned) {
      for (size_t i = 0; i < count; ++i) {
        InstState &st = S[i];
        if (!st.active) continue;
        // a unary-rooted solution of this size is verified already
        if (s >= min_size && st.found[s] != ~0ull) {
          finish(i, s);
          continue;
        }
        // small sizes are swept fused in one round (a round per block
        // group costs a graph replay and a sync)
        if (st.T->t[s][8] < learned_min) {
          order[i].clear();
          continue;
        }
        results[i].ordered_sizes++;
        std::vector<BlockRef> blocks = binary_blocks(*st.T, s);
        const float *sc = scores + i * 256;
        std::stable_sort(blocks.begin(), blocks.end(), [&](const BlockRef &x, const BlockRef &y) {
          return sc[x.op * 32 + x.split] > sc[y.op * 32 + y.split];
        });
        order[i].clear();
        for (const BlockRef &b : blocks)
          if (std::find(order[i].begin(), order[i].end(), b.pair_j) == order[i].end())
            order[i].push_back(b.pair_j);
        max_blocks = std::max(max_blocks, order[i].size());
      }
    }
    // learned order: rounds of the split of the best block, the next split,
    // then the rest, each split swept once with every operator instance fused
    // (a round costs a replay and a sync; sizes without a solution need
    // every split anyway)
    static const size_t kRoundEnd[3] = {1, 2, ~(size_t)0};
    for (int round = 0;; ++round) {
      const size_t rb = round == 0 ? 0 : kRoundEnd[round - 1];
      const size_t re = kRoundEnd[std::min(round, 2)];
      size_t nj = 0;
      uint64_t tiles = 0;
      const auto t_b0 = Clock::now();
      for (size_t i = 0; i < count; ++i) {
        InstState &st = S[i];
        if (!st.active) continue;
        auto push = [&](const BatchJob &J) {
          h_jobs[nj] = J;
          h_jobs[nj].tile0 = tiles;
          tiles += J.ntiles;
          ++nj;
        };
        const CountTable &T = *st.T;
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
          J.p.rb = kPairRB;
          J.ntiles = J.p.nTA * ((J.p.nB + kPairTB - 1) / kPairTB);
          J.lo_e = tg.lo_even;
          J.lo_o = tg.lo_odd;
          J.span = tg.span;
          J.inst = (uint32_t)i;
          J.kind = 1;
          push(J);
        };
        if (round == 0 && s == 1) {  // leaves and every unary chain over them
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
        if (round == 0 && s >= 2 && s <= st.Kt) {  // value table V[s]
          BatchJob J;
          std::memset(&J, 0, sizeof J);
          J.p.m = s;
          J.p.nB = T.t[s][8];
          J.ntiles = (J.p.nB + kBuildTile - 1) / kBuildTile;
          J.inst = (uint32_t)i;
          J.kind = 0;
          push(J);
        }
        if (s >= 3 && round == 0 && (!learned || T.t[s][8] < learned_min)) {
          for (int j = 1; j <= (s - 1) / 2; ++j) pair_job(j, kAllInst);
        }
        if (learned) {
          for (size_t k = rb; k < re && k < order[i].size(); ++k) pair_job(order[i][k], kAllInst);
        }
      }
      t_build += seconds_since(t_b0);
      // Runs jobs [j0, j1) of this step (one graph replay) and resolves their
      // hits; a replay whose hits overflow the buffer is run again as two
      // halves of its jobs (jobs are independent within a step, table builds
      // idempotent), so only one job that overflows it alone is an error.
      std::function<int(size_t, size_t)> run_jobs = [&](size_t j0, size_t j1) -> int {
        const uint64_t tbase = h_jobs[j0].tile0;
        const uint64_t tend = h_jobs[j1 - 1].tile0 + h_jobs[j1 - 1].ntiles;
        for (size_t k = j0; k < j1; ++k) h_jobs[k].tile0 -= tbase;
        const auto t_g0 = Clock::now();
        h_step->total_tiles = tend - tbase;
        h_step->njobs = (uint32_t)(j1 - j0);
        MBA_HIP(hipEventRecord(ctx.ev0, stq));
        // these jobs only (the graph then uploads the step header, resets the
        // queue, drains it and reads the hit count back)
        MBA_HIP(hipMemcpyAsync(d_jobs, h_jobs + j0, sizeof(BatchJob) * (j1 - j0),
                               hipMemcpyHostToDevice, stq));
        MBA_HIP(hipGraphLaunch(gexec, stq));
        MBA_HIP(hipEventRecord(ctx.ev1, stq));
        MBA_HIP(hipEventSynchronize(ctx.ev1));
        for (size_t k = j0; k < j1; ++k) h_jobs[k].tile0 += tbase;
        float ms = 0.f;
        MBA_HIP(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
        ktime += ms * 1e-3;
        ++klaunch;
        t_gpu += seconds_since(t_g0);
        if (o.verbose)
          std::fprintf(stderr, "batch: size %d round %d: %zu jobs, %llu tiles, %.3f ms\n", s,
                       round, j1 - j0, (unsigned long long)(tend - tbase), ms);
        const auto t_r0 = Clock::now();
        const uint32_t nh = *h_hc;
        if (nh > hit_cap) {
          ++replays;
          if (j1 - j0 == 1) {
            // one job: its tiles in two parts (a tile holds at most
            // 256 * 2048 * 7 hits, well under the buffer)
            BatchJob &J = h_jobs[j0];
            const BatchJob keep = J;
            if (keep.ntiles < 2) {
              ctx.last_error = "batch: one tile's window hits overflow the hit buffer";
              return MBA_ERR_LIMIT;
            }
            const uint64_t half = keep.ntiles / 2;
            J.ntiles = half;
            int rc = run_jobs(j0, j1);
            if (rc == 0) {
              J = keep;
              J.p.tile_begin += half;
              J.ntiles = keep.ntiles - half;
              rc = run_jobs(j0, j1);
            }
            J = keep;
            return rc;
          }
          const size_t jm = j0 + (j1 - j0) / 2;
          if (int rc = run_jobs(j0, jm)) return rc;
          return run_jobs(jm, j1);
        }
        n_hits += nh;
        if (nh > 0 && !hits_of.empty()) {  // verbose: hits per instance
          hrec.resize(nh);
          MBA_HIP(hipMemcpy(hrec.data(), d_hits, sizeof(HitRec) * nh, hipMemcpyDeviceToHost));
          for (const HitRec &h0 : hrec) hits_of[h0.inst >> 8]++;
        }
        if (nh > 0) {
          // resolve every hit on the device (screen on two samples, check on
          // every sample, per-slot minimum), then verify the lowered slots'
          // trees
          MBA_HIP(ctx.surv.ensure(sizeof(SurvRec) * 2 * (size_t)nh));
          MBA_HIP(hipMemsetAsync(d_ntouched, 0, 8, stq));  // touched and candidate counts
          MBA_HIP(launch_hit_resolve(W, d_tvs, ctx.leafvals.as<uint64_t>(),
                                     ctx.outputs.as<uint64_t>(),
                                     ctx.b_results.as<BatchInstData>(), d_hits, d_hc, nh,
                                     (nh + 63) / 64, std::min<uint32_t>(2 * nh, 8192),
                                     min_size, max_size, d_found, ctx.surv.as<SurvRec>(),
                                     d_ntouched + 1, d_flags, d_touched, d_ntouched, stq));
          MBA_HIP(hipMemcpyAsync(h_nt, d_ntouched, 4, hipMemcpyDeviceToHost, stq));
          MBA_HIP(hipStreamSynchronize(stq));
          const uint32_t nt = *h_nt;
          n_kept += nt;
          if (nt > 0) {
            MBA_HIP(ctx.vranks.ensure(sizeof(VerifyCand) * nt));
            MBA_HIP(ctx.vok.ensure(nt));
            MBA_HIP(launch_found_gather(d_touched, d_ntouched, nt, d_found, d_flags,
                                        ctx.vranks.as<VerifyCand>(), stq));
            MBA_HIP(launch_verify_batch(W, d_tvs, ctx.leafvals.as<uint64_t>(),
                                        ctx.outputs.as<uint64_t>(),
                                        ctx.b_results.as<BatchInstData>(),
                                        ctx.vranks.as<VerifyCand>(), nt,
                                        ctx.vok.as<uint8_t>(), stq));
            vc.resize(nt);
            okv.resize(nt);
            MBA_HIP(hipMemcpyAsync(vc.data(), ctx.vranks.p, sizeof(VerifyCand) * nt,
                                   hipMemcpyDeviceToHost, stq));
            MBA_HIP(hipMemcpyAsync(okv.data(), ctx.vok.p, nt, hipMemcpyDeviceToHost, stq));
            MBA_HIP(hipStreamSynchronize(stq));
            for (uint32_t c = 0; c < nt; ++c) {
              if (!okv[c]) continue;
              ++n_cands;
              auto &f = S[vc[c].inst].found[vc[c].size];
              f = std::min(f, vc[c].rank);
            }
          }
        }
        t_resolve += seconds_since(t_r0);
        return 0;
      };
      if (nj > 0) {
        if (int rc = run_jobs(0, nj)) return rc;
      }
      if (!learned) break;
      // learned order: any verified solution of this (minimal) size settles
      if (s >= min_size)
        for (size_t i = 0; i < count; ++i)
          if (S[i].active && S[i].found[s] != ~0ull) finish(i, s);
      if (re >= max_blocks) break;
    }
    // sizes completed by this step: 1..2 after the leaf scan, s for s >= 3
    if (s == 0 || s == 2) continue;
    const int lo = s == 1 ? 1 : s, hi = s == 1 ? std::min(2, max_size) : s;
    for (size_t i = 0; i < count; ++i) {
      InstState &st = S[i];
      if (!st.active) continue;
      for (int z = lo; z <= hi; ++z) {
        st.reached = z;
        if (z >= min_size && st.found[z] != ~0ull) {
          finish(i, z);
          break;
        }
      }
    }
  }
  const double total = seconds_since(t_start);
  if (o.verbose)
    std::fprintf(stderr,
                 "batch: jobs %.3f ms, device %.3f ms, hits %.3f ms (%llu hits, %llu lowered "
                 "slots, %llu verified), total %.3f ms\n",
                 1e3 * t_build, 1e3 * t_gpu, 1e3 * t_resolve, (unsigned long long)n_hits,
                 (unsigned long long)n_kept, (unsigned long long)n_cands, 1e3 * total);
  if (o.verbose && count > 0) {
    std::vector<uint32_t> idx(count);
    for (size_t i = 0; i < count; ++i) idx[i] = (uint32_t)i;
    std::partial_sort(idx.begin(), idx.begin() + std::min<size_t>(8, count), idx.end(),
                      [&](uint32_t a, uint32_t b) { return hits_of[a] > hits_of[b]; });
    for (size_t k = 0; k < std::min<size_t>(8, count); ++k) {
      const uint32_t i = idx[k];
      std::fprintf(stderr, "batch: instance %u: %u hits, L=%d n=%d, answer size %d (%s)\n", i,
                   hits_of[i], S[i].L, S[i].n, results[i].size, results[i].expr);
    }
  }
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
