// simba_kernels.h -- launch interface of the SIMBA gfx950 kernels.
#pragma once
#include "mba_common.h"
#include "pair_math.h"

namespace mba {

// Device-resident description of the per-size value tables V[1..Kt]
// (filter-sample value of every tree of that size, in rank order) and of the
// count table.  One instance per solve, read through uniform (scalar) loads.
struct TableView {
  uint64_t off[kMaxSize + 1];     // element offset of V[m] in the table buffer
  uint64_t T[kMaxSize + 1][9];    // count table
  const void *base;               // uint32_t or uint64_t elements
  int Kt;                         // largest tabulated size
  int L;                          // number of leaves
};

// One filter-sample hit: the value c = op(A[a], B[b]) (op-instance `inst`)
// lies in one of the two target windows.  Resolved to (size, rank) on the
// device (k_hit_resolve) and verified on all samples by the verify kernel.
struct HitRec {
  uint64_t a, b, v;
  uint32_t inst;  // 0..6: & | ^ + * (A-B) (B-A); 7: leaf/scan hit
  uint32_t pair;  // (sb << 16) | j
};

// Targets for the fused checks.  A value c is recorded when
//   (c - lo_even) mod 2^W <= span   (c = t0 + k, |k| <= U/2: even unary chains)
// or (c - lo_odd)  mod 2^W <= span   (c = -t0 + k, -(U+1)/2 <= k <= (U-1)/2:
// odd unary chains) with span = U, the number of unary nodes that still fit
// above the candidate.  U = 0 reduces the first window to c == t0.
struct CheckTargets {
  uint64_t lo_even, lo_odd, span;
};

// A block of B-side trees of size m > Kt whose filter-sample values each lane
// forms from tables: an inner value
//   v = op(VA[ra], VB[rb])   (op >= 0),   v = VA[ra]   (op < 0)
// for ra < nA, rb < nB, then the steps of the path from the inner node up to
// the B side's root, innermost first: an affine map x -> s*x + k (a chain of
// unary nodes: ~x = -x - 1, neg x = -x; s = -1 when neg), or a binary node
// whose other child is a fixed tree with filter-sample value V[coff] (kind 1:
// op(V[coff], x), kind 2: op(x, V[coff])).  The tree's rank within size m is
// start + stride * (ra * nB + rb).  A tile covers 2^logTB columns (over ra
// when trans, else over rb) by (256*RB) >> logTB rows; ntc tiles span a row
// band (ntc_div: the division by ntc).  When the columns span a whole tile
// (2^logTB = 256*RB) a tile covers rows_pt consecutive rows of its columns.
// generic: ranks start + stride * i, i < nA, each evaluated by a
// table-assisted unrank (children above Kt).
// Table references (offA, offB, the fixed siblings' c) name a table and an
// element of it, (m << kTabShift) | index, resolved through the instance's
// TableView (tab_off), so one block list serves every instance of the same
// leaf count and tabulated depth.  tile_pre: the B tiles of the blocks
// before this one in its list (a job's B side runs through a whole list).
constexpr int kMaxBSteps = 8;
constexpr uint32_t kBandRows = 8;  // rows of a row-band tile (at most)
constexpr int kTabShift = 56;
constexpr uint64_t kTabIndex = (1ull << kTabShift) - 1;
__host__ __device__ inline uint64_t tab_ref(int m, uint64_t index) {
  return ((uint64_t)m << kTabShift) | index;
}
__device__ inline uint64_t tab_off(const TableView *tv, uint64_t ref) {
  return tv->off[ref >> kTabShift] + (ref & kTabIndex);
}
struct BStep {
  int32_t kind;  // 0 affine, 1 op(c, x), 2 op(x, c)
  int32_t op;    // kinds 1-2: the binary operator; kind 0: neg flag
  uint64_t c;    // kind 0: k; kinds 1-2: table reference of the fixed tree
};
struct BBlock {
  uint64_t start, stride;
  uint64_t offA, offB;     // table references (tab_ref)
  uint64_t nA, nB;
  uint64_t ntc;
  int32_t op, generic, trans, logTB, nsteps, rows_pt;
  DivMagic ntc_div;
  uint32_t pad2;
  uint64_t tile_pre;       // B tiles of the earlier blocks of its list
  BStep steps[kMaxBSteps];
};

struct PairCore {
  int sb, j, m;            // candidate size sb = j + m + 1, j <= m
  uint64_t nA, nB;         // T[j], T[m]
  uint64_t tile_begin, tile_end;  // tile range of this launch
  uint64_t nTA;            // number of A tiles
  uint64_t nTB;            // number of B tiles (of the block when gen_b)
  bool gen_a, gen_b;       // side is above Kt: values computed per lane
  bool tri;                // j == m, both sides tabulated, checked as a j < m
                           // split (all seven instances) over the tiles that
                           // reach A index <= B index only
  uint32_t mask;           // operator instances to check (bit i = inst i)
  int rb;                  // B values per lane: 4 or 8 (tile = 256 * rb B values)
  uint32_t ta;             // A values per tile (<= kPairTA; nTA = ceil(nA / ta))
};

struct PairLaunch : PairCore {
  BBlock bb;               // the B block (gen_b, per-instance launches)
};

constexpr uint32_t kAllInst = 0x7f;

constexpr int kPairThreads = 256;
constexpr int kPairRB = 8;                      // B values per lane (batched kernel)
constexpr int kPairTB = kPairThreads * kPairRB; // B values per tile
constexpr int kPairTA = 256;                    // A values per tile

// ---- batched search --------------------------------------------------------
constexpr int kBuildTile = 1024;  // table values per build tile

// One unit of a batch step: kind 0 builds V[p.m] of instance `inst` (p.nB
// values), kind 1 checks the pair {p.j, p.m} at size p.sb, kind 2 scans the
// leaves.  Jobs own the global tile range [tile0, tile0 + ntiles): global
// tile t is the job's tile p.tile_begin + (t - tile0) (p.tile_begin is 0
// unless the host runs a job's tiles in parts).
struct BatchJob {
  uint64_t tile0, ntiles;
  uint64_t lo_e, lo_o, span;  // windows (kinds 1 and 2)
  PairCore p;                 // p.nTB: tiles of the B block when p.gen_b
  uint32_t inst;
  uint32_t kind;
  uint32_t bb;                // B side above Kt: its first block (index in the blocks)
  uint32_t nbb;               // blocks of the B side (0 or 1: block bb alone); the
                              // job's B tiles run through blocks [bb, bb + nbb)
  uint32_t bsub;              // tabulated B side: B tiles per job tile (0 or 1: one),
                              // so a job of few A values has tiles of more pairs
  uint32_t pad;
};

// One size step of a batch: its jobs are jobs_base[job_off, job_off + njobs)
// (global tiles [0, total_tiles)); once its hits are resolved, every size up
// to size_done is complete for the instances [inst_lo, inst_hi).
// The step's jobs with a tabulated B side (and table builds, leaf scans) come
// first, tiles [0, tiles_table); the jobs whose B side is formed from blocks
// follow, tiles [tiles_table, total_tiles): each part runs in a kernel of
// its own (the first needs far fewer registers, so more waves hide its
// memory latency).
struct BatchStep {
  uint64_t total_tiles;
  uint64_t tiles_table;
  uint32_t job_off;
  uint32_t njobs;
  int32_t size_done;
  uint32_t inst_lo, inst_hi;
  uint32_t index;  // position of the step in its launch sequence
};

// Device counters of a batch step: window hits, screened candidates,
// lowered slots; across a launch sequence, the largest hit count of a step
// above the buffer and the first step that overflowed it (~0: none).
struct BatchCounters {
  uint32_t hits, surv, touched, overflow, first_overflow, pad[3];
};

// A value table V[m] placed at element `off` of a batch's table pool for
// instance `inst`, whose largest tabulated size is then `kt`.
struct TvUpdate {
  uint32_t inst, m;
  uint64_t off;
  int32_t kt, pad;
};
// Applies n table placements to the device table views (records of one
// instance carry the same kt).
hipError_t launch_tv_update(TableView *d_tvs, const TvUpdate *d_up, uint32_t n,
                            hipStream_t stream);

struct BatchInstData {
  uint64_t leaf_off;   // [sample][leaf] values of this instance in the batch array
  uint64_t out_off;    // outputs of this instance
  uint32_t n_samples;
  uint32_t fs;         // filter sample (the tables' sample)
};

struct VerifyCand {
  uint32_t inst, size;
  uint64_t rank;
};

// Resets the step's work-queue head and the hit, survivor and lowered-slot
// counts (the overflow record is kept).
hipError_t launch_batch_reset(unsigned long long *d_head, BatchCounters *d_cnt,
                              hipStream_t stream);
// After a step's hits are resolved: for i in [step->inst_lo, step->inst_hi),
// settled[i] = the smallest size z in [min_size, step->size_done] whose slot
// d_found[i * (kMaxSize + 1) + z] holds a verified rank (instances already
// settled, settled[i] > 0, keep theirs; settled[i] = -y < 0 bounds z below
// y: the instance has swept part of size y only), and the step's hit count
// is recorded in d_cnt->overflow (first_overflow: the step's index) when it
// exceeded the buffer.  The grid covers `count` instances (the largest
// range).
hipError_t launch_batch_settle(const BatchStep *d_step, const unsigned long long *d_found,
                               int32_t *d_settled, uint32_t count, int min_size,
                               BatchCounters *d_cnt, uint32_t hit_cap, hipStream_t stream);
// V[1] (filter-sample leaf values) of every instance of a batch whose view
// has Kt > 0, one launch;
hipError_t launch_batch_leaves(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                               const BatchInstData *d_idat, uint32_t count,
                               hipStream_t stream);
// The persistent kernel of a step: jobs d_jobs_base + step->job_off, their
// B blocks d_blocks[job.bb] (at least one element: table jobs name block
// 0); jobs of instances with d_settled[inst] > 0 are skipped (d_settled may
// be NULL).
hipError_t launch_batch_step(int width, const BatchStep *d_step, const BatchJob *d_jobs_base,
                             const BBlock *d_blocks, const TableView *d_tvs,
                             unsigned long long *d_head, HitRec *d_hits,
                             uint32_t *d_hit_count, uint32_t hit_cap,
                             const int32_t *d_settled, int grid, hipStream_t stream);
// Full verification of (instance, size, rank) candidates on all samples, one
// wave per candidate (lane per sample, leaf values and RPN stacks in LDS, exit
// at the first mismatching sample): ok[i] = 1 when candidate i matches all.
// With d_n, only candidates [0, *d_n) (n_cand is then the bound the grid
// covers).
hipError_t launch_verify_batch(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                               const uint64_t *d_outputs, const BatchInstData *d_idat,
                               const VerifyCand *d_cand, uint32_t n_cand, uint8_t *d_ok,
                               hipStream_t stream, const uint32_t *d_n = nullptr);

// All launches are asynchronous on `stream`.
// V[m] from V[<m], one launch per (op, left size) block of size m.
hipError_t launch_table_build(int width, void *d_base, const uint64_t *off, const CountTable &T,
                              int m, hipStream_t stream);
// V[m_lo..m_hi] by one workgroup, size after size (small tables: one launch).
hipError_t launch_table_small(int width, const TableView *d_tv, int m_lo, int m_hi,
                              hipStream_t stream);
// d_base/off: the table buffer and the host copy of TableView::off.
hipError_t launch_pair_check(int width, const TableView *d_tv, const void *d_base,
                             const uint64_t *off, const PairLaunch &p,
                             const CheckTargets &tg, HitRec *d_hits,
                             uint32_t *d_hit_count, uint32_t hit_cap, int grid,
                             hipStream_t stream);
hipError_t launch_leaf_scan(int width, const TableView *d_tv, int L,
                            const CheckTargets &tg, HitRec *d_hits,
                            uint32_t *d_hit_count, uint32_t hit_cap,
                            hipStream_t stream);
// Candidate of a window hit that holds on two samples, awaiting the check on
// every sample: slot = instance * (kMaxSize + 1) + size, bit 31 = odd chain.
struct SurvRec {
  uint32_t hit, slot;
  uint64_t rank;
};

// Device resolution of window hits (d_tvs[hit.inst >> 8] is a hit's
// instance, d_idat its samples): each of the hits [0, min(*d_n_hits, cap))
// whose chain candidate holds on every sample lowers
// d_found[inst * (kMaxSize + 1) + size] to the candidate's rank; lowered slots
// are appended once to d_touched (count in *d_n_touched; d_flags all zero
// between launches).  d_surv holds up to 2 * cap candidates, *d_n_surv
// (zeroed by the caller, like *d_n_touched) counts them.  Both kernels are
// grid-stride over device-side counts: screen_blocks waves of 64 hits,
// confirm_blocks waves of one candidate.
hipError_t launch_hit_resolve(int width, const TableView *d_tvs, const uint64_t *d_leafvals,
                              const uint64_t *d_outputs, const BatchInstData *d_idat,
                              const HitRec *d_hits, const uint32_t *d_n_hits, uint32_t cap,
                              uint32_t screen_blocks, uint32_t confirm_blocks, int min_size,
                              int max_size, unsigned long long *d_found, SurvRec *d_surv,
                              uint32_t *d_n_surv, uint32_t *d_flags, uint32_t *d_touched,
                              uint32_t *d_n_touched, hipStream_t stream);
// (instance, size, found rank) of the listed slots [0, *d_n) (at most max_n);
// clears their flags.
hipError_t launch_found_gather(const uint32_t *d_touched, const uint32_t *d_n, uint32_t max_n,
                               const unsigned long long *d_found, uint32_t *d_flags,
                               VerifyCand *d_out, hipStream_t stream);
// SIMBA-RTid scan of threads [g0, g1) of the (size, op) block at `offset` with
// N trees: *d_best = min(*d_best, g) over threads whose tree matches every
// sample (t0 = the filter sample's output, the tables' sample).
hipError_t launch_rtid_scan(int width, const TableView *d_tv, const uint64_t *d_leafvals,
                            const uint64_t *d_outputs, int n_samples, int size,
                            uint64_t offset, uint64_t N, uint64_t g0, uint64_t g1,
                            uint64_t t0, unsigned long long *d_best, int grid,
                            hipStream_t stream);

}  // namespace mba
