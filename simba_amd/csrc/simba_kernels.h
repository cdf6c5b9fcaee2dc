// simba_kernels.h -- launch interface of the SIMBA gfx950 kernels.
#pragma once
#include "mba_common.h"

namespace mba {

// Device-resident description of the per-size value tables V[1..Kt]
// (sample-0 value of every tree of that size, in rank order) and of the
// count table.  One instance per solve, read through uniform (scalar) loads.
struct TableView {
  uint64_t off[kMaxSize + 1];     // element offset of V[m] in the table buffer
  uint64_t T[kMaxSize + 1][9];    // count table
  const void *base;               // uint32_t or uint64_t elements
  int Kt;                         // largest tabulated size
  int L;                          // number of leaves
};

// One sample-0 hit: the value c = op(A[a], B[b]) (op-instance `inst`) lies in
// one of the two target windows.  Resolved to (size, rank) on the host and
// verified on all samples by the verify kernel.
struct HitRec {
  uint64_t a, b, v;
  uint32_t inst;  // 0..6: & | ^ + * (A-B) (B-A); 7: leaf/scan hit
  uint32_t pair;  // (sb << 16) | j
};

// Targets for the fused checks.  A value c is recorded when
//   (c - lo_even) mod 2^W <= span   (c = t0 + k,  |k| <= U: even unary chains)
// or (c - lo_odd)  mod 2^W <= span   (c = -t0 + k: odd unary chains)
// with span = 2U.  U = 0 reduces the first window to c == t0.
struct CheckTargets {
  uint64_t lo_even, lo_odd, span;
};

struct PairLaunch {
  int sb, j, m;            // candidate size sb = j + m + 1, j <= m
  uint64_t nA, nB;         // T[j], T[m]
  uint64_t tile_begin, tile_end;  // tile range of this launch
  uint64_t nTA;            // number of A tiles
  bool gen_a, gen_b;       // side is above Kt: values computed per lane
};

constexpr int kPairThreads = 256;
constexpr int kPairRB = 4;                      // B values held per lane
constexpr int kPairTB = kPairThreads * kPairRB; // B values per tile
constexpr int kPairTA = 256;                    // A values per tile

// All launches are asynchronous on `stream`.
hipError_t launch_table_build(int width, const TableView *d_tv, void *d_base,
                              uint64_t off, int m, uint64_t n, hipStream_t stream);
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
// Full verification of (size, rank) candidates on all samples: unrank and
// evaluate every sample (leaf values [sample][leaf] staged in LDS), exit at
// the first mismatching sample.  ok[i] = 1 when candidate i matches all.
hipError_t launch_verify(int width, const TableView *d_tv, const uint64_t *d_leafvals,
                         const uint64_t *d_outputs, int n_samples,
                         const uint32_t *d_sizes, const uint64_t *d_ranks,
                         uint32_t n_cand, uint8_t *d_ok, hipStream_t stream);

}  // namespace mba
