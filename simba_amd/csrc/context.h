// context.h -- per-context device state shared by the SIMBA/VFB drivers.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <functional>
#include <string>

#include "../../include/mba_api.h"
#include "mba_common.h"

namespace mba {

// Grow-only device allocation owned by a context.  A fixed buffer is a
// view of memory owned elsewhere: it never allocates, and asking it for more
// than it holds fails.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (fixed) return hipErrorOutOfMemory;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p && !fixed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const { return (T *)p; }
};

struct Ctx {
  int device = 0;
  int cu_count = 256;
  size_t hbm_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string last_error;
  // SIMBA
  DevBuf tables, tview, hits, surv, hit_count, leafvals, outputs, vranks, vok;
  // VFB
  DevBuf vfb_rows, vfb_left, vfb_right, vfb_keys, vfb_idx, vfb_flags, vfb_misc,
      vfb_counts, vfb_target;
  // batched / learned
  DevBuf b_inst, b_data, b_queue, b_results, b_misc, b_tvup, mlp_w, mlp_misc;
  // instance features on the device (packed inputs/outputs, layout, rows)
  DevBuf f_in, f_out, f_meta, f_x;
  // pinned host staging for the per-step read backs of the search drivers
  // (first 4 KiB) and the job lists of fused steps (the rest)
  void *pinned = nullptr;
  static constexpr size_t kPinnedBytes = 64 << 10;
  // batched search: pinned job staging and the captured step graph, kept
  // across calls while the buffers the graph names are unchanged
  void *b_pinned = nullptr;
  size_t b_pinned_bytes = 0;
  hipGraph_t b_graph = nullptr;
  hipGraphExec_t b_gexec = nullptr;
  static constexpr int kGraphKey = 17;
  uint64_t b_graph_key[kGraphKey] = {};
};

using Clock = std::chrono::steady_clock;
inline double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

#define MBA_HIP(call)                                                     \
  do {                                                                    \
    hipError_t _e = (call);                                               \
    if (_e != hipSuccess) {                                               \
      ctx.last_error = std::string(#call) + ": " + hipGetErrorString(_e); \
      return MBA_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

// block_scores: NULL for the plain (rank-order, minimum-rank) search, else
// per (op, left size) scores at [op * 32 + split] that order the binary
// blocks of each size (learned proposal; any solution of minimal size).
// score_size: when given, fills block_scores for a size the first time the
// search orders that size by them (a search settled below runs no network).
// rtid: the SIMBA-RTid order instead (nocache_rtid.cu:85-95).
int simba_solve(Ctx &ctx, const mba_instance &inst, const mba_opts &opts,
                mba_result &res, const float *block_scores = nullptr, bool rtid = false,
                const std::function<int(int)> *score_size = nullptr);
int vfb_solve(Ctx &ctx, const mba_instance &inst, const mba_opts &opts,
              mba_result &res, mba_vfb_stats *stats);
// The SIMBA search with every device buffer carved from the library's
// static scratch array (see noalloc.hip); noalloc_setup finds the array's
// address on the current device (once per device, at context creation).
int noalloc_setup(int device);
int simba_solve_noalloc(Ctx &ctx, const mba_instance &inst, const mba_opts &opts,
                        mba_result &res);
// Device fingerprints of `count` signatures of n words (see mba_api.h).
int vfb_fingerprints(Ctx &ctx, int W, const uint64_t *values, int n, size_t count,
                     uint64_t *out);
// Random-access ceiling of the VFB insert pass (see mba_api.h).
int vfb_probe_ceiling(Ctx &ctx, int slots_log2, uint64_t count, uint64_t distinct,
                      uint64_t prefill, double *probe_s, double *insert_s, uint64_t *fresh);
// score_size: NULL (plain, minimum rank per instance) or a function that
// fills count x 256 block scores (as for simba_solve) for a size; learned
// order: rounds of the best-scored split, the next split, then the rest.
int batch_solve(Ctx &ctx, const mba_instance *insts, size_t count,
                const mba_opts &opts, mba_result *results,
                const std::function<int(int, float *)> *score_size = nullptr);

}  // namespace mba

struct mba_ctx {
  mba::Ctx c;
};

namespace mba {

}  // namespace mba
