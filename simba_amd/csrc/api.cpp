// api.cpp -- the extern "C" entry points declared in include/mba_api.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "context.h"

using namespace mba;


extern "C" {

void mba_default_opts(mba_opts *o) {
  std::memset(o, 0, sizeof *o);
  o->min_size = 1;
  o->max_size = 9;
}

const char *mba_strerror(int code) {
  switch (code) {
    case MBA_FOUND: return "solution found";
    case MBA_NOT_FOUND: return "Not found !";
    case MBA_BUDGET: return "time budget or cancel flag reached";
    case MBA_OVERFLOW: return "cache capacity exhausted";
    case MBA_ERR_ARG: return "invalid argument";
    case MBA_ERR_LIMIT: return "instance outside supported limits";
    case MBA_ERR_COUNT: return "count table overflows uint64";
    case MBA_ERR_HIP: return "HIP runtime error";
    case MBA_ERR_NOMEM: return "out of device memory";
    case MBA_ERR_IO: return "instance file missing or malformed";
    case MBA_ERR_NODEV: return "no gfx950 device";
    default: return "unknown status";
  }
}

int mba_ctx_create(int device, mba_ctx **out) {
  if (!out) return MBA_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return MBA_ERR_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MBA_ERR_NODEV;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MBA_ERR_NODEV;
  mba_ctx *c = new (std::nothrow) mba_ctx;
  if (!c) return MBA_ERR_NOMEM;
  c->c.device = device;
  c->c.cu_count = prop.multiProcessorCount;
  c->c.hbm_bytes = prop.totalGlobalMem;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->c.ev0) != hipSuccess || hipEventCreate(&c->c.ev1) != hipSuccess ||
      hipHostMalloc(&c->c.pinned, Ctx::kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
    delete c;
    return MBA_ERR_HIP;
  }
  // the no-alloc search's scratch array: its address on this device
  if (int rc = noalloc_setup(device)) {
    mba_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

void mba_ctx_destroy(mba_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->c.device);
  (void)hipStreamSynchronize(c->c.stream);
  DevBuf *bufs[] = {&c->c.tables, &c->c.tview, &c->c.hits, &c->c.surv, &c->c.hit_count,
                    &c->c.leafvals, &c->c.outputs, &c->c.vranks,
                    &c->c.vok, &c->c.vfb_rows, &c->c.vfb_left, &c->c.vfb_right,
                    &c->c.vfb_keys, &c->c.vfb_idx, &c->c.vfb_flags, &c->c.vfb_misc,
                    &c->c.vfb_counts, &c->c.vfb_target, &c->c.b_inst, &c->c.b_data,
                    &c->c.b_queue, &c->c.b_results, &c->c.b_misc, &c->c.b_tvup, &c->c.mlp_w,
                    &c->c.mlp_misc, &c->c.f_in, &c->c.f_out, &c->c.f_meta, &c->c.f_x};
  for (DevBuf *b : bufs) b->release();
  if (c->c.pinned) (void)hipHostFree(c->c.pinned);
  if (c->c.b_gexec) (void)hipGraphExecDestroy(c->c.b_gexec);
  if (c->c.b_graph) (void)hipGraphDestroy(c->c.b_graph);
  if (c->c.b_pinned) (void)hipHostFree(c->c.b_pinned);
  (void)hipEventDestroy(c->c.ev0);
  (void)hipEventDestroy(c->c.ev1);
  (void)hipStreamDestroy(c->c.stream);
  delete c;
}

// Last error text of a context (diagnostics; not part of the reference).
const char *mba_ctx_last_error(mba_ctx *c) { return c ? c->c.last_error.c_str() : ""; }

int mba_simba_solve(mba_ctx *c, const mba_instance *inst, const mba_opts *opts,
                    mba_result *res) {
  if (!c || !inst || !res) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return simba_solve(c->c, *inst, *opts, *res);
}

int mba_simba_solve_noalloc(mba_ctx *c, const mba_instance *inst, const mba_opts *opts,
                            mba_result *res) {
  if (!c || !inst || !res) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return simba_solve_noalloc(c->c, *inst, *opts, *res);
}

int mba_simba_rtid_solve(mba_ctx *c, const mba_instance *inst, const mba_opts *opts,
                         mba_result *res) {
  if (!c || !inst || !res) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return simba_solve(c->c, *inst, *opts, *res, nullptr, true);
}

int mba_vfb_solve(mba_ctx *c, const mba_instance *inst, const mba_opts *opts,
                  mba_result *res, mba_vfb_stats *stats) {
  if (!c || !inst || !res) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return vfb_solve(c->c, *inst, *opts, *res, stats);
}

int mba_vfb_fingerprints(mba_ctx *c, uint32_t width, const uint64_t *values,
                         uint32_t n_words, size_t count, uint64_t *out) {
  if (!c) return MBA_ERR_ARG;
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return vfb_fingerprints(c->c, (int)width, values, (int)n_words, count, out);
}

int mba_vfb_probe_ceiling(mba_ctx *c, uint32_t slots_log2, uint64_t count, uint64_t distinct,
                          uint64_t prefill, double *probe_s, double *insert_s,
                          uint64_t *fresh) {
  if (!c) return MBA_ERR_ARG;
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return vfb_probe_ceiling(c->c, (int)slots_log2, count, distinct, prefill, probe_s, insert_s,
                           fresh);
}

int mba_batch_solve(mba_ctx *c, const mba_instance *insts, size_t count,
                    const mba_opts *opts, mba_result *results) {
  if (!c || (!insts && count) || (!results && count)) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return batch_solve(c->c, insts, count, *opts, results);
}

int mba_rpn_to_infix(const char *rpn, uint32_t n_vars, const uint64_t *consts,
                     uint32_t n_consts, uint32_t width, char *out, size_t cap) {
  if (!rpn || !out || cap == 0) return MBA_ERR_ARG;
  static const char ops[] = "~&|^#+-*";
  uint8_t tok[MBA_MAX_SIZE];
  int n = 0;
  for (const char *p = rpn; *p && n < MBA_MAX_SIZE; ++p) {
    const char *q = std::strchr(ops, *p);
    if (q && *p) tok[n++] = (uint8_t)(kTokOp + (q - ops));
    else if (*p >= '0' && (uint32_t)(*p - '0') < n_vars + n_consts) tok[n++] = (uint8_t)(*p - '0');
    else return MBA_ERR_ARG;
  }
  std::string s = infix_simba(tok, n, (int)n_vars, consts, (int)width);
  std::snprintf(out, cap, "%s", s.c_str());
  return 0;
}

}  // extern "C"
