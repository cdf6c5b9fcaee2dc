// noalloc.hip -- the SIMBA search without device allocation.
//
// Callers that cannot allocate device memory during a call (tight serving
// loops, harnesses that forbid it) use mba_simba_solve_noalloc: every device
// buffer of the search -- value tables, window-hit records, screened
// candidates, the resolution slots and the fused steps' job staging -- is a
// fixed region of one file-scope device array, `scratch`.  Each device holds
// its own copy of the array (for the life of the process); its address on a
// device comes from hipGetSymbolAddress, called once per device when a
// context is created there, and only that address is ever handed to kernels
// or copies.  The search initialises every region it reads before reading
// it, so a call sees nothing of the previous one; calls that overlap on one
// device share the array and race.
#include <mutex>

#include "context.h"
#include "simba_kernels.h"

namespace mba {

namespace {

constexpr size_t kScratchBytes = 64ull << 20;
constexpr uint32_t kScratchHits = 1u << 17;  // window-hit records
// fixed regions (bytes); the value tables take the rest
constexpr size_t kViewBytes = 4096, kCountBytes = 256, kLeafBytes = 16384, kOutBytes = 1024,
                 kMiscBytes = 4096, kQueueBytes = 64 << 10;
constexpr size_t kHitBytes = sizeof(HitRec) * (size_t)kScratchHits;
constexpr size_t kSurvBytes = sizeof(SurvRec) * 2 * (size_t)kScratchHits;
constexpr size_t kFixedBytes =
    kViewBytes + kCountBytes + kLeafBytes + kOutBytes + kMiscBytes + kQueueBytes + kHitBytes +
    kSurvBytes;
constexpr size_t kTableBytes = kScratchBytes - kFixedBytes;
static_assert(kFixedBytes % 256 == 0 && kFixedBytes < kScratchBytes, "scratch layout");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // the array's address on each device

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return MBA_ERR_NODEV;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return 0;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MBA_ERR_HIP;
  g_scratch[device] = p;
  return 0;
}

int simba_solve_noalloc(Ctx &ctx, const mba_instance &inst, const mba_opts &o,
                        mba_result &res) {
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx.device >= 0 && ctx.device < kMaxDevices) base = g_scratch[ctx.device];
  }
  if (!base) {
    ctx.last_error = "no-alloc search: scratch address unknown on this device";
    return MBA_ERR_HIP;
  }
  if (o.table_bytes > kTableBytes) {
    ctx.last_error = "no-alloc search: table_bytes above the scratch's " +
                     std::to_string(kTableBytes) + " bytes of value tables";
    return MBA_ERR_LIMIT;
  }
  if (o.hit_cap < 0 || (uint32_t)o.hit_cap > kScratchHits) {
    ctx.last_error = "no-alloc search: hit_cap above the scratch's " +
                     std::to_string(kScratchHits) + " window-hit records";
    return MBA_ERR_LIMIT;
  }
  // a context view: the caller's stream, events and pinned staging, and
  // fixed device buffers over the regions of the scratch array
  Ctx s;
  s.device = ctx.device;
  s.cu_count = ctx.cu_count;
  s.hbm_bytes = ctx.hbm_bytes;
  s.stream = ctx.stream;
  s.ev0 = ctx.ev0;
  s.ev1 = ctx.ev1;
  s.pinned = ctx.pinned;
  char *p = (char *)base;
  auto carve = [&p](DevBuf &b, size_t n) {
    b.p = p;
    b.bytes = n;
    b.fixed = true;
    p += n;
  };
  carve(s.tview, kViewBytes);
  carve(s.hit_count, kCountBytes);
  carve(s.leafvals, kLeafBytes);
  carve(s.outputs, kOutBytes);
  carve(s.b_misc, kMiscBytes);
  carve(s.b_queue, kQueueBytes);
  carve(s.hits, kHitBytes);
  carve(s.surv, kSurvBytes);
  carve(s.tables, kTableBytes);
  mba_opts q = o;
  if (q.table_bytes == 0) q.table_bytes = kTableBytes;
  if (q.hit_cap == 0) q.hit_cap = (int32_t)kScratchHits;
  const int rc = simba_solve(s, inst, q, res);
  if (rc < 0) ctx.last_error = s.last_error;
  return rc;
}

}  // namespace mba
