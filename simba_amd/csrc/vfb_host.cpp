// vfb_host.cpp -- host driver of the VFB search (placeholder until the
// gfx950 VFB kernels land).
#include "context.h"

namespace mba {
int vfb_solve(Ctx &ctx, const mba_instance &, const mba_opts &, mba_result &,
              mba_vfb_stats *) {
  ctx.last_error = "VFB search not built yet";
  return MBA_ERR_ARG;
}
}  // namespace mba
