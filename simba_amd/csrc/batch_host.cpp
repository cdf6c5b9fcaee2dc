// batch_host.cpp -- batched SIMBA driver (placeholder until the persistent
// kernel lands).
#include "context.h"

namespace mba {
int batch_solve(Ctx &ctx, const mba_instance *, size_t, const mba_opts &,
                mba_result *) {
  ctx.last_error = "batched search not built yet";
  return MBA_ERR_ARG;
}
}  // namespace mba
