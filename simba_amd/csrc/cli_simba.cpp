// cli_simba.cpp -- simba_exe: drop-in for the reference's SIMBA CLI
// (simba/src/nocache.cu:291-342).  stdout: "<path>,<formula>,<seconds>".
// Built with MBA_CLI_RTID as simba_rtid_exe: the SIMBA-RTid order
// (simba/src/nocache_rtid.cu, same CLI).
#include <chrono>
#include <iostream>

#include "cli_common.h"

int main(int argc, char **argv) {
  CliArgs a = parse_cli(argc, argv);
  if (a.pos.size() != 2) {
    usage(argv[0]);
    return 0;
  }
  const int max_len = std::atoi(a.pos[1].c_str());
  if (max_len < 1 || max_len > 50) {
    std::fprintf(stderr, "Argument maxLen = %s should be between 1 and %d", a.pos[1].c_str(),
                 MBA_MAX_SIZE);
    return 0;
  }
  mba_instance inst;
  int rc = mba_instance_read_json(a.pos[0].c_str(), &inst);
  if (rc != 0) {
    std::fprintf(stderr, "Error: Could not open file %s (%s)\n", a.pos[0].c_str(),
                 mba_strerror(rc));
    return 1;
  }
  if (a.width) inst.width = (uint32_t)a.width;
  inst.n_consts = (uint32_t)a.consts.size();
  inst.consts = a.consts.empty() ? nullptr : a.consts.data();
  if ((a.native ? max_len : max_len - 1) > MBA_MAX_SIZE) {
    std::fprintf(stderr, "This version supports formulas of size at most %d\n", MBA_MAX_SIZE);
    return 1;
  }
  // Timing starts after parsing and includes device initialisation, as the
  // reference's (nocache.cu:323-338).
  auto t0 = std::chrono::high_resolution_clock::now();
  mba_ctx *ctx = nullptr;
  rc = mba_ctx_create(a.device, &ctx);
  if (rc != 0) {
    std::fprintf(stderr, "device init failed: %s\n", mba_strerror(rc));
    return 1;
  }
  mba_opts o;
  mba_default_opts(&o);
  o.max_size = max_len;
  o.compat = a.native ? 0 : 1;
  o.time_budget_s = a.budget;
  o.verbose = a.verbose;
  mba_result r;
#ifdef MBA_CLI_RTID
  rc = mba_simba_rtid_solve(ctx, &inst, &o, &r);
#else
  rc = mba_simba_solve(ctx, &inst, &o, &r);
#endif
  auto t1 = std::chrono::high_resolution_clock::now();
  if (rc < 0) {
    std::fprintf(stderr, "error: %s: %s\n", mba_strerror(rc), mba_ctx_last_error(ctx));
    mba_ctx_destroy(ctx);
    mba_instance_free(&inst);
    return 1;
  }
  double secs = std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / 1e6;
  std::fprintf(stderr, "\nRunning Time: %f s", secs);
  std::cout << a.pos[0] << "," << r.expr << "," << secs << std::endl;
  mba_ctx_destroy(ctx);
  inst.consts = nullptr;
  mba_instance_free(&inst);
  return 0;
}
