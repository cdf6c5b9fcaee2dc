// cli_vfb.cpp -- vfb_exe: drop-in for the reference's VFB CLI
// (vfb/src/full-cache.cu:776-853).  stdout: "<path>,<formula>,<seconds>", or
// with --print-cache-size one "level,rows" line per level
// (full-cache.cu:763-765).
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
  if (max_len < 1 || max_len > 100) {
    std::fprintf(stderr, "Argument maxLen = %s should be between 1 and 100", a.pos[1].c_str());
    return 0;
  }
  mba_instance inst;
  int rc = mba_instance_read_json(a.pos[0].c_str(), &inst);
  if (rc != 0) {
    std::fprintf(stderr, "Error: Could not open file %s\n", a.pos[0].c_str());
    return 1;
  }
  if (a.width) inst.width = (uint32_t)a.width;
  inst.n_consts = (uint32_t)a.consts.size();
  inst.consts = a.consts.empty() ? nullptr : a.consts.data();
  auto t0 = std::chrono::high_resolution_clock::now();
  std::string output;
  mba_vfb_stats st;
  mba_ctx *ctx = nullptr;
  if (inst.n_samples > 100) {
    std::fprintf(stderr, "This version supports at most %d samples.\n", 100);
    output = "see_the_error";
  } else {
    rc = mba_ctx_create(a.device, &ctx);
    if (rc != 0) {
      std::fprintf(stderr, "device init failed: %s\n", mba_strerror(rc));
      return 1;
    }
    mba_opts o;
    mba_default_opts(&o);
    o.max_size = max_len > MBA_MAX_SIZE ? MBA_MAX_SIZE : max_len;
    o.compat = 1;
    o.exhaustive = a.exhaustive;
    o.time_budget_s = a.budget;
    o.verbose = a.verbose;
    mba_result r;
    rc = mba_vfb_solve(ctx, &inst, &o, &r, &st);
    if (rc < 0) {
      std::fprintf(stderr, "error: %s: %s\n", mba_strerror(rc), mba_ctx_last_error(ctx));
      mba_ctx_destroy(ctx);
      return 1;
    }
    output = r.expr;
    if (a.print_cache) {
      uint64_t cum = 0;
      for (int l = 1; l <= o.max_size && l <= r.size_reached; ++l) {
        cum += st.level_new[l];
        if (l >= 2) std::cout << l << "," << cum << std::endl;
      }
    }
  }
  auto t1 = std::chrono::high_resolution_clock::now();
  double secs = std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / 1e6;
  std::fprintf(stderr, "Input file : %s\n", a.pos[0].c_str());
  std::fprintf(stderr, "Max length : %d\n", max_len);
  if (output == "Not found !" || output == "see_the_error")
    std::fprintf(stderr, "Result     : no expression found within length %d\n", max_len);
  else
    std::fprintf(stderr, "Result     : expression found\n");
  std::fprintf(stderr, "\nRunning Time: %f s", secs);
  if (!a.print_cache) std::cout << a.pos[0] << "," << output << "," << secs << std::endl;
  if (ctx) mba_ctx_destroy(ctx);
  inst.consts = nullptr;
  mba_instance_free(&inst);
  return 0;
}
