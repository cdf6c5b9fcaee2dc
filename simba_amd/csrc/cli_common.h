// cli_common.h -- argument handling shared by simba_exe and vfb_exe.
//
// Positional arguments follow the reference: <instance.json> <maxLen>
// (nocache.cu:291-313, full-cache.cu:776-798).  Extra long options:
//   --native          sizes 1..maxLen inclusive (default: reference ranges)
//   --consts a,b,...  constant leaves appended after the variables
//   --width 32|64     override the word width read from the instance
//   --budget S        time budget in seconds
//   --device N        HIP device
//   --verbose         per-size progress on stderr
//   --print-cache-size (vfb_exe) print "level,rows" lines instead of the CSV
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mba_api.h"

struct CliArgs {
  std::vector<std::string> pos;
  bool native = false, verbose = false, print_cache = false, exhaustive = false;
  int width = 0, device = 0;
  double budget = 0;
  std::vector<uint64_t> consts;
};

inline CliArgs parse_cli(int argc, char **argv) {
  CliArgs a;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto next = [&]() -> std::string { return i + 1 < argc ? argv[++i] : ""; };
    if (s == "--native") a.native = true;
    else if (s == "--verbose") a.verbose = true;
    else if (s == "--print-cache-size") a.print_cache = true;
    else if (s == "--exhaustive") a.exhaustive = true;
    else if (s == "--width") a.width = std::atoi(next().c_str());
    else if (s == "--device") a.device = std::atoi(next().c_str());
    else if (s == "--budget") a.budget = std::atof(next().c_str());
    else if (s == "--consts") {
      std::string v = next();
      size_t p = 0;
      while (p < v.size()) {
        size_t q = v.find(',', p);
        if (q == std::string::npos) q = v.size();
        a.consts.push_back((uint64_t)std::strtoll(v.substr(p, q - p).c_str(), nullptr, 0));
        p = q + 1;
      }
    } else a.pos.push_back(s);
  }
  return a;
}

inline void usage(const char *prog) {
  std::fprintf(stderr, "Arguments should be in the form of\n");
  std::fprintf(stderr, "-----------------------------------------------------------------\n");
  std::fprintf(stderr, "%s <input_file_address> <maxLen>\n", prog);
  std::fprintf(stderr, "-----------------------------------------------------------------\n");
}
