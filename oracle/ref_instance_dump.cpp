// Driver for the reference's SIMBA instance reader (simba/src/instance.cpp,
// MbaInstance::from_json_file, instance.h:19-29).  Prints every parsed value
// so tests can compare the product's JSON reader against the reference's.
// Test infrastructure only; built into oracle/_ref/ by oracle/Makefile.
#include "instance.h"
#include <cstdio>

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  MbaInstance m = MbaInstance::from_json_file(argv[1]);
  std::printf("%d %d\n", m.n_variables, m.n_examples);
  for (size_t i = 0; i < m.inputs.size(); ++i) std::printf("%u ", m.inputs[i]);
  std::printf("\n");
  for (size_t i = 0; i < m.outputs.size(); ++i) std::printf("%u ", m.outputs[i]);
  std::printf("\n");
  return 0;
}
