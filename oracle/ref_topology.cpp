// Driver of the reference SUS2-SH model generator -- TEST INFRASTRUCTURE ONLY
// (pins superlinear_amd/topology.py at body orders 2..6 and both factor
// prunings).  Linked against the reference's own sources
// (dev_src/sh_model_init.cpp, src/zbl_pair.cpp) by oracle/Makefile into
// oracle/_ref/; nothing in the product path uses it.
//
// Usage: ref-topology OUT.sus2 key=value ...
// Each key is an init-sh option name without its leading dashes (the option
// map WriteSphericalHarmonicModel reads, dev_src/sh_model_init.cpp:1191); a
// bare key stands for a flag ("true").
#include <cstdio>
#include <exception>
#include <map>
#include <string>

#include "sh_model_init.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT.sus2 key=value ...\n", argv[0]);
    return 2;
  }
  std::map<std::string, std::string> opts;
  for (int i = 2; i < argc; ++i) {
    const std::string a = argv[i];
    const size_t eq = a.find('=');
    if (eq == std::string::npos)
      opts[a] = "true";
    else
      opts[a.substr(0, eq)] = a.substr(eq + 1);
  }
  try {
    WriteSphericalHarmonicModel(argv[1], opts);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref-topology: %s\n", e.what());
    return 1;
  }
  return 0;
}
