// Model shapes: the table of kernel instances compiled into the library and,
// for any other (L, K, B, warp-sharing mode), a shape library compiled at
// first use with hipcc from the same kernel sources (csrc/ beside the
// library), cached on disk and loaded with dlopen.  The reference builds its
// kernels for any K and B (interfaces/lammps/ML-SUS2/sus2_sl_model.cpp:276,
// runtime loops); here every shape is a compile-time instance, so a new shape
// costs one compilation (about a minute) and is then as fast as the prebuilt
// ones.
//
// Cache: $SUS2_JIT_DIR, else `_jit/` beside libsus2amd.so; one file
// sus2shape_l<L>k<K>b<B>m<SM>_<source hash>.so per shape, so a rebuilt
// library never loads a shape library of other sources.  SUS2_JIT=0 loads
// cached shape libraries only (no compiler run).  $SUS2_HIPCC overrides the
// compiler (default $ROCM_PATH/bin/hipcc, else /opt/rocm/bin/hipcc).
#include <dlfcn.h>
#include <fcntl.h>
#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "sus2_common.hpp"

#ifndef SUS2_SRC_HASH
#define SUS2_SRC_HASH "dev"
#endif
#ifndef SUS2_OFFLOAD_ARCH
#define SUS2_OFFLOAD_ARCH "gfx950"
#endif

extern char** environ;

namespace sus2 {

#define SUS2_DECLARE(LL, KK, BB, SM)                                                                               \
  extern template void run_efs_shape<Cfg<float, LL, KK, BB, SM>>(const DevModelSet&, Workspace&, const Batch&,    \
                                                                 const EfsOut&, hipStream_t);                      \
  extern template void run_efs_shape<Cfg<double, LL, KK, BB, SM>>(const DevModelSet&, Workspace&, const Batch&,   \
                                                                  const EfsOut&, hipStream_t);                     \
  extern template void run_grad_shape<Cfg<float, LL, KK, BB, SM>>(const DevModelSet&, Workspace&, const Batch&,   \
                                                                  const GradWeights&, double*, double, bool, bool, \
                                                                  hipStream_t);                                    \
  extern template void run_grad_shape<Cfg<double, LL, KK, BB, SM>>(const DevModelSet&, Workspace&, const Batch&,  \
                                                                   const GradWeights&, double*, double, bool, bool, \
                                                                   hipStream_t);                                    \
  extern template void pair_stage_shape<Cfg<float, LL, KK, BB, SM>>(int, const DevModelSet&, Workspace&,           \
                                                                    const PairView&, hipStream_t);                 \
  extern template void pair_stage_shape<Cfg<double, LL, KK, BB, SM>>(int, const DevModelSet&, Workspace&,          \
                                                                     const PairView&, hipStream_t);
SUS2_FOR_EACH_SHAPE(SUS2_DECLARE)
#undef SUS2_DECLARE

namespace {

using Key = std::tuple<int, int, int, int>;

std::mutex g_mu;
std::map<Key, std::unique_ptr<ShapeOps>>& registry() {
  static std::map<Key, std::unique_ptr<ShapeOps>> r;
  return r;
}

bool prebuilt(const Key& k, ShapeOps* out) {
  const auto [L, K, B, SM] = k;
#define SUS2_SHAPE(LL, KK, BB, SMM)                 \
  if (L == LL && K == KK && B == BB && SM == SMM) { \
    *out = make_shape_ops<LL, KK, BB, SMM>();       \
    return true;                                    \
  }
  SUS2_FOR_EACH_SHAPE(SUS2_SHAPE)
#undef SUS2_SHAPE
  return false;
}

std::string env_or(const char* name, const std::string& d) {
  const char* v = std::getenv(name);
  return (v && *v) ? std::string(v) : d;
}

// Directory of this library (dladdr of a function defined here).
std::string library_dir() {
  Dl_info info{};
  if (!dladdr(reinterpret_cast<void*>(&library_dir), &info) || !info.dli_fname)
    throw Error(SUS2_E_UNSUPPORTED, "shape library: cannot locate libsus2amd.so");
  std::string p = info.dli_fname;
  const size_t s = p.rfind('/');
  return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

bool exists(const std::string& p) {
  struct stat st{};
  return ::stat(p.c_str(), &st) == 0;
}

std::string shape_name(const Key& k) {
  const auto [L, K, B, SM] = k;
  std::ostringstream o;
  o << "l" << L << "k" << K << "b" << B << "m" << SM;
  return o.str();
}

// Limits of the kernels: one lane per gate channel (F <= 64), the radial
// basis of at most BMAX terms, warp-sharing modes 0-3.
void check_shape(const Key& k) {
  const auto [L, K, B, SM] = k;
  if (L < 0 || L > 6 || K < 1 || K * (L + 1) > WAVE || B < 1 || B > BMAX || SM < 0 || SM > 3)
    throw Error(SUS2_E_UNSUPPORTED, "no kernels for L=" + std::to_string(L) + " K=" + std::to_string(K) +
                                        " B=" + std::to_string(B) + " scaling_map mode " + std::to_string(SM) +
                                        " (limits: L <= 6, K (L + 1) <= 64, B <= " + std::to_string(BMAX) + ")");
}

// Run argv[0] with argv as a child process, stdout and stderr to `log`;
// returns the exit status.
int run_logged(const std::vector<std::string>& args, const std::string& log) {
  std::vector<char*> argv;
  for (const auto& a : args) argv.push_back(const_cast<char*>(a.c_str()));
  argv.push_back(nullptr);
  posix_spawn_file_actions_t fa;
  posix_spawn_file_actions_init(&fa);
  posix_spawn_file_actions_addopen(&fa, 1, log.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  posix_spawn_file_actions_adddup2(&fa, 1, 2);
  pid_t pid = 0;
  const int rc = posix_spawnp(&pid, argv[0], &fa, nullptr, argv.data(), environ);
  posix_spawn_file_actions_destroy(&fa);
  if (rc != 0) return -1;
  int status = 0;
  while (waitpid(pid, &status, 0) < 0) {
    if (errno != EINTR) return -1;
  }
  return WIFEXITED(status) ? WEXITSTATUS(status) : -1;
}

std::string tail_of(const std::string& path, size_t n) {
  std::ifstream in(path);
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string s = ss.str();
  return s.size() > n ? s.substr(s.size() - n) : s;
}

}  // namespace

bool is_prebuilt_shape(int L, int K, int B, int SM) {
  ShapeOps o;
  return prebuilt(Key{L, K, B, SM}, &o);
}

// Path of the shape library of (L, K, B, SM) for these sources.
std::string shape_library_path(int L, int K, int B, int SM) {
  const std::string dir = env_or("SUS2_JIT_DIR", library_dir() + "/_jit");
  return dir + "/sus2shape_" + shape_name({L, K, B, SM}) + "_" + SUS2_SRC_HASH + ".so";
}

// Compile the shape library of (L, K, B, SM) unless it is cached; returns its
// path.  No GPU is needed.
std::string build_shape_library(int L, int K, int B, int SM) {
  const Key k{L, K, B, SM};
  check_shape(k);
  const std::string out = shape_library_path(L, K, B, SM);
  if (exists(out)) return out;
  if (env_or("SUS2_JIT", "1") == "0")
    throw Error(SUS2_E_UNSUPPORTED, "no kernels for shape " + shape_name(k) + " (" + out +
                                        " absent and runtime compilation disabled by SUS2_JIT=0)");
  const std::string libdir = library_dir();
  const std::string dir = out.substr(0, out.rfind('/'));
  ::mkdir(dir.c_str(), 0755);
  const std::string tag = std::to_string(::getpid());
  const std::string src = out + "." + tag + ".hip", tmp = out + "." + tag + ".tmp", log = out + "." + tag + ".log";
  {
    std::ofstream f(src);
    f << "// shape library " << shape_name(k) << " (generated)\n"
      << "#include \"sus2_grad_impl.hpp\"\n"
      << "extern \"C\" __attribute__((visibility(\"default\"))) int sus2_shape_ops_fill(sus2::ShapeOps* o, int abi) {\n"
      << "  if (abi != sus2::SHAPE_ABI) return 1;\n"
      << "  *o = sus2::make_shape_ops<" << L << ", " << K << ", " << B << ", " << SM << ">();\n"
      << "  return 0;\n}\n";
    if (!f) throw Error(SUS2_E_IO, "shape library: cannot write " + src);
  }
  const std::string hipcc =
      env_or("SUS2_HIPCC", env_or("ROCM_PATH", "/opt/rocm") + "/bin/hipcc");
  const std::vector<std::string> args = {hipcc, "-O3", "-std=c++17", "-fPIC", "-shared",
                                         std::string("--offload-arch=") + SUS2_OFFLOAD_ARCH, "-Xclang",
                                         "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics",
                                         "-I" + libdir + "/csrc", src, "-L" + libdir, "-lsus2amd",
                                         "-Wl,-rpath," + libdir, "-o", tmp};
  const int rc = run_logged(args, log);
  ::unlink(src.c_str());
  if (rc != 0) {
    const std::string msg = tail_of(log, 2000);
    ::unlink(tmp.c_str());
    throw Error(SUS2_E_UNSUPPORTED, "shape library " + shape_name(k) + ": compilation failed (" + hipcc +
                                        ", status " + std::to_string(rc) + ")\n" + msg);
  }
  ::unlink(log.c_str());
  if (::rename(tmp.c_str(), out.c_str()) != 0) throw Error(SUS2_E_IO, "shape library: cannot place " + out);
  return out;
}

const ShapeOps& shape_ops(int L, int K, int B, int SM) {
  const Key k{L, K, B, SM};
  std::lock_guard<std::mutex> lock(g_mu);
  auto& reg = registry();
  auto it = reg.find(k);
  if (it != reg.end()) return *it->second;
  auto ops = std::make_unique<ShapeOps>();
  if (!prebuilt(k, ops.get())) {
    const std::string path = build_shape_library(L, K, B, SM);
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) throw Error(SUS2_E_UNSUPPORTED, std::string("shape library: ") + dlerror());
    using Fill = int (*)(ShapeOps*, int);
    auto fill = reinterpret_cast<Fill>(dlsym(h, "sus2_shape_ops_fill"));
    if (!fill || fill(ops.get(), SHAPE_ABI) != 0)
      throw Error(SUS2_E_UNSUPPORTED, "shape library " + path + " does not match this library");
    // the handle stays open for the life of the process (its kernels are in use)
  }
  return *(reg[k] = std::move(ops));
}

}  // namespace sus2
