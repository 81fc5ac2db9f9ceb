// Shared pieces of the SUS2-SH HIP sources: neighbour records, launch
// macro with the optional per-kernel HIP-event profiler.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <string>

#include "../../include/sus2amd.h"
#include "sus2_device.hpp"
#include "sus2_model.hpp"
#include "sus2_runtime.hpp"

namespace sus2 {

// One directed edge i -> j: d = x_j - x_i (+ image), and the neighbour index
// j with its species t_j packed into the bits of the fourth component (one
// 16 / 32-byte load per edge, no dependent species lookup).  FP32: j in the
// low 26 bits, t_j in the high 6; FP64: j in the low 32, t_j in the high 32.
template <class T>
struct alignas(4 * sizeof(T)) Rec {
  T x, y, z, jb;
};
constexpr int REC32_JBITS = 26;
constexpr long REC32_MAX_ATOMS = 1L << REC32_JBITS;
constexpr int REC32_MAX_SPECIES = 1 << (32 - REC32_JBITS);
__device__ __forceinline__ int rec_j(const Rec<float>& r) {
  return __float_as_int(r.jb) & ((1 << REC32_JBITS) - 1);
}
__device__ __forceinline__ int rec_t(const Rec<float>& r) {
  return (int)((unsigned)__float_as_int(r.jb) >> REC32_JBITS);
}
__device__ __forceinline__ int rec_j(const Rec<double>& r) {
  return (int)(__double_as_longlong(r.jb) & 0xffffffffLL);
}
__device__ __forceinline__ int rec_t(const Rec<double>& r) { return (int)(__double_as_longlong(r.jb) >> 32); }
__device__ __forceinline__ float rec_jbits(int j, int t, float) {
  return __int_as_float((int)((unsigned)j | ((unsigned)t << REC32_JBITS)));
}
__device__ __forceinline__ double rec_jbits(int j, int t, double) {
  return __longlong_as_double(((long long)t << 32) | (unsigned)j);
}
// The FP32 packing bounds the batch (2^26 atoms, far beyond one GPU's memory
// at FP32) and the species count (64).
inline void check_rec_limits(int precision, long natoms, int species) {
  if (precision == SUS2_FP32 && (natoms >= REC32_MAX_ATOMS || species > REC32_MAX_SPECIES))
    throw Error(SUS2_E_UNSUPPORTED, "FP32 batches hold fewer than 2^26 atoms and at most 64 species (use SUS2_FP64)");
}

// Frame of atom i: the largest f with frame_off[f] <= i (binary search).
__device__ __forceinline__ int frame_of(const int64_t* off, int nframes, int i) {
  int lo = 0, hi = nframes;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

#define SUS2_CHECK(x)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) throw ::sus2::Error(SUS2_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Launch with per-kernel event timing when the workspace profiler is on.
// Launch a kernel; with the profiler on, its start/stop events are attached to
// the dispatch packet itself (hipExtLaunchKernel), so the timings are the
// kernel's own execution span rather than that of marker packets around it.
#define SUS2_LAUNCH(ws, st, name, kern, grid, block, lds, ...)                                      \
  do {                                                                                              \
    const int ev_ = (ws).prof.begin((name));                                                        \
    if (ev_ < 0) {                                                                                  \
      hipLaunchKernelGGL(kern, grid, block, lds, st, __VA_ARGS__);                                  \
    } else {                                                                                        \
      hipExtLaunchKernelGGL(kern, grid, block, lds, st, (ws).prof.start_event(ev_),                 \
                            (ws).prof.stop_event(ev_), 0, __VA_ARGS__);                             \
    }                                                                                               \
    SUS2_CHECK(hipGetLastError());                                                                  \
  } while (0)

template <class T> constexpr const char* tname();
template <> constexpr const char* tname<float>() { return "f32"; }
template <> constexpr const char* tname<double>() { return "f64"; }

template <class T>
void build_neighbors(Workspace& ws, const Batch& b, double rc, hipStream_t st);
// Reads the overflow report of a build that sized its edge buffers from earlier
// calls (waits for the stream); true when the lists did not fit and the work
// queued after the build has to be redone.  False when no report is pending.
bool neighbors_verify(Workspace& ws, hipStream_t st);
// Per-species site-energy constant species[t] + shift[t] (FP64 coefficients).
struct SiteConst {
  const double* theta;
  int o_species, o_shift, C;
  __device__ __forceinline__ double operator()(int t) const { return theta[o_species + t] + theta[o_shift + t]; }
};
// Forces of atoms [0, b.natoms), site energies of atoms [0, nsite) (all when
// nsite < 0), per-frame energy and virial sums over the frames' atoms.
template <class T>
void finish_efs(Workspace& ws, const Batch& b, const EfsOut& out, const SiteConst& sc, hipStream_t st,
                int nsite = -1);
// Zero n values of type T on the stream (fill kernel; no memset node).
template <class T>
void fill_zero(Workspace& ws, T* p, size_t n, hipStream_t st);

// Device view of one pair-style evaluation (owned atoms [0, nowned), ghosts
// [nowned, nall)).
struct PairView {
  int nowned = 0, nall = 0;
  const int32_t* types = nullptr;   // [nall]
  const int* in_off = nullptr;      // [nall + 1]
  const int64_t* frame_off = nullptr;  // {0, nowned}
  double *energy = nullptr, *force = nullptr, *virial = nullptr, *site = nullptr;  // device outputs
};
template <class Cf>
void pair_stage_shape(int stage, const DevModelSet& dm, Workspace& ws, const PairView& pv, hipStream_t st);
template <class T>
void build_pair_lists(Workspace& ws, int nowned, int nall, long nedges, const int32_t* ei, const int32_t* ej,
                      const double* d, const int32_t* types, int* in_off, hipStream_t st);
void pair_stage_dispatch(int stage, int precision, const DevModelSet& dm, Workspace& ws, const PairView& pv,
                         hipStream_t st);
template <class Cf>
void run_efs_shape(const DevModelSet& dm, Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st);

// Weights of the EFS combination whose parameter gradient is taken (device).
struct GradWeights {
  const double* ene_w;  // [N]
  const double* frc_w;  // [N*3]
  const double* str_w;  // [nf*9] row-major
  const int* atom_frame;  // [N] frame of every atom
};
template <class Cf>
void run_grad_shape(const DevModelSet& dm, Workspace& ws, const Batch& b, const GradWeights& gw, double* grad,
                    double scale, bool add, bool reuse_neighbors, hipStream_t st);

// Host entry points of one model shape (L, K, B, warp-sharing mode), index 0
// FP32 and 1 FP64.  The shapes of SUS2_FOR_EACH_SHAPE are compiled into the
// library; any other shape's are compiled at first use into a shape library
// that is loaded beside it (shape_ops).
struct ShapeOps {
  void (*efs[2])(const DevModelSet&, Workspace&, const Batch&, const EfsOut&, hipStream_t);
  void (*grad[2])(const DevModelSet&, Workspace&, const Batch&, const GradWeights&, double*, double, bool, bool,
                  hipStream_t);
  void (*pair[2])(int, const DevModelSet&, Workspace&, const PairView&, hipStream_t);
};
// Interface version between the library and its shape libraries.
constexpr int SHAPE_ABI = 1;

template <int L, int K, int B, int SM>
ShapeOps make_shape_ops() {
  using C32 = Cfg<float, L, K, B, SM>;
  using C64 = Cfg<double, L, K, B, SM>;
  ShapeOps o;
  o.efs[0] = &run_efs_shape<C32>;
  o.efs[1] = &run_efs_shape<C64>;
  o.grad[0] = &run_grad_shape<C32>;
  o.grad[1] = &run_grad_shape<C64>;
  o.pair[0] = &pair_stage_shape<C32>;
  o.pair[1] = &pair_stage_shape<C64>;
  return o;
}

// The entry points of a model's shape: prebuilt, from a shape library loaded
// earlier, or compiled now (hipcc, cached on disk); throws SUS2_E_UNSUPPORTED
// when the shape cannot be built.
const ShapeOps& shape_ops(int L, int K, int B, int SM);
inline const ShapeOps& shape_ops(const Model& m) { return shape_ops(m.L, m.K, m.B, m.smode); }
bool is_prebuilt_shape(int L, int K, int B, int SM);
// Cache path of a shape library, and its compilation when absent (no GPU use).
std::string shape_library_path(int L, int K, int B, int SM);
std::string build_shape_library(int L, int K, int B, int SM);
template <class T>
constexpr int prec_index() { return sizeof(T) == 8 ? 1 : 0; }

}  // namespace sus2
