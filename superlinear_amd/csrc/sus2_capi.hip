// extern "C" boundary of the SUS2-SH MI355X path (declared in include/sus2amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sus2amd.h"
#include "sus2_model.hpp"
#include "sus2_optim.hpp"
#include "sus2_common.hpp"
#include "sus2_runtime.hpp"

struct sus2_model {
  sus2::Model model;
  sus2::DevModelSet dev;
  sus2::Workspace ws;
  sus2::ObjectiveOpts obj;  // objective options of the loss, trainers and graphs
  bool uploaded = false;
  int device = -1;
};

namespace {
thread_local std::string g_error;
// Candidate launch shapes of the autotuner: centres per edge-backward group x
// grid multiple (0 = one workgroup per atom / group).
constexpr int kTuneEb[3] = {1, 2, 4};
constexpr int kTuneGrid[4] = {0, 1, 2, 4};
static_assert(SUS2_TUNE_CANDIDATES == 12, "candidate table");

int guard(const std::function<void()>& f) {
  try {
    f();
    return SUS2_OK;
  } catch (const sus2::Error& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_error = "out of host memory";
    return SUS2_E_DEVICE;
  } catch (const std::exception& e) {
    g_error = e.what();
    return SUS2_E_INVALID;
  }
}

void ensure_device(sus2_model* m) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "no HIP device available");
  if (!m->uploaded) {
    m->dev.upload(m->model);
    m->uploaded = true;
    m->device = dev;
  } else if (dev != m->device) {
    throw sus2::Error(SUS2_E_DEVICE, "model handle was set up on another device");
  }
}

void check_batch(const sus2_model* m, int nframes, int natoms, const void* frame_off, const void* types,
                 const void* pos, const void* cell, const void* periodic) {
  if (nframes < 0 || natoms < 0) throw sus2::Error(SUS2_E_INVALID, "negative frame or atom count");
  if (nframes == 0) return;
  if (!frame_off || !cell || !periodic || (natoms > 0 && (!types || !pos)))
    throw sus2::Error(SUS2_E_INVALID, "null batch buffer");
  (void)m;
}

sus2::Batch make_batch(int nframes, int natoms, const int64_t* frame_off, const int32_t* types, const double* pos,
                       const double* cell, const int32_t* periodic) {
  sus2::Batch b;
  b.nframes = nframes;
  b.natoms = natoms;
  b.frame_off = frame_off;
  b.types = types;
  b.pos = pos;
  b.cell = cell;
  b.periodic = periodic;
  return b;
}

// Host-side validation of species and frame offsets (the reference rejects
// out-of-range species in Kernel::begin_impl, sus2_sl_kernel.cpp:422).
void validate_host(const sus2::Model& m, int nframes, int natoms, const int64_t* off, const int32_t* types,
                   const double* pos, const double* cell, const int32_t* periodic) {
  if (off[0] != 0 || off[nframes] != natoms) throw sus2::Error(SUS2_E_INVALID, "frame offsets must span [0, natoms]");
  for (int f = 0; f < nframes; ++f)
    if (off[f + 1] < off[f]) throw sus2::Error(SUS2_E_INVALID, "frame offsets must be non-decreasing");
  for (int i = 0; i < natoms; ++i)
    if (types[i] < 0 || types[i] >= m.C) throw sus2::Error(SUS2_E_RANGE, "atom species index out of range");
  for (long i = 0; i < 3L * natoms; ++i)
    if (!std::isfinite(pos[i])) throw sus2::Error(SUS2_E_INVALID, "non-finite coordinate");
  // periodic lattice rows (the non-zero ones) must be finite and independent
  for (int f = 0; f < nframes; ++f) {
    if (!periodic[f]) continue;
    const double* a = cell + 9 * f;
    double basis[3][3];
    int nb = 0;
    for (int q = 0; q < 3; ++q) {
      double v[3] = {a[3 * q], a[3 * q + 1], a[3 * q + 2]};
      if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2]))
        throw sus2::Error(SUS2_E_INVALID, "non-finite lattice vector");
      if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0) continue;
      const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int p = 0; p < nb; ++p) {
        const double d = v[0] * basis[p][0] + v[1] * basis[p][1] + v[2] * basis[p][2];
        for (int k = 0; k < 3; ++k) v[k] -= d * basis[p][k];
      }
      const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (!(r > 1e-9 * len)) throw sus2::Error(SUS2_E_INVALID, "singular periodic cell (dependent lattice rows)");
      for (int k = 0; k < 3; ++k) basis[nb][k] = v[k] / r;
      ++nb;
    }
  }
}

struct Dev {
  void* p = nullptr;
  explicit Dev(size_t n) {
    if (hipMalloc(&p, std::max<size_t>(n, 8)) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "hipMalloc failed");
  }
  ~Dev() { (void)hipFree(p); }
  template <class U> U* get() { return static_cast<U*>(p); }
};
void h2d(void* d, const void* h, size_t n) {
  if (n && hipMemcpy(d, h, n, hipMemcpyHostToDevice) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "hipMemcpy H2D failed");
}
void d2h(void* h, const void* d, size_t n) {
  if (n && hipMemcpy(h, d, n, hipMemcpyDeviceToHost) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "hipMemcpy D2H failed");
}
}  // namespace

extern "C" {

int sus2_abi_version(void) { return SUS2_ABI_VERSION; }
const char* sus2_last_error(void) { return g_error.c_str(); }

int sus2_model_load(const char* path, sus2_model** out) {
  return guard([&] {
    if (!path || !out) throw sus2::Error(SUS2_E_INVALID, "null argument");
    std::unique_ptr<sus2_model> m(new sus2_model());
    m->model.load(path);
    *out = m.release();
  });
}

int sus2_model_save(const sus2_model* m, const char* path) {
  return guard([&] {
    if (!m || !path) throw sus2::Error(SUS2_E_INVALID, "null argument");
    m->model.save(path);
  });
}

void sus2_model_free(sus2_model* m) { delete m; }

int sus2_model_dims(const sus2_model* m, int* d) {
  return guard([&] {
    if (!m || !d) throw sus2::Error(SUS2_E_INVALID, "null argument");
    const auto& x = m->model;
    const int v[16] = {x.C, x.L, x.K, x.B, x.F, x.M, x.Q, x.R, x.O, x.Qg, x.Kw, (int)x.bk.size(), (int)x.prod.size(),
                       x.gated ? 1 : 0, x.species_readout ? 1 : 0, x.H};
    std::memcpy(d, v, sizeof(v));
  });
}

double sus2_model_cutoff(const sus2_model* m) { return m ? m->model.rc : 0.0; }

int sus2_shape_build(int L, int K, int B, int scaling_mode, char* path, int cap) {
  return guard([&] {
    if (!path || cap <= 0) throw sus2::Error(SUS2_E_INVALID, "null argument");
    const sus2::ShapeOps& ops = sus2::shape_ops(L, K, B, scaling_mode);  // builds and loads when not prebuilt
    (void)ops;
    std::string p = sus2::is_prebuilt_shape(L, K, B, scaling_mode) ? std::string()
                                                                     : sus2::shape_library_path(L, K, B, scaling_mode);
    if ((int)p.size() + 1 > cap) throw sus2::Error(SUS2_E_RANGE, "path buffer too small");
    std::memcpy(path, p.c_str(), p.size() + 1);
  });
}
int sus2_coeff_count(const sus2_model* m) { return m ? m->model.ncoeff : -1; }

int sus2_coeff_get(const sus2_model* m, double* theta) {
  return guard([&] {
    if (!m || !theta) throw sus2::Error(SUS2_E_INVALID, "null argument");
    std::memcpy(theta, m->model.theta.data(), sizeof(double) * m->model.theta.size());
  });
}

int sus2_coeff_set(sus2_model* m, const double* theta) {
  return guard([&] {
    if (!m || !theta) throw sus2::Error(SUS2_E_INVALID, "null argument");
    for (int i = 0; i < m->model.ncoeff; ++i)
      if (!std::isfinite(theta[i])) throw sus2::Error(SUS2_E_INVALID, "non-finite coefficient");
    std::memcpy(m->model.theta.data(), theta, sizeof(double) * m->model.theta.size());
    if (m->uploaded) m->dev.upload_theta(m->model);
  });
}

int sus2_coeff_offsets(const sus2_model* m, int* o) {
  return guard([&] {
    if (!m || !o) throw sus2::Error(SUS2_E_INVALID, "null argument");
    const auto& x = m->model;
    const int v[13] = {x.o_shift, x.o_scal, x.o_rad, x.o_grad, x.o_gscal, x.o_alpha, x.o_gtype,
                       x.o_W,     x.o_gw,   x.o_ow,  x.o_species, x.o_read, x.ncoeff};
    std::memcpy(o, v, sizeof(v));
  });
}

int sus2_efs(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off, const int32_t* types,
             const double* pos, const double* cell, const int32_t* periodic, double* energy, double* force,
             double* virial, double* site, void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) return;
    if (!energy || !virial || (natoms > 0 && !force)) throw sus2::Error(SUS2_E_INVALID, "null output buffer");
    ensure_device(m);
    sus2::EfsOut out{energy, force, virial, site};
    sus2::efs_dispatch(m->dev, m->ws, precision, make_batch(nframes, natoms, frame_off, types, pos, cell, periodic),
                       out, (hipStream_t)stream);
  });
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free EFS
// Device scratch of sus2_efs_noalloc: one static array per GPU (each device
// holds its own copy once this module is loaded there).  Host code never
// passes the array's name to a kernel: the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[SUS2_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Address of this device's scratch array (set up on the device's first call).
void* scratch_of(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) throw sus2::Error(SUS2_E_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw sus2::Error(SUS2_E_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int sus2_efs_noalloc(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                     const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                     double* energy, double* force, double* virial, double* site, void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) return;
    if (!energy || !virial || (natoms > 0 && !force)) throw sus2::Error(SUS2_E_INVALID, "null output buffer");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "no HIP device available");
    if (!m->uploaded || dev != m->device)
      throw sus2::Error(SUS2_E_INVALID, "sus2_efs_noalloc needs the model's device copy on the current device: "
                                        "make one sus2_efs call (which may allocate) on this device first");
    sus2::Arena arena;
    arena.base = static_cast<char*>(scratch_of(dev));
    arena.cap = SUS2_NOALLOC_SCRATCH_BYTES;
    sus2::ArenaScope scope(&arena);
    sus2::Workspace ws;  // every buffer carved from the scratch array; nothing outlives the call
    sus2::EfsOut out{energy, force, virial, site};
    sus2::efs_dispatch(m->dev, ws, precision, make_batch(nframes, natoms, frame_off, types, pos, cell, periodic),
                       out, (hipStream_t)stream);
  });
}

int sus2_efs_host(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                  const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                  double* energy, double* force, double* virial, double* site) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) return;
    validate_host(m->model, nframes, natoms, frame_off, types, pos, cell, periodic);
    ensure_device(m);
    Dev doff(sizeof(int64_t) * (nframes + 1)), dty(sizeof(int32_t) * natoms), dpos(sizeof(double) * 3 * natoms),
        dcell(sizeof(double) * 9 * nframes), dper(sizeof(int32_t) * nframes), de(sizeof(double) * nframes),
        df(sizeof(double) * 3 * natoms), dv(sizeof(double) * 6 * nframes), ds(sizeof(double) * natoms);
    h2d(doff.p, frame_off, sizeof(int64_t) * (nframes + 1));
    h2d(dty.p, types, sizeof(int32_t) * natoms);
    h2d(dpos.p, pos, sizeof(double) * 3 * natoms);
    h2d(dcell.p, cell, sizeof(double) * 9 * nframes);
    h2d(dper.p, periodic, sizeof(int32_t) * nframes);
    sus2::EfsOut out{de.get<double>(), df.get<double>(), dv.get<double>(), ds.get<double>()};
    sus2::efs_dispatch(m->dev, m->ws, precision,
                       make_batch(nframes, natoms, doff.get<int64_t>(), dty.get<int32_t>(), dpos.get<double>(),
                                  dcell.get<double>(), dper.get<int32_t>()),
                       out, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "kernel execution failed");
    d2h(energy, de.p, sizeof(double) * nframes);
    if (force) d2h(force, df.p, sizeof(double) * 3 * natoms);
    if (virial) d2h(virial, dv.p, sizeof(double) * 6 * nframes);
    if (site) d2h(site, ds.p, sizeof(double) * natoms);
  });
}

int sus2_efs_grad(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                  const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                  const double* ene_weight, const double* frc_weight, const double* str_weight, double* grad,
                  void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) return;
    if (!ene_weight || !frc_weight || !str_weight || !grad) throw sus2::Error(SUS2_E_INVALID, "null weight buffer");
    ensure_device(m);
    sus2::grad_dispatch(m->dev, m->ws, precision,
                        make_batch(nframes, natoms, frame_off, types, pos, cell, periodic), ene_weight, frc_weight,
                        str_weight, grad, 1.0, true, false, (hipStream_t)stream);
  });
}

int sus2_loss_grad(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                   const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                   const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                   double wS, double* loss_out, double* grad, void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (!loss_out || !grad) throw sus2::Error(SUS2_E_INVALID, "null loss or gradient buffer");
    if (nframes > 0 && (!e0 || !s0 || (natoms > 0 && !f0))) throw sus2::Error(SUS2_E_INVALID, "null label buffer");
    ensure_device(m);
    sus2::loss_grad_dispatch(m->dev, m->ws, precision,
                             make_batch(nframes, natoms, frame_off, types, pos, cell, periodic), e0, f0, s0,
                             weighting, wE, wF, wS, m->obj, loss_out, grad, (hipStream_t)stream);
  });
}

// ------------------------------------------------------------------ training
// L-BFGS over the training objective: host FP64 two-loop recursion with
// history `history`, Armijo backtracking (c1 = 1e-4, halving), curvature
// guard s.y > 1e-10 |s||y| (updates failing it are skipped), restart on a
// non-descent direction; every objective/gradient evaluation runs on the GPU
// (sus2_loss_grad).  Follows the protocol of LBFGS::IterateArmijoWithGradient
// (src/common/lbfgs.h:14-20, lbfgs.cpp:488-542) as used by the trainer's
// evaluation loop (dev_src/mtpr_trainer.cpp:4685-4760).
int sus2_train_lbfgs(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                     const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                     const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                     double wS, int max_iter, int history, double gtol, double* loss_trace, int* iters_out,
                     void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (max_iter < 0 || history < 1) throw sus2::Error(SUS2_E_INVALID, "max_iter >= 0 and history >= 1 required");
    ensure_device(m);
    const hipStream_t st = (hipStream_t)stream;
    const sus2::Batch b = make_batch(nframes, natoms, frame_off, types, pos, cell, periodic);
    const int P = m->model.ncoeff;
    Dev gdev(sizeof(double) * P);
    std::vector<double> x = m->model.theta;
    const sus2::Objective obj = [&](const std::vector<double>& th, std::vector<double>& grad) {
      std::memcpy(m->model.theta.data(), th.data(), sizeof(double) * P);
      m->dev.upload_theta(m->model);
      double f = 0;
      sus2::loss_grad_dispatch(m->dev, m->ws, precision, b, e0, f0, s0, weighting, wE, wF, wS, m->obj, &f, gdev.get<double>(),
                               st);
      // the gradient kernels finish on the stream after the objective is known
      if (hipStreamSynchronize(st) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "stream synchronize failed");
      d2h(grad.data(), gdev.get<double>(), sizeof(double) * P);
      return f;
    };
    const sus2::LbfgsResult r = sus2::lbfgs_minimize(obj, x, max_iter, history, gtol);
    const int it = r.iters;
    if (loss_trace)
      for (size_t k = 0; k < r.trace.size(); ++k) loss_trace[k] = r.trace[k];
    // leave the model at the final accepted point
    std::memcpy(m->model.theta.data(), x.data(), sizeof(double) * P);
    m->dev.upload_theta(m->model);
    if (iters_out) *iters_out = it;
  });
}

// ------------------------------------------------------------------ HIP graphs
}  // extern "C"

struct sus2_graph {
  sus2_model* model = nullptr;
  sus2::Workspace ws;  // the graph's own buffers, sized before the capture
  sus2::DevBuf flag;   // sticky neighbour-overflow flag (int)
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~sus2_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {
// Three runs of `run` on the graph's private stream: host-sized (eager), then
// capture-safe (eager: sizes every buffer of that path and checks that the
// batch fits the edge capacity), then the same calls captured into a graph.
template <class Fn>
void build_graph(sus2_graph* g, int nframes, const int64_t* frame_off, hipStream_t user, Fn&& run) {
  SUS2_CHECK(hipStreamSynchronize(user));  // inputs written on the caller's stream are complete
  SUS2_CHECK(hipStreamCreate(&g->cap));
  std::vector<int64_t> off(nframes + 1);
  SUS2_CHECK(hipMemcpy(off.data(), frame_off, sizeof(int64_t) * (nframes + 1), hipMemcpyDeviceToHost));
  long mx = 0;
  for (int f = 0; f < nframes; ++f) mx = std::max<long>(mx, off[f + 1] - off[f]);
  g->ws.known_max_frame = mx;
  g->flag.ensure(sizeof(int));
  SUS2_CHECK(hipMemset(g->flag.p, 0, sizeof(int)));
  g->ws.overflow_flag = g->flag.get<int>();
  run(g->cap);
  SUS2_CHECK(hipStreamSynchronize(g->cap));
  g->ws.capture = true;
  run(g->cap);
  SUS2_CHECK(hipStreamSynchronize(g->cap));
  int over = 0;
  SUS2_CHECK(hipMemcpy(&over, g->flag.p, sizeof(int), hipMemcpyDeviceToHost));
  if (over) throw sus2::Error(SUS2_E_RANGE, "neighbour lists outgrew the capacity sized for the graph");
  {
    sus2::CaptureScope scope;
    SUS2_CHECK(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
    try {
      run(g->cap);
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(g->cap, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      throw;
    }
    SUS2_CHECK(hipStreamEndCapture(g->cap, &g->graph));
  }
  SUS2_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
}
}  // namespace

extern "C" {

int sus2_graph_efs_create(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                          const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                          double* energy, double* force, double* virial, double* site, void* stream,
                          sus2_graph** out) {
  return guard([&] {
    if (!m || !out) throw sus2::Error(SUS2_E_INVALID, "null argument");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) throw sus2::Error(SUS2_E_INVALID, "empty batch");
    if (!energy || !virial || (natoms > 0 && !force)) throw sus2::Error(SUS2_E_INVALID, "null output buffer");
    ensure_device(m);
    std::unique_ptr<sus2_graph> g(new sus2_graph());
    g->model = m;
    const sus2::Batch b = make_batch(nframes, natoms, frame_off, types, pos, cell, periodic);
    const sus2::EfsOut o{energy, force, virial, site};
    build_graph(g.get(), nframes, frame_off, (hipStream_t)stream,
                [&](hipStream_t st) { sus2::efs_dispatch(m->dev, g->ws, precision, b, o, st); });
    *out = g.release();
  });
}

int sus2_graph_loss_grad_create(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                                const int32_t* types, const double* pos, const double* cell,
                                const int32_t* periodic, const double* e0, const double* f0, const double* s0,
                                int weighting, double wE, double wF, double wS, double* loss, double* grad,
                                double* energy, double* force, double* virial, void* stream, sus2_graph** out) {
  return guard([&] {
    if (!m || !out || !loss || !grad) throw sus2::Error(SUS2_E_INVALID, "null argument");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) throw sus2::Error(SUS2_E_INVALID, "empty batch");
    if (!e0 || !s0 || (natoms > 0 && !f0)) throw sus2::Error(SUS2_E_INVALID, "null label buffer");
    ensure_device(m);
    std::unique_ptr<sus2_graph> g(new sus2_graph());
    g->model = m;
    const sus2::Batch b = make_batch(nframes, natoms, frame_off, types, pos, cell, periodic);
    const sus2::EfsOut o{energy, force, virial, nullptr};
    build_graph(g.get(), nframes, frame_off, (hipStream_t)stream, [&](hipStream_t st) {
      sus2::loss_grad_dispatch(m->dev, g->ws, precision, b, e0, f0, s0, weighting, wE, wF, wS, m->obj, nullptr, grad, st,
                               loss, &o);
    });
    *out = g.release();
  });
}

int sus2_graph_launch(sus2_graph* g, void* stream) {
  return guard([&] {
    if (!g || !g->exec) throw sus2::Error(SUS2_E_INVALID, "null graph");
    SUS2_CHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
  });
}

int sus2_graph_overflow(sus2_graph* g, void* stream, int* overflow) {
  return guard([&] {
    if (!g || !overflow) throw sus2::Error(SUS2_E_INVALID, "null argument");
    SUS2_CHECK(hipStreamSynchronize((hipStream_t)stream));
    int v = 0;
    SUS2_CHECK(hipMemcpy(&v, g->flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (v) SUS2_CHECK(hipMemset(g->flag.p, 0, sizeof(int)));
    *overflow = v;
  });
}

void sus2_graph_free(sus2_graph* g) { delete g; }

// ------------------------------------------------------------------ launch shape and autotuner
namespace {
sus2::EfsTune& tune_slot(sus2_model* m, int precision) {
  if (precision != SUS2_FP64 && precision != SUS2_FP32) throw sus2::Error(SUS2_E_INVALID, "precision must be FP64 or FP32");
  return m->dev.tune[precision == SUS2_FP32 ? 1 : 0];
}
}  // namespace

int sus2_tune_set(sus2_model* m, int precision, int eb_atoms, int grid_mult) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    if (eb_atoms != 1 && eb_atoms != 2 && eb_atoms != 4) throw sus2::Error(SUS2_E_INVALID, "eb_atoms must be 1, 2 or 4");
    if (grid_mult < 0 || grid_mult > 64) throw sus2::Error(SUS2_E_INVALID, "grid_mult must be in [0, 64]");
    sus2::EfsTune& t = tune_slot(m, precision);
    t.eb_atoms = eb_atoms;
    t.grid_mult = grid_mult;
  });
}

int sus2_tune_messages(sus2_model* m, int precision, int mode, int* current) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    if (mode < -1 || mode > 3)
      throw sus2::Error(SUS2_E_INVALID,
                        "mode must be -1 (query), 0 (combined), 1 (recompute), 2 (split) or 3 (automatic)");
    sus2::EfsTune& t = tune_slot(m, precision);
    if (mode >= 0) t.msg_recompute = mode == 3 ? -1 : mode;
    if (current) *current = t.msg_recompute < 0 ? 3 : t.msg_recompute;
  });
}

int sus2_tune_get(sus2_model* m, int precision, int* eb_atoms, int* grid_mult) {
  return guard([&] {
    if (!m || !eb_atoms || !grid_mult) throw sus2::Error(SUS2_E_INVALID, "null argument");
    const sus2::EfsTune& t = tune_slot(m, precision);
    *eb_atoms = t.eb_atoms;
    *grid_mult = t.grid_mult;
  });
}

// Times the batch EFS under every launch shape of the candidate table
// (centres per edge-backward group x grid multiple), `reps` calls each after
// one warm-up call, with HIP events on `stream`, and keeps the fastest
// (median call time).  ms_out [SUS2_TUNE_CANDIDATES] receives each shape's
// median (may be NULL); best [2] the chosen eb_atoms, grid_mult.
int sus2_autotune(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                  const int32_t* types, const double* pos, const double* cell, const int32_t* periodic, int reps,
                  void* stream, int* best, double* ms_out) {
  return guard([&] {
    if (!m || !best) throw sus2::Error(SUS2_E_INVALID, "null argument");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (reps < 1) throw sus2::Error(SUS2_E_INVALID, "reps >= 1 required");
    ensure_device(m);
    sus2::EfsTune& slot = tune_slot(m, precision);
    const sus2::EfsTune keep = slot;
    const hipStream_t st = (hipStream_t)stream;
    Dev de(sizeof(double) * std::max(1, nframes)), df(sizeof(double) * 3 * std::max(1, natoms)),
        dv(sizeof(double) * 6 * std::max(1, nframes));
    const sus2::Batch b = make_batch(nframes, natoms, frame_off, types, pos, cell, periodic);
    const sus2::EfsOut out{de.get<double>(), df.get<double>(), dv.get<double>(), nullptr};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      throw sus2::Error(SUS2_E_DEVICE, "hipEventCreate failed");
    double best_ms = 1e300;
    sus2::EfsTune chosen = keep;
    try {
      for (int c = 0; c < SUS2_TUNE_CANDIDATES; ++c) {
        sus2::EfsTune t = keep;
        t.eb_atoms = kTuneEb[c / 4];
        t.grid_mult = kTuneGrid[c % 4];
        slot = t;
        sus2::efs_dispatch(m->dev, m->ws, precision, b, out, st);  // warm-up (sizes, compiles nothing)
        std::vector<double> ms(reps);
        for (int r = 0; r < reps; ++r) {
          SUS2_CHECK(hipEventRecord(e0, st));
          sus2::efs_dispatch(m->dev, m->ws, precision, b, out, st);
          SUS2_CHECK(hipEventRecord(e1, st));
          SUS2_CHECK(hipEventSynchronize(e1));
          float v = 0;
          SUS2_CHECK(hipEventElapsedTime(&v, e0, e1));
          ms[r] = v;
        }
        std::sort(ms.begin(), ms.end());
        const double med = ms[reps / 2];
        if (ms_out) ms_out[c] = med;
        if (med < best_ms) {
          best_ms = med;
          chosen = t;
        }
      }
    } catch (...) {
      slot = keep;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    slot = chosen;
    best[0] = chosen.eb_atoms;
    best[1] = chosen.grid_mult;
  });
}

// BFGS over the training objective: host FP64 dense inverse Hessian and
// strong-Wolfe line search (sus2_optim.hpp; protocol of BFGS::Iterate,
// src/common/bfgs.cpp:251-330, as driven by MTPR_trainer::Train,
// dev_src/mtpr_trainer.cpp:4685-4760); every objective and gradient runs on
// the GPU (sus2_loss_grad).
int sus2_train_bfgs(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                    const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                    const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                    double wS, int max_iter, double gtol, double* loss_trace, int* iters_out, void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (max_iter < 0) throw sus2::Error(SUS2_E_INVALID, "max_iter >= 0 required");
    ensure_device(m);
    const hipStream_t st = (hipStream_t)stream;
    const sus2::Batch b = make_batch(nframes, natoms, frame_off, types, pos, cell, periodic);
    const int P = m->model.ncoeff;
    Dev gdev(sizeof(double) * P);
    sus2::Objective fn = [&](const std::vector<double>& th, std::vector<double>& grad) {
      std::memcpy(m->model.theta.data(), th.data(), sizeof(double) * P);
      m->dev.upload_theta(m->model);
      double f = 0;
      sus2::loss_grad_dispatch(m->dev, m->ws, precision, b, e0, f0, s0, weighting, wE, wF, wS, m->obj, &f, gdev.get<double>(),
                               st);
      if (hipStreamSynchronize(st) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "stream synchronize failed");
      d2h(grad.data(), gdev.get<double>(), sizeof(double) * P);
      return f;
    };
    std::vector<double> x = m->model.theta;
    const sus2::BfgsResult r = sus2::bfgs_minimize(fn, x, max_iter, gtol);
    std::memcpy(m->model.theta.data(), x.data(), sizeof(double) * P);  // the last accepted point
    m->dev.upload_theta(m->model);
    if (loss_trace)
      for (size_t k = 0; k < r.trace.size() && k <= (size_t)max_iter; ++k) loss_trace[k] = r.trace[k];
    if (iters_out) *iters_out = (int)r.trace.size() - 1;
  });
}

// Linear readout solve (TrainLinear, dev_src/mtpr_trainer.cpp:1130-1455).
int sus2_train_linear(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                      const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                      const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                      double wS, double ridge, double* objective_out, void* stream) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    check_batch(m, nframes, natoms, frame_off, types, pos, cell, periodic);
    if (nframes == 0) throw sus2::Error(SUS2_E_INVALID, "empty batch");
    if (!e0 || !s0 || (natoms > 0 && !f0)) throw sus2::Error(SUS2_E_INVALID, "null label buffer");
    if (precision != SUS2_FP64 && precision != SUS2_FP32) throw sus2::Error(SUS2_E_INVALID, "precision must be FP64 or FP32");
    ensure_device(m);
    const double v = sus2::train_linear(m->model, m->dev, m->ws, precision,
                                        make_batch(nframes, natoms, frame_off, types, pos, cell, periodic), e0, f0,
                                        s0, weighting, wE, wF, wS, ridge, m->obj, (hipStream_t)stream);
    if (objective_out) *objective_out = v;
  });
}

// ------------------------------------------------------------------ objective options
void sus2_objective_default(sus2_objective* out) {
  if (!out) return;
  const sus2::ObjectiveOpts d;
  *out = sus2_objective{d.rel_forces, d.norm_by_forces, d.force_loss, d.log_cosh_scale,
                        d.std_weight, d.stdd_weight, d.constr_weight};
}

int sus2_objective_get(const sus2_model* m, sus2_objective* out) {
  return guard([&] {
    if (!m || !out) throw sus2::Error(SUS2_E_INVALID, "null model or options");
    const sus2::ObjectiveOpts& o = m->obj;
    *out = sus2_objective{o.rel_forces, o.norm_by_forces, o.force_loss, o.log_cosh_scale,
                          o.std_weight, o.stdd_weight, o.constr_weight};
  });
}

int sus2_objective_set(sus2_model* m, const sus2_objective* in) {
  return guard([&] {
    if (!m || !in) throw sus2::Error(SUS2_E_INVALID, "null model or options");
    sus2::ObjectiveOpts o;
    o.rel_forces = in->rel_forces;
    o.norm_by_forces = in->norm_by_forces;
    o.force_loss = in->force_loss;
    o.log_cosh_scale = in->log_cosh_scale;
    o.std_weight = in->std_weight;
    o.stdd_weight = in->stdd_weight;
    o.constr_weight = in->constr_weight;
    o.validate();
    if (o.std_terms() && m->model.C > 64)
      throw sus2::Error(SUS2_E_UNSUPPORTED, "the std terms support at most 64 species");
    m->obj = o;
  });
}

// ------------------------------------------------------------------ pair style
}  // extern "C"

struct sus2_pair {
  sus2_model* model = nullptr;
  int precision = SUS2_FP64;
  sus2::Workspace ws;
  sus2::DevBuf types, ei, ej, d, in_off, frame_off, outE, outF, outV, outS, rows, idx;
  int nowned = 0, nall = 0;
  int state = 0;  // 0 idle, 1 begun, 2 outer done
  int width = 0;
};

namespace {
// Rows of `src` (T, width F) for the listed atoms -> FP64 host buffer, or the
// reverse (set or add).  mode 0 pack, 1 set, 2 add.
template <class T>
__global__ void k_rows(int n, int F, const int* __restrict__ atoms, T* __restrict__ dev, double* __restrict__ buf,
                       int mode) {
  const int k = blockIdx.x * blockDim.y + threadIdx.y, mu = threadIdx.x;
  if (k >= n || mu >= F) return;
  T* row = dev + (size_t)atoms[k] * F;
  if (mode == 0) buf[(size_t)k * F + mu] = (double)row[mu];
  else if (mode == 1) row[mu] = (T)buf[(size_t)k * F + mu];
  else row[mu] = (T)((double)row[mu] + buf[(size_t)k * F + mu]);
}

void pair_rows(sus2_pair* p, bool adj, int n, const int32_t* atoms, double* hbuf, int mode) {
  if (n < 0) throw sus2::Error(SUS2_E_INVALID, "negative atom count");
  if (p->width == 0 || n == 0) return;
  if (!atoms || !hbuf) throw sus2::Error(SUS2_E_INVALID, "null atom list or buffer");
  const int F = p->width;
  std::vector<int32_t> list(atoms, atoms + n);
  std::vector<double> vals;
  if (mode == 2) {  // aggregate repeated atoms on the host in list order, then one parallel add
    std::vector<int32_t> uniq;
    std::vector<int> slot(p->nall, -1);
    for (int k = 0; k < n; ++k) {
      const int a = list[k];
      if (a < 0 || a >= p->nall) throw sus2::Error(SUS2_E_RANGE, "atom index out of range");
      if (slot[a] < 0) {
        slot[a] = (int)uniq.size();
        uniq.push_back(a);
        vals.insert(vals.end(), hbuf + (size_t)k * F, hbuf + (size_t)(k + 1) * F);
      } else {
        for (int mu = 0; mu < F; ++mu) vals[(size_t)slot[a] * F + mu] += hbuf[(size_t)k * F + mu];
      }
    }
    list.swap(uniq);
    n = (int)list.size();
  } else {
    for (int k = 0; k < n; ++k)
      if (list[k] < 0 || list[k] >= p->nall) throw sus2::Error(SUS2_E_RANGE, "atom index out of range");
    if (mode == 1) vals.assign(hbuf, hbuf + (size_t)n * F);
  }
  p->idx.ensure(sizeof(int) * n);
  p->rows.ensure(sizeof(double) * (size_t)n * F);
  h2d(p->idx.get<int>(), list.data(), sizeof(int) * n);
  if (mode != 0) h2d(p->rows.get<double>(), vals.data(), sizeof(double) * (size_t)n * F);
  const int per = std::max(1, 256 / F);
  const dim3 block(F, per), grid((n + per - 1) / per);
  if (p->precision == SUS2_FP32) {
    float* dev = adj ? p->ws.hbar.get<float>() : p->ws.h.get<float>();
    hipLaunchKernelGGL(k_rows<float>, grid, block, 0, nullptr, n, F, p->idx.get<int>(), dev, p->rows.get<double>(),
                       mode);
  } else {
    double* dev = adj ? p->ws.hbar.get<double>() : p->ws.h.get<double>();
    hipLaunchKernelGGL(k_rows<double>, grid, block, 0, nullptr, n, F, p->idx.get<int>(), dev, p->rows.get<double>(),
                       mode);
  }
  if (hipGetLastError() != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "row copy kernel failed");
  if (mode == 0) d2h(hbuf, p->rows.get<double>(), sizeof(double) * (size_t)n * F);
  else if (hipStreamSynchronize(nullptr) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "stream synchronize failed");
}

sus2::PairView pair_view(sus2_pair* p) {
  sus2::PairView v;
  v.nowned = p->nowned;
  v.nall = p->nall;
  v.types = p->types.get<int32_t>();
  v.in_off = p->in_off.get<int>();
  v.frame_off = p->frame_off.get<int64_t>();
  v.energy = p->outE.get<double>();
  v.force = p->outF.get<double>();
  v.virial = p->outV.get<double>();
  v.site = p->outS.get<double>();
  return v;
}
}  // namespace

extern "C" {

int sus2_pair_create(sus2_model* m, int precision, sus2_pair** out) {
  return guard([&] {
    if (!m || !out) throw sus2::Error(SUS2_E_INVALID, "null argument");
    if (precision != SUS2_FP64 && precision != SUS2_FP32) throw sus2::Error(SUS2_E_INVALID, "bad precision");
    ensure_device(m);
    std::unique_ptr<sus2_pair> p(new sus2_pair());
    p->model = m;
    p->precision = precision;
    p->width = m->model.gated ? m->model.F : 0;
    *out = p.release();
  });
}

void sus2_pair_free(sus2_pair* p) { delete p; }

int sus2_pair_comm_width(const sus2_pair* p) { return p ? p->width : 0; }

int sus2_pair_begin(sus2_pair* p, int nowned, int nall, const int32_t* types, int64_t nedges, const int32_t* ei,
                    const int32_t* ej, const double* d) {
  return guard([&] {
    if (!p) throw sus2::Error(SUS2_E_INVALID, "null pair handle");
    const sus2::Model& M = p->model->model;
    if (nowned < 0 || nall < nowned || nedges < 0) throw sus2::Error(SUS2_E_INVALID, "invalid atom layout");
    if ((nall > 0 && !types) || (nedges > 0 && (!ei || !ej || !d))) throw sus2::Error(SUS2_E_INVALID, "null buffer");
    if (nedges > INT32_MAX - 1) throw sus2::Error(SUS2_E_RANGE, "too many edges");
    sus2::check_rec_limits(p->precision, nall, M.C);
    for (int k = 0; k < nall; ++k)
      if (types[k] < 0 || types[k] >= M.C) throw sus2::Error(SUS2_E_RANGE, "atom species index out of range");
    // validate every edge; edges at or beyond the cutoff contribute nothing and are dropped
    // (Kernel::begin_impl and the r >= max_dist skips, sus2_sl_kernel.cpp:435, 173-176)
    std::vector<int32_t> vi, vj;
    std::vector<double> vd;
    vi.reserve(nedges); vj.reserve(nedges); vd.reserve(3 * nedges);
    const double rc2 = M.rc * M.rc;
    for (int64_t k = 0; k < nedges; ++k) {
      const double r2 = d[3 * k] * d[3 * k] + d[3 * k + 1] * d[3 * k + 1] + d[3 * k + 2] * d[3 * k + 2];
      if (ei[k] < 0 || ei[k] >= nowned || ej[k] < 0 || ej[k] >= nall || !(r2 > 0) || !std::isfinite(r2))
        throw sus2::Error(SUS2_E_INVALID, "distributed EFS neighborhood has an invalid edge");
      if (r2 >= rc2) continue;
      vi.push_back(ei[k]); vj.push_back(ej[k]);
      vd.insert(vd.end(), d + 3 * k, d + 3 * k + 3);
    }
    const long ne = (long)vi.size();
    ensure_device(p->model);
    p->nowned = nowned;
    p->nall = nall;
    const size_t a1 = std::max(1, nall);
    p->types.ensure(sizeof(int32_t) * a1);
    p->ei.ensure(sizeof(int32_t) * std::max<long>(1, ne));
    p->ej.ensure(sizeof(int32_t) * std::max<long>(1, ne));
    p->d.ensure(sizeof(double) * 3 * std::max<long>(1, ne));
    p->in_off.ensure(sizeof(int) * (a1 + 1));
    p->frame_off.ensure(sizeof(int64_t) * 2);
    p->outE.ensure(sizeof(double));
    p->outF.ensure(sizeof(double) * 3 * a1);
    p->outV.ensure(sizeof(double) * 6);
    p->outS.ensure(sizeof(double) * a1);
    h2d(p->types.get<int32_t>(), types, sizeof(int32_t) * nall);
    h2d(p->ei.get<int32_t>(), vi.data(), sizeof(int32_t) * ne);
    h2d(p->ej.get<int32_t>(), vj.data(), sizeof(int32_t) * ne);
    h2d(p->d.get<double>(), vd.data(), sizeof(double) * 3 * ne);
    const int64_t fo[2] = {0, nowned};
    h2d(p->frame_off.get<int64_t>(), fo, sizeof(fo));
    if (p->precision == SUS2_FP32)
      sus2::build_pair_lists<float>(p->ws, nowned, nall, ne, p->ei.get<int32_t>(), p->ej.get<int32_t>(),
                                    p->d.get<double>(), p->types.get<int32_t>(), p->in_off.get<int>(), nullptr);
    else
      sus2::build_pair_lists<double>(p->ws, nowned, nall, ne, p->ei.get<int32_t>(), p->ej.get<int32_t>(),
                                     p->d.get<double>(), p->types.get<int32_t>(), p->in_off.get<int>(), nullptr);
    sus2::pair_stage_dispatch(0, p->precision, p->model->dev, p->ws, pair_view(p), nullptr);
    p->state = 1;
  });
}

int sus2_pair_pack_forward(sus2_pair* p, int n, const int32_t* atoms, double* buf) {
  return guard([&] {
    if (!p || p->state < 1) throw sus2::Error(SUS2_E_INVALID, "pack_forward before begin");
    pair_rows(p, false, n, atoms, buf, 0);
  });
}

int sus2_pair_unpack_forward(sus2_pair* p, int n, const int32_t* atoms, const double* buf) {
  return guard([&] {
    if (!p || p->state != 1) throw sus2::Error(SUS2_E_INVALID, "unpack_forward outside begin .. outer");
    pair_rows(p, false, n, atoms, const_cast<double*>(buf), 1);
  });
}

int sus2_pair_outer(sus2_pair* p) {
  return guard([&] {
    if (!p || p->state != 1) throw sus2::Error(SUS2_E_INVALID, "outer before begin");
    sus2::pair_stage_dispatch(1, p->precision, p->model->dev, p->ws, pair_view(p), nullptr);
    p->state = 2;
  });
}

int sus2_pair_pack_reverse(sus2_pair* p, int n, const int32_t* atoms, double* buf) {
  return guard([&] {
    if (!p || p->state != 2) throw sus2::Error(SUS2_E_INVALID, "pack_reverse before outer");
    pair_rows(p, true, n, atoms, buf, 0);
  });
}

int sus2_pair_add_reverse(sus2_pair* p, int n, const int32_t* atoms, const double* buf) {
  return guard([&] {
    if (!p || p->state != 2) throw sus2::Error(SUS2_E_INVALID, "add_reverse before outer");
    pair_rows(p, true, n, atoms, const_cast<double*>(buf), 2);
  });
}

int sus2_pair_finish(sus2_pair* p, double* energy, double* site, double* force, double* virial) {
  return guard([&] {
    if (!p || p->state != 2) throw sus2::Error(SUS2_E_INVALID, "finish before outer");
    if (!energy || !force || !virial) throw sus2::Error(SUS2_E_INVALID, "null output buffer");
    sus2::pair_stage_dispatch(2, p->precision, p->model->dev, p->ws, pair_view(p), nullptr);
    d2h(energy, p->outE.get<double>(), sizeof(double));
    d2h(virial, p->outV.get<double>(), sizeof(double) * 6);
    d2h(force, p->outF.get<double>(), sizeof(double) * 3 * p->nall);
    if (site) d2h(site, p->outS.get<double>(), sizeof(double) * p->nowned);
    p->state = 0;
  });
}

int sus2_profile_enable(sus2_model* m, int on) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    m->ws.prof.on = on != 0;
  });
}

int sus2_profile_reset(sus2_model* m) {
  return guard([&] {
    if (!m) throw sus2::Error(SUS2_E_INVALID, "null model");
    m->ws.prof.reset();
  });
}

int sus2_profile_read(sus2_model* m, int cap, char* names, double* total_ms, int64_t* launches, int* n) {
  return guard([&] {
    if (!m || !n || cap < 0 || (cap > 0 && (!names || !total_ms || !launches)))
      throw sus2::Error(SUS2_E_INVALID, "null argument");
    auto& p = m->ws.prof;
    p.collect_all();
    *n = (int)p.names.size();
    for (int k = 0; k < cap && k < *n; ++k) {
      std::strncpy(names + 64 * k, p.names[k].c_str(), 63);
      names[64 * k + 63] = 0;
      total_ms[k] = p.total_ms[k];
      launches[k] = p.launches[k];
    }
  });
}

int sus2_last_edge_count(const sus2_model* m, int64_t* n) {
  return guard([&] {
    if (!m || !n) throw sus2::Error(SUS2_E_INVALID, "null argument");
    *n = m->ws.nedges;
  });
}

}  // extern "C"
