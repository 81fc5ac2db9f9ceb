// extern "C" boundary of the SUS2-SH MI355X path (declared in include/sus2amd.h).
#include <hip/hip_runtime.h>

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
#include "sus2_common.hpp"
#include "sus2_runtime.hpp"

struct sus2_model {
  sus2::Model model;
  sus2::DevModelSet dev;
  sus2::Workspace ws;
  bool uploaded = false;
  int device = -1;
};

namespace {
thread_local std::string g_error;

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
                             weighting, wE, wF, wS, loss_out, grad, (hipStream_t)stream);
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
    std::vector<double> x = m->model.theta, g(P), xn(P), gn(P), p(P), q(P);
    auto eval = [&](const std::vector<double>& th, std::vector<double>& grad) {
      std::memcpy(m->model.theta.data(), th.data(), sizeof(double) * P);
      m->dev.upload_theta(m->model);
      double f = 0;
      sus2::loss_grad_dispatch(m->dev, m->ws, precision, b, e0, f0, s0, weighting, wE, wF, wS, &f, gdev.get<double>(),
                               st);
      // the gradient kernels finish on the stream after the objective is known
      if (hipStreamSynchronize(st) != hipSuccess) throw sus2::Error(SUS2_E_DEVICE, "stream synchronize failed");
      d2h(grad.data(), gdev.get<double>(), sizeof(double) * P);
      return f;
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
      double s = 0;
      for (int k = 0; k < P; ++k) s += a[k] * c[k];
      return s;
    };
    std::vector<std::vector<double>> S, Y;
    std::vector<double> rho;
    double f = eval(x, g);
    if (loss_trace) loss_trace[0] = f;
    int it = 0;
    for (; it < max_iter; ++it) {
      double gmax = 0;
      for (double v : g) gmax = std::max(gmax, std::fabs(v));
      if (gmax <= gtol) break;
      // two-loop recursion: p = -H g
      q = g;
      std::vector<double> alpha(S.size());
      for (int k = (int)S.size() - 1; k >= 0; --k) {
        alpha[k] = rho[k] * dot(S[k], q);
        for (int t = 0; t < P; ++t) q[t] -= alpha[k] * Y[k][t];
      }
      double gamma = 1.0;
      if (!S.empty()) gamma = dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
      for (int t = 0; t < P; ++t) q[t] *= gamma;
      for (size_t k = 0; k < S.size(); ++k) {
        const double beta = rho[k] * dot(Y[k], q);
        for (int t = 0; t < P; ++t) q[t] += (alpha[k] - beta) * S[k][t];
      }
      for (int t = 0; t < P; ++t) p[t] = -q[t];
      double pg = dot(p, g);
      if (!(pg < 0)) {  // not a descent direction: restart from steepest descent
        S.clear(); Y.clear(); rho.clear();
        const double gn2 = dot(g, g);
        for (int t = 0; t < P; ++t) p[t] = -g[t] / std::sqrt(gn2);
        pg = dot(p, g);
      }
      double step = S.empty() ? std::min(1.0, 1.0 / std::max(std::sqrt(dot(g, g)), 1e-300)) : 1.0;
      double fn = 0;
      bool ok = false;
      for (int bt = 0; bt < 40; ++bt) {
        for (int t = 0; t < P; ++t) xn[t] = x[t] + step * p[t];
        fn = eval(xn, gn);
        if (std::isfinite(fn) && fn <= f + 1e-4 * step * pg) { ok = true; break; }
        step *= 0.5;
      }
      if (!ok) {  // line search failed: keep the last accepted point
        eval(x, g);
        break;
      }
      std::vector<double> s(P), y(P);
      for (int t = 0; t < P; ++t) { s[t] = xn[t] - x[t]; y[t] = gn[t] - g[t]; }
      const double sy = dot(s, y);
      if (sy > 1e-10 * std::sqrt(dot(s, s) * dot(y, y))) {
        S.push_back(s); Y.push_back(y); rho.push_back(1.0 / sy);
        if ((int)S.size() > history) { S.erase(S.begin()); Y.erase(Y.begin()); rho.erase(rho.begin()); }
      }
      x.swap(xn);
      g.swap(gn);
      f = fn;
      if (loss_trace) loss_trace[it + 1] = f;
    }
    // leave the model at the final accepted point
    std::memcpy(m->model.theta.data(), x.data(), sizeof(double) * P);
    m->dev.upload_theta(m->model);
    if (iters_out) *iters_out = it;
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
