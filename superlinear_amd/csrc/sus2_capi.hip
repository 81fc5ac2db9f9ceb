// extern "C" boundary of the SUS2-SH MI355X path (declared in include/sus2amd.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sus2amd.h"
#include "sus2_model.hpp"
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
