// Parameter gradient and training objective of the SUS2-SH path: shape
// dispatch of the forward-over-reverse kernels (sus2_grad_impl.hpp) and the
// per-frame loss with its adjoints (dev_src/non_linear_regression.cpp:936-1065).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sus2amd.h"
#include "sus2_common.hpp"

namespace sus2 {

template <class T>
void run_grad(const DevModelSet& dm, Workspace& ws, const Batch& b, const GradWeights& gw, double* grad,
              double scale, bool add, bool reuse, hipStream_t st) {
  shape_ops(*dm.host).grad[prec_index<T>()](dm, ws, b, gw, grad, scale, add, reuse, st);
}

void grad_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* ene_w,
                   const double* frc_w, const double* str_w, double* grad, double scale, bool add,
                   bool reuse_neighbors, hipStream_t st) {
  check_rec_limits(precision, b.natoms, dm.host->C);
  const GradWeights gw{ene_w, frc_w, str_w, nullptr};  // atom_frame is the workspace's
  if (precision == SUS2_FP64) run_grad<double>(dm, ws, b, gw, grad, scale, add, reuse_neighbors, st);
  else if (precision == SUS2_FP32) run_grad<float>(dm, ws, b, gw, grad, scale, add, reuse_neighbors, st);
  else throw Error(SUS2_E_INVALID, "precision must be SUS2_FP64 or SUS2_FP32");
  ws.prof.collect_ready();
}

// ------------------------------------------------------------------ loss
// Frame-term options of the objective on the device (ObjectiveOpts without
// the penalty).
struct LossOpts {
  double rel, lcs, stdw, stddw;
  int nbf, kind;
};
constexpr int STD_TYPES = 64;  // species of the std terms' per-frame tables

// Force residual loss and its derivative (dev_src/force_loss.h): L2 d^2, 2 d;
// log-cosh s^2 log cosh(d / s) in the overflow-free form
// |x| + log1p(exp(-2|x|)) - log 2, and s tanh(d / s).
__device__ __forceinline__ double force_rho(double d, int kind, double s) {
  if (kind == 0) return d * d;
  const double x = fabs(d / s);
  return s * s * (x + log1p(exp(-2.0 * x)) - 0.69314718055994530942);
}
__device__ __forceinline__ double force_drho(double d, int kind, double s) {
  return kind == 0 ? 2.0 * d : s * tanh(d / s);
}

// Sum of v over the 256 threads in a fixed tree order, returned to every thread.
__device__ __forceinline__ double block_sum256(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// One 256-thread block per frame, FP64, fixed reduction order.  Follows
// NonLinearRegression::AddLossGrad (dev_src/non_linear_regression.cpp:936-1052):
// weighting 0 vibrations, 1 structures (energy, force and stress weights / N),
// 2 molecules (energy and stress weights * N); force weight per atom
// wF rel / (|F0_i|^2 + rel) when rel > 0; energy and force terms scaled by
// 0.1 / (0.1 + nbf mean |F0|^2); force residual L2 or log-cosh; with the std
// terms, per-species site-energy spread regularisers from the site energies
// `se`.  Writes the adjoints handed to the model: ene_w_i = dL/dE_i, frc_w_i =
// dL/dF_i, str_w = -2 w_S (S - S0) / N (the reference's sign convention,
// :1027), and the frame loss.
__global__ __launch_bounds__(256) void k_loss_frames(int nframes, const int64_t* __restrict__ frame_off,
                                                     const int32_t* __restrict__ types, int C,
                                                     const double* __restrict__ E, const double* __restrict__ F,
                                                     const double* __restrict__ V, const double* __restrict__ se,
                                                     const double* __restrict__ e0, const double* __restrict__ f0,
                                                     const double* __restrict__ s0, int mode, double wE, double wF,
                                                     double wS, LossOpts o, double* __restrict__ ene_w,
                                                     double* __restrict__ frc_w, double* __restrict__ str_w,
                                                     double* __restrict__ loss_f) {
  __shared__ double red[256];
  __shared__ double tmean[STD_TYPES], tjoint[STD_TYPES], tcnt[STD_TYPES];
  const int f = blockIdx.x, t = threadIdx.x;
  const long b = frame_off[f], e = frame_off[f + 1];
  const int n = (int)(e - b);
  auto f0sq = [&](long i) {
    return f0[3 * i] * f0[3 * i] + f0[3 * i + 1] * f0[3 * i + 1] + f0[3 * i + 2] * f0[3 * i + 2];
  };
  // energy and force scale of norm_by_forces (1 when it is off)
  double nb = 1.0;
  if (o.nbf != 0 && n > 0) {
    double a = 0;
    for (long i = b + t; i < e; i += 256) a += f0sq(i) / n;
    nb = 0.1 / (0.1 + o.nbf * block_sum256(a, red));
  }
  double fl = 0;
  for (long i = b + t; i < e; i += 256) {
    double wi = o.rel > 0 ? wF * o.rel / (f0sq(i) + o.rel) : wF;
    if (mode == 1) wi /= n;
    const double fs = wi * nb;
    for (int a = 0; a < 3; ++a) {
      const double d = F[3 * i + a] - f0[3 * i + a];
      if (wF != 0) {
        fl += fs * force_rho(d, o.kind, o.lcs);
        frc_w[3 * i + a] = fs * force_drho(d, o.kind, o.lcs);
      } else {
        frc_w[3 * i + a] = 0;
      }
    }
  }
  const double floss = block_sum256(fl, red);
  double we = wE, ws = wS;
  if (mode == 1 && n > 0) { we /= n; ws /= n; }
  else if (mode == 2) { we *= n; ws *= n; }
  const double dE = (we != 0 && n > 0) ? 2.0 * we * (E[f] - e0[f]) * nb / n : 0.0;
  // per-species site-energy spread terms
  const bool std_terms = (o.stdw != 0 || o.stddw != 0) && n > 0;
  double stdl = 0, stddl = 0;
  const double mean0 = n > 0 ? e0[f] / n : 0.0;
  if (std_terms) {
    for (int ty = 0; ty < C; ++ty) {
      double c = 0;
      for (long i = b + t; i < e; i += 256) c += types[i] == ty ? 1.0 : 0.0;
      const double cnt = block_sum256(c, red);
      if (t == 0) tcnt[ty] = cnt;
      if (cnt == 0) continue;
      double s = 0;
      for (long i = b + t; i < e; i += 256)
        if (types[i] == ty) s += se[i] / cnt;
      const double mt = block_sum256(s, red);
      double j = 0;
      for (long i = b + t; i < e; i += 256)
        if (types[i] == ty) j += (se[i] - mt) * 200.0 / (200.0 + f0sq(i)) / cnt / n;
      const double jt = block_sum256(j, red);
      if (t == 0) {
        tmean[ty] = mt;
        tjoint[ty] = jt;
      }
      stddl += (mt - mean0) * (mt - mean0);
    }
    __syncthreads();
    double sv = 0;
    for (long i = b + t; i < e; i += 256) {
      const double d = se[i] - tmean[types[i]];
      sv += d * d * 200.0 / (200.0 + f0sq(i)) / n;
    }
    stdl = block_sum256(sv, red);
  }
  for (long i = b + t; i < e; i += 256) {
    double w = dE;
    if (std_terms) {
      const int ty = types[i];
      w += o.stdw * 2.0 * ((se[i] - tmean[ty]) * 200.0 / (200.0 + f0sq(i)) / n - tjoint[ty]) +
           o.stddw * 2.0 * (tmean[ty] - mean0) / tcnt[ty];
    }
    ene_w[i] = w;
  }
  if (t == 0) {
    double loss = floss;
    // PlusStress as a full matrix from Voigt xx yy zz xy xz yz
    const int vi[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    double fro = 0;
    for (int q = 0; q < 9; ++q) {
      const double d = V[6 * f + vi[q]] - s0[6 * f + vi[q]];
      fro += d * d;
      str_w[9 * f + q] = (ws != 0 && n > 0) ? -2.0 * ws * d / n : 0.0;
    }
    if (ws != 0 && n > 0) loss += ws * fro / n;
    if (we != 0 && n > 0) loss += we * (E[f] - e0[f]) * (E[f] - e0[f]) * nb / n;
    if (std_terms) loss += o.stdw * stdl + o.stddw * stddl;
    loss_f[f] = loss;
  }
}

// Legacy radial penalty of MLMTPR::AddPenaltyGrad (dev_src/mtpr.cpp:14050)
// over the F outer radial rows c_mu (B basis coefficients each, rows `stride`
// apart from theta[o_rad]): value w (sum_mu (|c_mu|^2 - 1)^2 + sum_{mu<nu}
// (c_mu . c_nu)^2) into value[0], its gradient by (mu, b) into pgrad [F B],
// each entry's terms added in the reference's loop order.  One block.
constexpr int PEN_ROWS = 64;
__global__ __launch_bounds__(256) void k_penalty(const double* __restrict__ theta, int F, int B, int stride,
                                                 int o_rad, double w, double* __restrict__ value,
                                                 double* __restrict__ pgrad) {
  __shared__ double nrm[PEN_ROWS];
  __shared__ double sc[PEN_ROWS * PEN_ROWS];
  const int t = threadIdx.x;
  const double* c = theta + o_rad;
  for (int k = t; k < F; k += 256) {
    double s = 0;
    for (int l = 0; l < B; ++l) s += c[k * stride + l] * c[k * stride + l];
    nrm[k] = s;
  }
  for (int p = t; p < F * F; p += 256) {
    const int k1 = p / F, k2 = p % F;
    if (k1 >= k2) continue;
    double s = 0;
    for (int l = 0; l < B; ++l) s += c[k1 * stride + l] * c[k2 * stride + l];
    sc[p] = s;
  }
  __syncthreads();
  if (t == 0) {
    double v = 0;
    for (int k = 0; k < F; ++k) v += w * (nrm[k] - 1) * (nrm[k] - 1);
    for (int k1 = 0; k1 < F; ++k1)
      for (int k2 = k1 + 1; k2 < F; ++k2) v += w * sc[k1 * F + k2] * sc[k1 * F + k2];
    value[0] = v;
  }
  for (int p = t; p < F * B; p += 256) {
    const int k = p / B, l = p % B;
    double g = w * 4 * (nrm[k] - 1) * c[k * stride + l];
    for (int k1 = 0; k1 < k; ++k1) g += w * 2 * sc[k1 * F + k] * c[k1 * stride + l];
    for (int k2 = k + 1; k2 < F; ++k2) g += w * 2 * sc[k * F + k2] * c[k2 * stride + l];
    pgrad[p] = g;
  }
}

__global__ void k_penalty_add(int F, int B, int stride, int o_rad, const double* __restrict__ pgrad,
                              double* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= F * B) return;
  grad[o_rad + (p / B) * stride + p % B] += pgrad[p];
}

// out[0] = (1/nframes) sum_f loss_f (fixed order) + penalty[0] (when given); with report, also its low
// and high 32 bits in report[0..1], each word tagged with seq in its upper half.
__global__ __launch_bounds__(256) void k_loss_sum(int nframes, const double* __restrict__ loss_f,
                                                  const double* __restrict__ penalty, double* __restrict__ out,
                                                  int64_t* report, int64_t seq) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0;
  for (int f = t; f < nframes; f += 256) s += loss_f[f];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) {
    double v = nframes > 0 ? red[0] / nframes : 0.0;
    if (penalty) v += penalty[0];
    out[0] = v;
    if (report) {
      const uint64_t bits = (uint64_t)__double_as_longlong(v), tag = (uint64_t)(seq & 0x7fffffff) << 32;
      __hip_atomic_store(report, (int64_t)(tag | (bits & 0xffffffffu)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(report + 1, (int64_t)(tag | (bits >> 32)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

__global__ void k_fill(int n, double v, double* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// FP64 edge records -> FP32 (same order, neighbour and species bits repacked).
__global__ void k_rec_narrow(long n, const Rec<double>* __restrict__ in, Rec<float>* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const Rec<double> q = in[e];
  out[e] = Rec<float>{(float)q.x, (float)q.y, (float)q.z, rec_jbits(rec_j(q), rec_t(q), 0.0f)};
}

void ObjectiveOpts::validate() const {
  if (!(rel_forces >= 0) || !std::isfinite(rel_forces)) throw Error(SUS2_E_INVALID, "rel_forces must be finite and >= 0");
  if (norm_by_forces < 0) throw Error(SUS2_E_INVALID, "norm_by_forces must be >= 0");
  if (force_loss != SUS2_FORCE_L2 && force_loss != SUS2_FORCE_LOG_COSH)
    throw Error(SUS2_E_INVALID, "force_loss must be SUS2_FORCE_L2 or SUS2_FORCE_LOG_COSH");
  if (!(log_cosh_scale > 0) || !std::isfinite(log_cosh_scale))
    throw Error(SUS2_E_INVALID, "log_cosh_scale must be a finite positive value");
  if (!(std_weight >= 0) || !std::isfinite(std_weight) || !(stdd_weight >= 0) || !std::isfinite(stdd_weight))
    throw Error(SUS2_E_INVALID, "std_weight and stdd_weight must be finite and >= 0");
  if (!std::isfinite(constr_weight)) throw Error(SUS2_E_INVALID, "constr_weight must be finite");
}

double radial_penalty(const Model& m, const double* theta, double w, double* grad) {
  if (w == 0) return 0.0;
  const int F = m.F, B = m.B, stride = m.B + m.C;
  const double* c = theta + m.o_rad;
  std::vector<double> nrm(F, 0.0);
  double v = 0;
  for (int k = 0; k < F; ++k) {
    for (int l = 0; l < B; ++l) nrm[k] += c[k * stride + l] * c[k * stride + l];
    v += w * (nrm[k] - 1) * (nrm[k] - 1);
    if (grad)
      for (int l = 0; l < B; ++l) grad[m.o_rad + k * stride + l] += w * 4 * (nrm[k] - 1) * c[k * stride + l];
  }
  for (int k1 = 0; k1 < F; ++k1)
    for (int k2 = k1 + 1; k2 < F; ++k2) {
      double s = 0;
      for (int l = 0; l < B; ++l) s += c[k1 * stride + l] * c[k2 * stride + l];
      v += w * s * s;
      if (grad)
        for (int l = 0; l < B; ++l) {
          grad[m.o_rad + k1 * stride + l] += w * 2 * s * c[k2 * stride + l];
          grad[m.o_rad + k2 * stride + l] += w * 2 * s * c[k1 * stride + l];
        }
    }
  return v;
}

void loss_grad_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* e0,
                        const double* f0, const double* s0, int mode, double wE, double wF, double wS,
                        const ObjectiveOpts& obj, double* loss_host, double* grad, hipStream_t st, double* loss_dev,
                        const EfsOut* efs_out) {
  if (mode < 0 || mode > 2) throw Error(SUS2_E_INVALID, "weighting must be 0 (vibrations), 1 (structures) or 2 (molecules)");
  obj.validate();
  if (obj.std_terms() && dm.host->C > STD_TYPES)
    throw Error(SUS2_E_UNSUPPORTED, "the std terms support at most 64 species");
  if (precision != SUS2_FP64 && precision != SUS2_FP32 && precision != SUS2_MIXED)
    throw Error(SUS2_E_INVALID, "precision must be SUS2_FP64, SUS2_FP32 or SUS2_MIXED");
  // mixed: the loss residuals come from an FP64 EFS, the sweeps run in FP32
  const int efs_prec = precision == SUS2_MIXED ? SUS2_FP64 : precision;
  const int grad_prec = precision == SUS2_MIXED ? SUS2_FP32 : precision;
  if (precision == SUS2_MIXED) check_rec_limits(SUS2_FP32, b.natoms, dm.host->C);
  const int nf = b.nframes, N = b.natoms;
  if (nf == 0) {  // empty batch: zero objective and gradient
    const int P = dm.host->ncoeff;
    SUS2_LAUNCH(ws, st, "fill", k_fill, dim3((P + 255) / 256), dim3(256), 0, P, 0.0, grad);
    if (loss_dev) SUS2_LAUNCH(ws, st, "fill", k_fill, dim3(1), dim3(256), 0, 1, 0.0, loss_dev);
    if (loss_host) *loss_host = 0.0;
    return;
  }
  const size_t n1 = std::max(1, N), f1 = std::max(1, nf);
  ws.lE.ensure(sizeof(double) * f1);
  ws.lF.ensure(sizeof(double) * 3 * n1);
  ws.lV.ensure(sizeof(double) * 6 * f1);
  ws.lew.ensure(sizeof(double) * n1);
  ws.lfw.ensure(sizeof(double) * 3 * n1);
  ws.lsw.ensure(sizeof(double) * 9 * f1);
  ws.lloss.ensure(sizeof(double) * (f1 + 1));
  const Model& hm = *dm.host;
  const bool pen = obj.constr_weight != 0;
  if (pen && hm.F > PEN_ROWS) throw Error(SUS2_E_UNSUPPORTED, "the radial penalty supports at most 64 radial rows");
  if (pen) ws.lpen.ensure(sizeof(double) * (1 + (size_t)hm.F * hm.B));
  if (obj.std_terms()) ws.lsite.ensure(sizeof(double) * n1);
  EfsOut out;
  out.energy = efs_out && efs_out->energy ? efs_out->energy : ws.lE.get<double>();
  out.force = efs_out && efs_out->force ? efs_out->force : ws.lF.get<double>();
  out.virial = efs_out && efs_out->virial ? efs_out->virial : ws.lV.get<double>();
  out.site = efs_out && efs_out->site ? efs_out->site : obj.std_terms() ? ws.lsite.get<double>() : nullptr;
  efs_dispatch(dm, ws, efs_prec, b, out, st);
  double* lf = ws.lloss.get<double>();
  const LossOpts lo{obj.rel_forces, obj.log_cosh_scale, obj.std_weight, obj.stdd_weight, obj.norm_by_forces,
                    obj.force_loss};
  SUS2_LAUNCH(ws, st, "loss_frames", k_loss_frames, dim3(f1), dim3(256), 0, nf, b.frame_off, b.types, hm.C,
              out.energy, out.force, out.virial, out.site, e0, f0, s0, mode, wE, wF, wS, lo, ws.lew.get<double>(),
              ws.lfw.get<double>(), ws.lsw.get<double>(), lf + 1);
  // the penalty of the FP64 coefficients (every precision)
  const int stride = hm.B + hm.C;
  if (pen)
    SUS2_LAUNCH(ws, st, "penalty", k_penalty, dim3(1), dim3(256), 0, dm.get<double>().theta, hm.F, hm.B, stride,
                hm.o_rad, obj.constr_weight, ws.lpen.get<double>(), ws.lpen.get<double>() + 1);
  // capture-safe calls write the objective to the caller's device word only
  const bool polled = !ws.capture && loss_host;
  const int64_t seq = polled ? ++ws.report_seq : 0;
  SUS2_LAUNCH(ws, st, "loss_sum", k_loss_sum, dim3(1), dim3(256), 0, nf, lf + 1,
              pen ? ws.lpen.get<double>() : nullptr, loss_dev ? loss_dev : lf, polled ? report_word(ws, 4) : nullptr,
              seq);
  const bool narrow = precision == SUS2_MIXED && ws.nedges > 0;
  if (narrow) {  // the neighbour list of the FP64 EFS, narrowed for the FP32 sweeps
    ws.rec_alt.ensure(sizeof(Rec<float>) * (size_t)ws.nedges);
    SUS2_LAUNCH(ws, st, "rec_narrow", k_rec_narrow, dim3((unsigned)((ws.nedges + 255) / 256)), dim3(256), 0,
                ws.nedges, ws.rec.get<Rec<double>>(), ws.rec_alt.get<Rec<float>>());
    ws.rec.swap(ws.rec_alt);
  }
  grad_dispatch(dm, ws, grad_prec, b, ws.lew.get<double>(), ws.lfw.get<double>(), ws.lsw.get<double>(), grad,
                1.0 / nf, false, true, st);
  if (narrow) ws.rec.swap(ws.rec_alt);  // the FP64 records stay in place for the next build
  if (pen)
    SUS2_LAUNCH(ws, st, "penalty_add", k_penalty_add, dim3((hm.F * hm.B + 255) / 256), dim3(256), 0, hm.F, hm.B,
                stride, hm.o_rad, ws.lpen.get<double>() + 1, grad);
  if (!polled) return;
  // the objective is known once the EFS and the loss kernels have run: the
  // call returns then, and the gradient kernels finish on the stream
  uint64_t w[2];
  poll_report(ws, 4, 2, seq, st, "training objective", w);
  const uint64_t bits = (w[0] & 0xffffffffu) | (w[1] << 32);
  double v;
  std::memcpy(&v, &bits, sizeof(double));
  *loss_host = v;
}

}  // namespace sus2
