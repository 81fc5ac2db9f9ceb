// Parameter gradient and training objective of the SUS2-SH path: shape
// dispatch of the forward-over-reverse kernels (sus2_grad_impl.hpp) and the
// per-frame loss with its adjoints (dev_src/non_linear_regression.cpp:936-1065).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

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
// One 256-thread block per frame, FP64, fixed reduction order.  Follows
// NonLinearRegression::AddLoss / AddLossGrad with norm_by_forces = 0 and the
// L2 force residual: weighting 0 vibrations, 1 structures (energy, force and
// stress weights / N), 2 molecules (energy and stress weights * N).  Writes the
// adjoints handed to the model: ene_w_i = dL/dE (same for every atom of the
// frame), frc_w_i = dL/dF_i, str_w = -2 w_S (S - S0) / N (the reference's
// sign convention, non_linear_regression.cpp:1027), and the frame loss.
__global__ __launch_bounds__(256) void k_loss_frames(int nframes, const int64_t* __restrict__ frame_off,
                                                     const double* __restrict__ E, const double* __restrict__ F,
                                                     const double* __restrict__ V, const double* __restrict__ e0,
                                                     const double* __restrict__ f0, const double* __restrict__ s0,
                                                     int mode, double wE, double wF, double wS,
                                                     double* __restrict__ ene_w, double* __restrict__ frc_w,
                                                     double* __restrict__ str_w, double* __restrict__ loss_f) {
  __shared__ double red[256];
  const int f = blockIdx.x, t = threadIdx.x;
  const long b = frame_off[f], e = frame_off[f + 1];
  const int n = (int)(e - b);
  double w = wF;
  if (mode == 1 && n > 0) w /= n;
  double fl = 0;
  for (long i = b + t; i < e; i += 256)
    for (int a = 0; a < 3; ++a) {
      const double d = F[3 * i + a] - f0[3 * i + a];
      if (wF != 0) {
        fl += d * d;
        frc_w[3 * i + a] = w * 2 * d;
      } else {
        frc_w[3 * i + a] = 0;
      }
    }
  red[t] = fl;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  double we = wE, ws = wS;
  if (mode == 1 && n > 0) { we /= n; ws /= n; }
  else if (mode == 2) { we *= n; ws *= n; }
  const double dE = (we != 0 && n > 0) ? 2.0 * we * (E[f] - e0[f]) / n : 0.0;
  for (long i = b + t; i < e; i += 256) ene_w[i] = dE;
  if (t == 0) {
    double loss = w * red[0];
    // PlusStress as a full matrix from Voigt xx yy zz xy xz yz
    const int vi[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    double fro = 0;
    for (int q = 0; q < 9; ++q) {
      const double d = V[6 * f + vi[q]] - s0[6 * f + vi[q]];
      fro += d * d;
      str_w[9 * f + q] = (ws != 0 && n > 0) ? -2.0 * ws * d / n : 0.0;
    }
    if (ws != 0 && n > 0) loss += ws * fro / n;
    if (we != 0 && n > 0) loss += we * (E[f] - e0[f]) * (E[f] - e0[f]) / n;
    loss_f[f] = loss;
  }
}

// out[0] = (1/nframes) sum_f loss_f (fixed order); with report, also its low
// and high 32 bits in report[0..1], each word tagged with seq in its upper half.
__global__ __launch_bounds__(256) void k_loss_sum(int nframes, const double* __restrict__ loss_f,
                                                  double* __restrict__ out, int64_t* report, int64_t seq) {
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
    const double v = nframes > 0 ? red[0] / nframes : 0.0;
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

void loss_grad_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* e0,
                        const double* f0, const double* s0, int mode, double wE, double wF, double wS,
                        double* loss_host, double* grad, hipStream_t st, double* loss_dev, const EfsOut* efs_out) {
  if (mode < 0 || mode > 2) throw Error(SUS2_E_INVALID, "weighting must be 0 (vibrations), 1 (structures) or 2 (molecules)");
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
  EfsOut out;
  out.energy = efs_out && efs_out->energy ? efs_out->energy : ws.lE.get<double>();
  out.force = efs_out && efs_out->force ? efs_out->force : ws.lF.get<double>();
  out.virial = efs_out && efs_out->virial ? efs_out->virial : ws.lV.get<double>();
  out.site = efs_out ? efs_out->site : nullptr;
  efs_dispatch(dm, ws, efs_prec, b, out, st);
  double* lf = ws.lloss.get<double>();
  SUS2_LAUNCH(ws, st, "loss_frames", k_loss_frames, dim3(f1), dim3(256), 0, nf, b.frame_off, out.energy, out.force,
              out.virial, e0, f0, s0, mode, wE, wF, wS, ws.lew.get<double>(), ws.lfw.get<double>(),
              ws.lsw.get<double>(), lf + 1);
  // capture-safe calls write the objective to the caller's device word only
  const bool polled = !ws.capture && loss_host;
  const int64_t seq = polled ? ++ws.report_seq : 0;
  SUS2_LAUNCH(ws, st, "loss_sum", k_loss_sum, dim3(1), dim3(256), 0, nf, lf + 1, loss_dev ? loss_dev : lf,
              polled ? report_word(ws, 4) : nullptr, seq);
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
