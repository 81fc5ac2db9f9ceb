// Linear readout solve of the training objective (the reference's
// TrainLinear: normal equations over the coefficients that enter the energy
// linearly, dev_src/mtpr_trainer.cpp:1130-1455).
//
// With every other coefficient fixed, a per-element readout model's energy is
//   E_f = sum_i shift[t_i] + sum_t species[t] n_t(f) + sum_{t,q} read[t,q] S_{tq}(f),
// S_{tq}(f) = sum over atoms i of type t of the scalar v_i[q], so E, the
// forces and the virial are linear in x = (species, read).  Column c of the
// design matrix J is the EFS of the model with x = e_c minus the EFS with
// x = 0 (a readout column's forces and virial come from that EFS; a species
// column has n_t(f) in the energy rows only).  The rows of frame f, in order:
// its energy, 3 N forces, 9 virial entries, scaled by the square roots of the
// objective's weights (AddToSLAE, dev_src/mtpr_trainer.cpp:1186-1448):
//   energy  wE / N   (structures: / N again; molecules: * N)
//   forces  wF       (structures: / N)
//   virial  wS / N   (structures: / N again; molecules: * N)
// with the energy and force rows times 0.1 / (0.1 + norm_by_forces mean |F0|^2);
// the objective's other options are not part of these equations.
// A = J^T J is formed on the device by FP64 MFMA tiles (16 x 16 x 4) over
// row chunks and reduced in a fixed order; b = J^T r likewise in FP64 VALU;
// the n x n system (n = C + C Q) is solved on the host by Cholesky.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/sus2amd.h"
#include "sus2_common.hpp"

namespace sus2 {

namespace {

constexpr int TILE = 16;        // MFMA output tile (16 x 16 FP64)
constexpr int CHUNK_ROWS = 2048;  // rows of J per partial product

typedef double d4_t __attribute__((ext_vector_type(4)));

// Rows of J of frame f start at row_off[f]: 1 energy row, 3 N force rows and
// (with stress) 9 virial rows.
struct RowMap {
  const int64_t* frame_off;
  const int64_t* row_off;
  const double* wrow;  // [nframes][3] square-root weights: energy, force, virial
  int nframes, stress;
};

// Column c of J (ld columns, row-major) from the EFS outputs of x = e_c
// (E, F, V) minus those of x = 0 (E0, F0, V0); with rhs, the label column r
// (labels minus the x = 0 EFS) instead.  One thread per atom, plus one per
// frame for its energy and virial rows.
__global__ void k_design_column(RowMap rm, int natoms, int ld, int c, const double* __restrict__ E,
                                const double* __restrict__ F, const double* __restrict__ V,
                                const double* __restrict__ E0, const double* __restrict__ F0,
                                const double* __restrict__ V0, double* __restrict__ J, int rhs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < natoms) {
    const int f = frame_of(rm.frame_off, rm.nframes, t);
    const long k = t - rm.frame_off[f];
    const double w = rm.wrow[3 * f + 1];
    for (int a = 0; a < 3; ++a) {
      const long row = rm.row_off[f] + 1 + 3 * k + a;
      const double v = w * (F[3 * (size_t)t + a] - F0[3 * (size_t)t + a]);
      if (rhs) J[row] = v;
      else J[(size_t)row * ld + c] = v;
    }
    return;
  }
  const int f = t - natoms;
  if (f >= rm.nframes) return;
  const long r0 = rm.row_off[f];
  const double we = rm.wrow[3 * f] * (E[f] - E0[f]);
  if (rhs) J[r0] = we;
  else J[(size_t)r0 * ld + c] = we;
  if (!rm.stress) return;
  const long n = rm.frame_off[f + 1] - rm.frame_off[f];
  const int vi[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};  // full 3 x 3 from Voigt xx yy zz xy xz yz
  for (int q = 0; q < 9; ++q) {
    const long row = r0 + 1 + 3 * n + q;
    const double v = rm.wrow[3 * f + 2] * (V[6 * f + vi[q]] - V0[6 * f + vi[q]]);
    if (rhs) J[row] = v;
    else J[(size_t)row * ld + c] = v;
  }
}

// Species column: n_t(f) in the energy rows (zero forces and virial).
__global__ void k_species_column(RowMap rm, const int32_t* __restrict__ types, int ld, int c, int t,
                                 double* __restrict__ J) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= rm.nframes) return;
  long cnt = 0;
  for (long i = rm.frame_off[f]; i < rm.frame_off[f + 1]; ++i) cnt += types[i] == t;
  J[(size_t)rm.row_off[f] * ld + c] = rm.wrow[3 * f] * (double)cnt;
}

// Partial J^T J over rows [z CHUNK_ROWS, (z + 1) CHUNK_ROWS) for output tile
// (bi, bj), bi <= bj: one wavefront, one FP64 MFMA 16 x 16 x 4 per four rows.
// Operands: lane l supplies A(i = l % 16, k = l / 16) = J[r + k][16 bi + i]
// and B(k = l / 16, j = l % 16) = J[r + k][16 bj + j]; the FP64 accumulator's
// element v of lane l is D(4 v + l / 16, l % 16) (rows interleaved over the
// four lane groups, one row per register).  The solver checks the diagonal
// against an FP64 VALU sum of squares before use.  ld is a multiple of 16 and
// the rows of J are padded with zeros to a multiple of 4.
__global__ __launch_bounds__(WAVE) void k_gram_mfma(long rows, int ld, const double* __restrict__ J,
                                                    double* __restrict__ part) {
  const int bi = blockIdx.x, bj = blockIdx.y, z = blockIdx.z;
  if (bi > bj) return;
  const int lane = threadIdx.x;
  const long r0 = (long)z * CHUNK_ROWS, r1 = min(rows, r0 + CHUNK_ROWS);
  const int ci = TILE * bi + lane % 16, cj = TILE * bj + lane % 16, kk = lane / 16;
  d4_t acc = {0.0, 0.0, 0.0, 0.0};
  for (long r = r0; r < r1; r += 4) {
    const double* row = J + (size_t)(r + kk) * ld;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(row[ci], row[cj], acc, 0, 0, 0);
  }
  double* P = part + (size_t)z * ld * ld;
#pragma unroll
  for (int v = 0; v < 4; ++v) P[(size_t)(TILE * bi + 4 * v + kk) * ld + cj] = acc[v];
}

// diag[c] = sum_r J[r][c]^2 in row order (FP64 VALU; the check of the MFMA Gram matrix).
__global__ void k_gram_diag(long rows, int ld, const double* __restrict__ J, double* __restrict__ diag) {
  __shared__ double red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  double s = 0;
  for (long k = t; k < rows; k += 256) {
    const double v = J[(size_t)k * ld + c];
    s += v * v;
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) diag[c] = red[0];
}

// A[i][j] = sum_z part[z][i][j] in chunk order for the upper tiles, mirrored
// into the lower triangle; b[i] = sum_r J[r][i] r[r] in row order per column.
__global__ void k_gram_reduce(int ld, int nz, const double* __restrict__ part, double* __restrict__ A) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ld) return;
  const int ui = min(i, j), uj = max(i, j);
  if ((ui / TILE) > (uj / TILE)) return;
  double s = 0;
  for (int z = 0; z < nz; ++z) s += part[(size_t)z * ld * ld + (size_t)ui * ld + uj];
  A[(size_t)i * ld + j] = s;
}

__global__ void k_gram_rhs(long rows, int ld, const double* __restrict__ J, const double* __restrict__ r,
                           double* __restrict__ b) {
  __shared__ double red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  double s = 0;
  for (long k = t; k < rows; k += 256) s += J[(size_t)k * ld + c] * r[k];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) b[c] = red[0];
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(SUS2_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Cholesky solve of the SPD system A x = b (n x n, leading dimension ld); a
// failed factorisation is retried with a diagonal shift of 1e-12 trace / n,
// growing tenfold.
std::vector<double> spd_solve(std::vector<double> A, int n, int ld, std::vector<double> b) {
  double tr = 0;
  for (int i = 0; i < n; ++i) tr += A[(size_t)i * ld + i];
  const std::vector<double> A0 = A;
  for (double shift = 0; shift < 1e-2 * std::max(tr, 1e-300);
       shift = shift == 0 ? 1e-12 * std::max(tr, 1e-300) / std::max(n, 1) : shift * 10) {
    A = A0;
    bool ok = true;
    for (int j = 0; j < n && ok; ++j) {
      double d = A[(size_t)j * ld + j] + shift;
      for (int k = 0; k < j; ++k) d -= A[(size_t)j * ld + k] * A[(size_t)j * ld + k];
      if (!(d > 0)) { ok = false; break; }
      const double ljj = std::sqrt(d);
      A[(size_t)j * ld + j] = ljj;
      for (int i = j + 1; i < n; ++i) {
        double s = A[(size_t)i * ld + j];
        for (int k = 0; k < j; ++k) s -= A[(size_t)i * ld + k] * A[(size_t)j * ld + k];
        A[(size_t)i * ld + j] = s / ljj;
      }
    }
    if (!ok) continue;
    std::vector<double> y(n);
    for (int i = 0; i < n; ++i) {
      double s = b[i];
      for (int k = 0; k < i; ++k) s -= A[(size_t)i * ld + k] * y[k];
      y[i] = s / A[(size_t)i * ld + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * ld + i] * y[k];
      y[i] = s / A[(size_t)i * ld + i];
    }
    return y;
  }
  throw Error(SUS2_E_INVALID, "linear readout system is not positive definite");
}

}  // namespace

double train_linear(Model& host, DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* e0,
                    const double* f0, const double* s0, int mode, double wE, double wF, double wS, double ridge,
                    const ObjectiveOpts& obj, hipStream_t st) {
  obj.validate();
  if (!host.species_readout)
    throw Error(SUS2_E_UNSUPPORTED, "linear readout solve needs a per-element readout (readout_mode = species)");
  if (mode < 0 || mode > 2) throw Error(SUS2_E_INVALID, "weighting must be 0, 1 or 2");
  if (!(ridge >= 0) || !std::isfinite(ridge)) throw Error(SUS2_E_INVALID, "ridge must be finite and >= 0");
  const int nf = b.nframes, N = b.natoms, C = host.C, nread = C * host.Q, n = C + nread;
  const int ld = ((n + TILE - 1) / TILE) * TILE;
  // frame offsets, species and labels on the host
  std::vector<int64_t> off(nf + 1);
  check(hipMemcpy(off.data(), b.frame_off, sizeof(int64_t) * (nf + 1), hipMemcpyDeviceToHost), "frame offsets");
  const int stress = wS != 0 ? 1 : 0;
  // norm_by_forces: the frame's energy and force rows scaled by
  // 0.1 / (0.1 + nbf mean |F0|^2) (dev_src/mtpr_trainer.cpp:1203-1212)
  std::vector<double> lab;
  if (obj.norm_by_forces != 0 && N > 0) {
    lab.resize(3 * (size_t)N);
    check(hipMemcpy(lab.data(), f0, sizeof(double) * lab.size(), hipMemcpyDeviceToHost), "force labels");
  }
  std::vector<int64_t> roff(nf + 1, 0);
  std::vector<double> wrow(3 * (size_t)std::max(nf, 1));
  for (int f = 0; f < nf; ++f) {
    const double na = (double)(off[f + 1] - off[f]);
    roff[f + 1] = roff[f] + 1 + 3 * (off[f + 1] - off[f]) + 9 * stress;
    double we = na > 0 ? wE / na : 0, wf = wF, wsv = na > 0 ? wS / na : 0;
    if (mode == 1 && na > 0) { we /= na; wf /= na; wsv /= na; }
    else if (mode == 2) { we *= na; wsv *= na; }
    if (!lab.empty() && na > 0) {
      double avef = 0;
      for (int64_t i = off[f]; i < off[f + 1]; ++i)
        avef += (lab[3 * i] * lab[3 * i] + lab[3 * i + 1] * lab[3 * i + 1] + lab[3 * i + 2] * lab[3 * i + 2]) / na;
      const double nb = 0.1 / (0.1 + obj.norm_by_forces * avef);
      we *= nb;
      wf *= nb;
    }
    wrow[3 * f] = std::sqrt(std::max(we, 0.0));
    wrow[3 * f + 1] = std::sqrt(std::max(wf, 0.0));
    wrow[3 * f + 2] = std::sqrt(std::max(wsv, 0.0));
  }
  const long rows = roff[nf], rows4 = ((rows + 3) / 4) * 4;
  DevBuf d_roff, d_wrow, d_J, d_r, d_part, d_A, d_b, d_out, d_diag;
  d_roff.ensure(sizeof(int64_t) * (nf + 1));
  d_wrow.ensure(sizeof(double) * wrow.size());
  d_J.ensure(sizeof(double) * (size_t)rows4 * ld);
  d_r.ensure(sizeof(double) * (size_t)rows4);
  d_out.ensure(sizeof(double) * (size_t)(2 * (size_t)nf + 6 * (size_t)nf + 3 * (size_t)N) * 2);
  check(hipMemcpy(d_roff.p, roff.data(), sizeof(int64_t) * (nf + 1), hipMemcpyHostToDevice), "row offsets");
  check(hipMemcpy(d_wrow.p, wrow.data(), sizeof(double) * wrow.size(), hipMemcpyHostToDevice), "row weights");
  check(hipMemsetAsync(d_J.p, 0, sizeof(double) * (size_t)rows4 * ld, st), "clear J");
  check(hipMemsetAsync(d_r.p, 0, sizeof(double) * (size_t)rows4, st), "clear r");
  double* out = d_out.get<double>();
  // EFS outputs of x = 0 (base) and of the current column
  double *E0 = out, *V0 = E0 + nf, *F0 = V0 + 6 * (size_t)nf;
  double *E1 = F0 + 3 * (size_t)N, *V1 = E1 + nf, *F1 = V1 + 6 * (size_t)nf;
  const RowMap rm{b.frame_off, d_roff.get<int64_t>(), d_wrow.get<double>(), nf, stress};
  const std::vector<double> theta = host.theta;
  auto set_x = [&](int col) {  // x = 0, or x = e_col over (species, read)
    for (int k = 0; k < C; ++k) host.theta[host.o_species + k] = 0.0;
    for (int k = 0; k < nread; ++k) host.theta[host.o_read + k] = 0.0;
    if (col >= 0) host.theta[col < C ? host.o_species + col : host.o_read + col - C] = 1.0;
    dm.upload_theta(host);
  };
  const unsigned gcol = (unsigned)((N + nf + 255) / 256);
  try {
    set_x(-1);
    efs_dispatch(dm, ws, precision, b, EfsOut{E0, F0, V0, nullptr}, st);
    // label column r = sqrt(w) (labels - EFS at x = 0); s0 is Voigt, like V
    hipLaunchKernelGGL(k_design_column, dim3(gcol), dim3(256), 0, st, rm, N, ld, 0, e0, f0, s0, E0, F0, V0,
                       d_r.get<double>(), 1);
    check(hipGetLastError(), "label column");
    for (int t = 0; t < C; ++t) {
      hipLaunchKernelGGL(k_species_column, dim3((nf + 255) / 256), dim3(256), 0, st, rm, b.types, ld, t, t,
                         d_J.get<double>());
      check(hipGetLastError(), "species column");
    }
    for (int q = 0; q < nread; ++q) {
      set_x(C + q);
      efs_dispatch(dm, ws, precision, b, EfsOut{E1, F1, V1, nullptr}, st);
      hipLaunchKernelGGL(k_design_column, dim3(gcol), dim3(256), 0, st, rm, N, ld, C + q, E1, F1, V1, E0, F0, V0,
                         d_J.get<double>(), 0);
      check(hipGetLastError(), "design column");
    }
  } catch (...) {
    host.theta = theta;
    dm.upload_theta(host);
    throw;
  }
  // Gram matrix on MFMA tiles and the right-hand side
  const int nt = ld / TILE, nz = (int)((rows4 + CHUNK_ROWS - 1) / CHUNK_ROWS);
  d_part.ensure(sizeof(double) * (size_t)std::max(nz, 1) * ld * ld);
  d_A.ensure(sizeof(double) * (size_t)ld * ld);
  d_b.ensure(sizeof(double) * ld);
  if (nz > 0) {
    hipLaunchKernelGGL(k_gram_mfma, dim3(nt, nt, nz), dim3(WAVE), 0, st, rows4, ld, d_J.get<double>(),
                       d_part.get<double>());
    check(hipGetLastError(), "gram");
  }
  hipLaunchKernelGGL(k_gram_reduce, dim3((ld + 63) / 64, ld), dim3(64), 0, st, ld, nz, d_part.get<double>(),
                     d_A.get<double>());
  hipLaunchKernelGGL(k_gram_rhs, dim3(ld), dim3(256), 0, st, rows4, ld, d_J.get<double>(), d_r.get<double>(),
                     d_b.get<double>());
  d_diag.ensure(sizeof(double) * ld);
  hipLaunchKernelGGL(k_gram_diag, dim3(ld), dim3(256), 0, st, rows4, ld, d_J.get<double>(), d_diag.get<double>());
  check(hipGetLastError(), "gram reduce");
  std::vector<double> A((size_t)ld * ld), rhs(ld), r(rows4), diag(ld);
  check(hipStreamSynchronize(st), "linear system");
  check(hipMemcpy(A.data(), d_A.p, sizeof(double) * A.size(), hipMemcpyDeviceToHost), "A");
  check(hipMemcpy(rhs.data(), d_b.p, sizeof(double) * ld, hipMemcpyDeviceToHost), "b");
  check(hipMemcpy(r.data(), d_r.p, sizeof(double) * rows4, hipMemcpyDeviceToHost), "r");
  check(hipMemcpy(diag.data(), d_diag.p, sizeof(double) * ld, hipMemcpyDeviceToHost), "diag");
  for (int k = 0; k < n; ++k)
    if (!(std::fabs(A[(size_t)k * ld + k] - diag[k]) <= 1e-10 * std::fabs(diag[k]) + 1e-300))
      throw Error(SUS2_E_DEVICE, "MFMA Gram matrix disagrees with its VALU diagonal at column " + std::to_string(k));
  for (int k = C; k < n; ++k) A[(size_t)k * ld + k] += ridge;  // ridge on the readout weights
  rhs.resize(n);
  const std::vector<double> x = spd_solve(A, n, ld, rhs);
  host.theta = theta;
  for (int k = 0; k < C; ++k) host.theta[host.o_species + k] = x[k];
  for (int k = 0; k < nread; ++k) host.theta[host.o_read + k] = x[C + k];
  dm.upload_theta(host);
  // objective at the solution: (r^T r - 2 b^T x + x^T A x) / nframes (without the ridge term)
  double rr = 0, bx = 0, xAx = 0;
  for (long k = 0; k < rows; ++k) rr += r[k] * r[k];
  for (int i = 0; i < n; ++i) {
    bx += rhs[i] * x[i];
    double s = 0;
    for (int j = 0; j < n; ++j) s += A[(size_t)i * ld + j] * x[j];
    xAx += x[i] * (s - (i >= C ? ridge * x[i] : 0.0));
  }
  // the radial penalty does not depend on the solved coefficients
  return (nf > 0 ? (rr - 2 * bx + xAx) / nf : 0.0) + radial_penalty(host, host.theta.data(), obj.constr_weight, nullptr);
}

}  // namespace sus2
