// SUS2-SH energy / force / virial on MI355X (gfx950).
//
// Pass structure (one 64-lane workgroup = one wavefront per centre atom):
//   k_neighbors      full directed neighbour list with periodic images, one
//                    record {dx, dy, dz, j} per edge, in the oracle's order
//   k_moments<gate>  edge phase: each lane evaluates one edge's radial
//                    functions R_mu and harmonics Y_lm and stages w_mu R_mu,
//                    Y_lm in LDS; moment phase: each lane owns moments (h, k)
//                    and sums the 64 staged edges; then the K->rank
//                    contraction and the layered product graph run over LDS.
//                    gate: h_i = W S_i.  outer: site energy and the basic
//                    moment adjoints (graph reverse + transposed contraction).
//   k_force<pass>    gather-based, atomic-free backward: for every edge i->j
//                    the lane evaluates both directed edges, dE_i/dd_ij with
//                    the centre's adjoints and dE_j/dd_ji with the
//                    neighbour's, so F_i = sum_j (g_ij - g_ji) needs no
//                    scatter; the outer pass also gathers the gate adjoint
//                    h_bar_i and runs the gate graph reverse.
//   k_finalize       per-frame energy and virial sums in FP64, fixed order.
//
// Math follows interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp (edge basis
// :166-216, type factors :319-338, contraction :217-221, graph :308-313,
// passes :385-562) and sus2_sl_radial_math.cpp:256-335,468-503.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sus2amd.h"
#include "sus2_common.hpp"

namespace sus2 {

// ------------------------------------------------------------------ neighbours
__device__ __forceinline__ int frame_of(const int64_t* off, int nframes, int i) {
  int lo = 0, hi = nframes;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Full neighbour list of atom i: every atom j of its frame and every periodic
// image with |d| < rc, skipping the atom itself (same order and image ranges
// as interfaces/lammps/tests/embedded_efs.cpp:44-66).  Count pass when rec is
// null, fill pass otherwise.
template <class T>
__global__ void k_neighbors(int natoms, int nframes, const int64_t* __restrict__ frame_off,
                            const double* __restrict__ pos, const double* __restrict__ cell,
                            const int32_t* __restrict__ periodic, double rc, const int* __restrict__ off,
                            int* __restrict__ cnt, Rec<T>* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const int f = frame_of(frame_off, nframes, i);
  const long base = frame_off[f], n = frame_off[f + 1] - base;
  const double* a = cell + 9 * f;
  double inv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, reach[3] = {0, 0, 0};
  const bool per = periodic[f] != 0;
  if (per) {
    double b[3][3];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q)
        b[q][p] = a[((p + 1) % 3) * 3 + (q + 1) % 3] * a[((p + 2) % 3) * 3 + (q + 2) % 3] -
                  a[((p + 1) % 3) * 3 + (q + 2) % 3] * a[((p + 2) % 3) * 3 + (q + 1) % 3];
    double det = a[0] * b[0][0] + a[1] * b[1][0] + a[2] * b[2][0];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) inv[p][q] = b[p][q] / det;
    for (int q = 0; q < 3; ++q) reach[q] = rc * sqrt(inv[0][q] * inv[0][q] + inv[1][q] * inv[1][q] + inv[2][q] * inv[2][q]);
  }
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const double rc2 = rc * rc;
  int count = 0;
  const int w0 = rec ? off[i] : 0;
  for (long jj = 0; jj < n; ++jj) {
    const long j = base + jj;
    const double d0 = pos[3 * j] - xi, d1 = pos[3 * j + 1] - yi, d2 = pos[3 * j + 2] - zi;
    int lo0 = 0, lo1 = 0, lo2 = 0, hi0 = 0, hi1 = 0, hi2 = 0;
    if (per) {
      const double f0 = d0 * inv[0][0] + d1 * inv[1][0] + d2 * inv[2][0];
      const double f1 = d0 * inv[0][1] + d1 * inv[1][1] + d2 * inv[2][1];
      const double f2 = d0 * inv[0][2] + d1 * inv[1][2] + d2 * inv[2][2];
      lo0 = (int)ceil(-f0 - reach[0]); hi0 = (int)floor(-f0 + reach[0]);
      lo1 = (int)ceil(-f1 - reach[1]); hi1 = (int)floor(-f1 + reach[1]);
      lo2 = (int)ceil(-f2 - reach[2]); hi2 = (int)floor(-f2 + reach[2]);
    }
    for (int p = lo0; p <= hi0; ++p)
      for (int q = lo1; q <= hi1; ++q)
        for (int s = lo2; s <= hi2; ++s) {
          if (j == i && p == 0 && q == 0 && s == 0) continue;
          const double e0 = d0 + p * a[0] + q * a[3] + s * a[6];
          const double e1 = d1 + p * a[1] + q * a[4] + s * a[7];
          const double e2 = d2 + p * a[2] + q * a[5] + s * a[8];
          if (e0 * e0 + e1 * e1 + e2 * e2 < rc2) {
            if (rec) rec[w0 + count] = Rec<T>{T(e0), T(e1), T(e2), rec_jbits((int)j, T(0))};
            ++count;
          }
        }
  }
  if (!rec) cnt[i] = count;
}

// ------------------------------------------------------------------ finalize
// One 256-thread block per frame: FP64 sums of site energies and virial
// partials in a fixed order (strided per thread, then a fixed tree).
template <class T>
__global__ __launch_bounds__(256) void k_finalize(int nframes, const int64_t* __restrict__ frame_off,
                                                  const T* __restrict__ site, const T* __restrict__ fw,
                                                  double* __restrict__ energy, double* __restrict__ virial) {
  __shared__ double red[7][256];
  const int f = blockIdx.x, t = threadIdx.x;
  const long b = frame_off[f], e = frame_off[f + 1];
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (long i = b + t; i < e; i += 256) {
    s[0] += (double)site[i];
    for (int q = 0; q < 6; ++q) s[1 + q] += (double)fw[i * 9 + 3 + q];
  }
  for (int q = 0; q < 7; ++q) red[q][t] = s[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 7; ++q) red[q][t] += red[q][t + w];
    __syncthreads();
  }
  if (t == 0) {
    energy[f] = red[0][0];
    for (int q = 0; q < 6; ++q) virial[6 * f + q] = red[1 + q][0];
  }
}

template <class T>
__global__ void k_atom_out(int natoms, const T* __restrict__ site, const T* __restrict__ fw,
                           double* __restrict__ force, double* __restrict__ site_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  force[3 * i] = (double)fw[9 * (size_t)i];
  force[3 * i + 1] = (double)fw[9 * (size_t)i + 1];
  force[3 * i + 2] = (double)fw[9 * (size_t)i + 2];
  if (site_out) site_out[i] = (double)site[i];
}

// ------------------------------------------------------------------ driver
template <class T>
void build_neighbors(Workspace& ws, const Batch& b, double rc, hipStream_t st) {
  ws.ensure_atoms(b.natoms);
  const int nb = (b.natoms + 255) / 256;
  SUS2_LAUNCH(ws, st, std::string("neighbors_count<") + tname<T>() + ">", k_neighbors<T>,
              dim3(nb), dim3(256), 0, b.natoms, b.nframes, b.frame_off, b.pos, b.cell, b.periodic, rc, nullptr,
              ws.nbr_cnt.get<int>(), nullptr);
  const long total = exclusive_scan(ws, b.natoms, st);
  if (total > INT32_MAX - 1) throw Error(SUS2_E_RANGE, "neighbour list exceeds 2^31 edges");
  ws.nedges = total;
  ws.rec.ensure(sizeof(Rec<T>) * std::max<long>(total, 1));
  SUS2_LAUNCH(ws, st, std::string("neighbors_fill<") + tname<T>() + ">", k_neighbors<T>,
              dim3(nb), dim3(256), 0, b.natoms, b.nframes, b.frame_off, b.pos, b.cell, b.periodic, rc,
              ws.nbr_off.get<int>(), nullptr, ws.rec.get<Rec<T>>());
}

template <class T>
void finish_efs(Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st) {
  const int N = b.natoms;
  if (N > 0)
    SUS2_LAUNCH(ws, st, "atom_out", k_atom_out<T>, dim3((N + 255) / 256), dim3(256), 0, N, ws.site.get<T>(),
                ws.fw.get<T>(), out.force, out.site);
  SUS2_LAUNCH(ws, st, "finalize", k_finalize<T>, dim3(b.nframes), dim3(256), 0, b.nframes, b.frame_off,
              ws.site.get<T>(), ws.fw.get<T>(), out.energy, out.virial);
}

template void build_neighbors<float>(Workspace&, const Batch&, double, hipStream_t);
template void build_neighbors<double>(Workspace&, const Batch&, double, hipStream_t);
template void finish_efs<float>(Workspace&, const Batch&, const EfsOut&, hipStream_t);
template void finish_efs<double>(Workspace&, const Batch&, const EfsOut&, hipStream_t);

#define SUS2_DECLARE(LL, KK)                                                                                \
  extern template void run_efs_shape<float, LL, KK>(const DevModelSet&, Workspace&, const Batch&,          \
                                                    const EfsOut&, hipStream_t);                           \
  extern template void run_efs_shape<double, LL, KK>(const DevModelSet&, Workspace&, const Batch&,         \
                                                     const EfsOut&, hipStream_t);
SUS2_FOR_EACH_SHAPE(SUS2_DECLARE)
#undef SUS2_DECLARE

template <class T>
void run_efs(const DevModelSet& dm, Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st) {
  const int L = dm.host->L, K = dm.host->K;
  if (dm.host->B > BMAX) throw Error(SUS2_E_UNSUPPORTED, "radial basis size above 16");
#define SUS2_SHAPE(LL, KK) \
  if (L == LL && K == KK) return run_efs_shape<T, LL, KK>(dm, ws, b, out, st);
  SUS2_FOR_EACH_SHAPE(SUS2_SHAPE)
#undef SUS2_SHAPE
  throw Error(SUS2_E_UNSUPPORTED, "no kernel instance for L=" + std::to_string(L) + " K=" + std::to_string(K));
}

void efs_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const EfsOut& out,
                  hipStream_t st) {
  if (precision == SUS2_FP64) run_efs<double>(dm, ws, b, out, st);
  else if (precision == SUS2_FP32) run_efs<float>(dm, ws, b, out, st);
  else throw Error(SUS2_E_INVALID, "precision must be SUS2_FP64 or SUS2_FP32");
  ws.prof.collect_ready();
}

}  // namespace sus2
