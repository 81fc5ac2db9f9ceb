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
#include "sus2_device.hpp"
#include "sus2_model.hpp"
#include "sus2_runtime.hpp"

namespace sus2 {

template <class T>
struct alignas(4 * sizeof(T)) Rec {
  T x, y, z, jb;
};
__device__ __forceinline__ int rec_j(const Rec<float>& r) { return __float_as_int(r.jb); }
__device__ __forceinline__ int rec_j(const Rec<double>& r) { return (int)__double_as_longlong(r.jb); }
__device__ __forceinline__ float rec_jbits(int j, float) { return __int_as_float(j); }
__device__ __forceinline__ double rec_jbits(int j, double) { return __longlong_as_double((long long)j); }

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

// ------------------------------------------------------------------ moments
template <class T, int LL, int KK>
struct Shape {
  static constexpr int L1 = LL + 1, H = L1 * L1, F = KK * L1, KH = KK * H, NM = (KH + WAVE - 1) / WAVE;
  static constexpr int LDSR = WAVE + 1;  // padded staging row
};

template <class T, int LL, int KK>
size_t moments_lds(int M) {
  using S = Shape<T, LL, KK>;
  return sizeof(T) * ((S::F + S::H) * S::LDSR + S::KH + 2 * (size_t)M);
}

template <class T, int LL, int KK, bool GATE>
__global__ __launch_bounds__(WAVE) void k_moments(DModel<T> m, const int32_t* __restrict__ types,
                                                  const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                                  const T* __restrict__ h, T* __restrict__ h_out,
                                                  T* __restrict__ gbasic_out, T* __restrict__ site_out,
                                                  T* __restrict__ abar_out) {
  using S = Shape<T, LL, KK>;
  constexpr int L1 = S::L1, H = S::H, F = S::F, KH = S::KH, NM = S::NM, LR = S::LDSR;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sR = reinterpret_cast<T*>(smem_raw);  // [F][LR]
  T* sY = sR + F * LR;                     // [H][LR]
  T* mom = sY + H * LR;                    // [KH]
  T* v = mom + KH;                         // [M]
  T* a = v + m.M;                          // [M]
  const int i = blockIdx.x, lane = threadIdx.x;
  const int ti = types[i];
  const int e0 = nbr_off[i], e1 = nbr_off[i + 1];
  T acc[NM];
#pragma unroll
  for (int t = 0; t < NM; ++t) acc[t] = T(0);
  for (int base = e0; base < e1; base += WAVE) {
    const int e = base + lane;
    if (e < e1) {
      const Rec<T> q = rec[e];
      const int j = rec_j(q);
      const T r = dsqrt(q.x * q.x + q.y * q.y + q.z * q.z);
      const T inv = T(1) / r;
      const int tj = types[j];
      T Rv[F], Y[H];
      radial<T, LL, KK, false>(m, GATE, ti * m.C + tj, r, Rv, nullptr);
      sh_values<T, LL>(q.x * inv, q.y * inv, q.z * inv, Y);
#pragma unroll
      for (int mu = 0; mu < F; ++mu) {
        T w;
        if (GATE) w = m.gtype(ti, mu) * m.gtype(tj, mu);
        else {
          T x1 = m.tau(tj, mu);
          if (m.gated) x1 += m.alpha(tj, mu) * h[(size_t)j * F + mu];
          w = m.tau(ti, mu) * x1;
        }
        sR[mu * LR + lane] = w * Rv[mu];
      }
#pragma unroll
      for (int hh = 0; hh < H; ++hh) sY[hh * LR + lane] = Y[hh];
    }
    __syncthreads();
    const int cnt = min(WAVE, e1 - base);
#pragma unroll
    for (int t = 0; t < NM; ++t) {
      const int c = lane + WAVE * t;
      if (c < KH) {
        const int hh = c / KK, k = c % KK;
        int l = 0;
        while ((l + 1) * (l + 1) <= hh) ++l;
        const T* pr = sR + (k * L1 + l) * LR;
        const T* py = sY + hh * LR;
        T s = acc[t];
        for (int ee = 0; ee < cnt; ++ee) s += pr[ee] * py[ee];
        acc[t] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < NM; ++t) {
    const int c = lane + WAVE * t;
    if (c < KH) mom[c] = acc[t];
  }
  __syncthreads();
  load_and_contract(m, GATE, mom, v);
  graph_forward(GATE ? m.gp : m.op, v);
  if (GATE) {
    for (int mu = lane; mu < F; mu += WAVE) {
      T s = T(0);
      for (int qq = 0; qq < m.Qg; ++qq) s += m.W(mu, qq) * v[m.gate_nodes[qq]];
      h_out[(size_t)i * F + mu] = s;
    }
    for (int c = lane; c < KH; c += WAVE) gbasic_out[(size_t)i * KH + c] = mom[c];
    return;
  }
  T part = T(0);
  for (int qq = lane; qq < m.Q; qq += WAVE) part += m.readout(ti, qq) * v[m.scalar_map[qq]];
  const T E = wave_sum(part) + m.site0(ti);
  if (lane == 0) site_out[i] = E;
  for (int n = lane; n < m.M; n += WAVE) a[n] = T(0);
  __syncthreads();
  if (lane == 0)
    for (int qq = 0; qq < m.Q; ++qq) a[m.scalar_map[qq]] += m.readout(ti, qq);
  __syncthreads();
  graph_reverse(m.op, v, a);
  adjoint_to_basic(m, false, a, abar_out + (size_t)i * KH);
}

// ------------------------------------------------------------------ forces
template <class T, int LL, int KK>
size_t force_lds(int M) {
  using S = Shape<T, LL, KK>;
  return sizeof(T) * (2 * S::KH + S::F + 2 * (size_t)M);
}

// adj: basic adjoints of this pass for every atom ([N][H][K]); fw: per-atom
// force (3) and virial (6) partials.  The outer pass writes them and, for a
// gated model, gathers h_bar_i and leaves the gate basic adjoints in gbar_out;
// the gate pass adds its part.
template <class T, int LL, int KK, bool GATE>
__global__ __launch_bounds__(WAVE) void k_force(DModel<T> m, const int32_t* __restrict__ types,
                                                const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                                const T* __restrict__ h, const T* __restrict__ adj,
                                                const T* __restrict__ gbasic, T* __restrict__ gbar_out,
                                                T* __restrict__ fw) {
  using S = Shape<T, LL, KK>;
  constexpr int L1 = S::L1, H = S::H, F = S::F, KH = S::KH;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sA = reinterpret_cast<T*>(smem_raw);  // [KH]
  T* mom = sA + KH;                        // [KH]
  T* x3i = mom + KH;                       // [F]
  T* v = x3i + F;                          // [M]
  T* a = v + m.M;                          // [M]
  const int i = blockIdx.x, lane = threadIdx.x;
  const int ti = types[i];
  const bool hgate = !GATE && m.gated;
  for (int c = lane; c < KH; c += WAVE) sA[c] = adj[(size_t)i * KH + c];
  for (int mu = lane; mu < F; mu += WAVE) {
    if (GATE) x3i[mu] = m.gtype(ti, mu);
    else x3i[mu] = m.tau(ti, mu) + (m.gated ? m.alpha(ti, mu) * h[(size_t)i * F + mu] : T(0));
  }
  __syncthreads();
  T fx = T(0), fy = T(0), fz = T(0);
  T w0 = T(0), w1 = T(0), w2 = T(0), w3 = T(0), w4 = T(0), w5 = T(0);
  T hb[F];
#pragma unroll
  for (int mu = 0; mu < F; ++mu) hb[mu] = T(0);
  const int e0 = nbr_off[i], e1 = nbr_off[i + 1];
  for (int e = e0 + lane; e < e1; e += WAVE) {
    const Rec<T> q = rec[e];
    const int j = rec_j(q);
    const T dx = q.x, dy = q.y, dz = q.z;
    const T r = dsqrt(dx * dx + dy * dy + dz * dz);
    const T inv = T(1) / r;
    const T ux = dx * inv, uy = dy * inv, uz = dz * inv;
    const int tj = types[j];
    T Ro[F], dRo[F], Rr[F], dRr[F], cr[F];
    radial<T, LL, KK, true>(m, GATE, ti * m.C + tj, r, Ro, dRo);
    if (ti == tj) {
#pragma unroll
      for (int mu = 0; mu < F; ++mu) { Rr[mu] = Ro[mu]; dRr[mu] = dRo[mu]; }
    } else {
      radial<T, LL, KK, true>(m, GATE, tj * m.C + ti, r, Rr, dRr);
    }
#pragma unroll
    for (int mu = 0; mu < F; ++mu) {
      T wo, wr;
      if (GATE) { wo = x3i[mu] * m.gtype(tj, mu); wr = wo; }
      else {
        const T tauj = m.tau(tj, mu);
        T x3j = tauj;
        if (m.gated) x3j += m.alpha(tj, mu) * h[(size_t)j * F + mu];
        wo = m.tau(ti, mu) * x3j;
        wr = tauj * x3i[mu];
        cr[mu] = hgate ? tauj * m.alpha(ti, mu) * Rr[mu] : T(0);
      }
      Ro[mu] *= wo; dRo[mu] *= wo; Rr[mu] *= wr; dRr[mu] *= wr;
    }
    const T* aj = adj + (size_t)j * KH;
    T sYo = T(0), gox = T(0), goy = T(0), goz = T(0);
    T sYr = T(0), grx = T(0), gry = T(0), grz = T(0);
    sh_stream_grad<T, LL>(ux, uy, uz, inv, [&](int l, int mm, T Y, T gx, T gy, T gz) {
      const int hh = hidx(l, mm);
      const T par = (l & 1) ? T(-1) : T(1);
#pragma unroll
      for (int k = 0; k < KK; ++k) {
        const int mu = k * L1 + l;
        const T ao = sA[hh * KK + k];
        const T ar = aj[hh * KK + k] * par;
        sYo += ao * dRo[mu] * Y;
        const T go = ao * Ro[mu];
        gox += go * gx; goy += go * gy; goz += go * gz;
        sYr += ar * dRr[mu] * Y;
        const T gr = ar * Rr[mu];
        grx -= gr * gx; gry -= gr * gy; grz -= gr * gz;
        if (hgate) hb[mu] += cr[mu] * ar * Y;
      }
    });
    // g_own = dE_i/dd_ij ; g_rev = dE_j/dd_ji with d_ji = -d_ij
    const T gx0 = ux * sYo + gox, gy0 = uy * sYo + goy, gz0 = uz * sYo + goz;
    const T gx1 = -ux * sYr + grx, gy1 = -uy * sYr + gry, gz1 = -uz * sYr + grz;
    fx += gx0 - gx1; fy += gy0 - gy1; fz += gz0 - gz1;
    w0 -= dx * gx0; w1 -= dy * gy0; w2 -= dz * gz0;
    w3 -= dx * gy0; w4 -= dx * gz0; w5 -= dy * gz0;
  }
  fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
  w0 = wave_sum(w0); w1 = wave_sum(w1); w2 = wave_sum(w2);
  w3 = wave_sum(w3); w4 = wave_sum(w4); w5 = wave_sum(w5);
  if (lane == 0) {
    T* o = fw + (size_t)i * 9;
    if (GATE) {
      o[0] += fx; o[1] += fy; o[2] += fz;
      o[3] += w0; o[4] += w1; o[5] += w2; o[6] += w3; o[7] += w4; o[8] += w5;
    } else {
      o[0] = fx; o[1] = fy; o[2] = fz;
      o[3] = w0; o[4] = w1; o[5] = w2; o[6] = w3; o[7] = w4; o[8] = w5;
    }
  }
  if (!hgate) return;
#pragma unroll
  for (int mu = 0; mu < F; ++mu) hb[mu] = wave_sum(hb[mu]);
  // gate reverse of atom i: node values from its stored gate moments
  for (int c = lane; c < KH; c += WAVE) mom[c] = gbasic[(size_t)i * KH + c];
  __syncthreads();
  load_and_contract(m, true, mom, v);
  graph_forward(m.gp, v);
  for (int n = lane; n < m.M; n += WAVE) a[n] = T(0);
  __syncthreads();
  for (int qq = lane; qq < m.Qg; qq += WAVE) {
    T s = T(0);
#pragma unroll
    for (int mu = 0; mu < F; ++mu) s += hb[mu] * m.W(mu, qq);
    a[m.gate_nodes[qq]] = s;
  }
  __syncthreads();
  graph_reverse(m.gp, v, a);
  adjoint_to_basic(m, true, a, gbar_out + (size_t)i * KH);
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
#define SUS2_CHECK(x)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw Error(SUS2_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
void build_neighbors(Workspace& ws, const Batch& b, double rc, hipStream_t st) {
  ws.ensure_atoms(b.natoms);
  const int nb = (b.natoms + 255) / 256;
  hipLaunchKernelGGL(k_neighbors<T>, dim3(nb), dim3(256), 0, st, b.natoms, b.nframes, b.frame_off, b.pos, b.cell,
                     b.periodic, rc, nullptr, ws.nbr_cnt.get<int>(), nullptr);
  SUS2_CHECK(hipGetLastError());
  const long total = exclusive_scan(ws, b.natoms, st);
  if (total > INT32_MAX - 1) throw Error(SUS2_E_RANGE, "neighbour list exceeds 2^31 edges");
  ws.nedges = total;
  ws.rec.ensure(sizeof(Rec<T>) * std::max<long>(total, 1));
  hipLaunchKernelGGL(k_neighbors<T>, dim3(nb), dim3(256), 0, st, b.natoms, b.nframes, b.frame_off, b.pos, b.cell,
                     b.periodic, rc, ws.nbr_off.get<int>(), nullptr, ws.rec.get<Rec<T>>());
  SUS2_CHECK(hipGetLastError());
}

template <class T, int LL, int KK>
void run_efs_shape(const DevModelSet& dm, Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st) {
  using S = Shape<T, LL, KK>;
  const DModel<T>& m = dm.get<T>();
  const int N = b.natoms;
  build_neighbors<T>(ws, b, dm.host->rc, st);
  ws.h.ensure(sizeof(T) * (size_t)std::max(1, N) * S::F);
  ws.gbasic.ensure(sizeof(T) * (size_t)std::max(1, N) * S::KH);
  ws.abar.ensure(sizeof(T) * (size_t)std::max(1, N) * S::KH);
  ws.gbar.ensure(sizeof(T) * (size_t)std::max(1, N) * S::KH);
  ws.site.ensure(sizeof(T) * (size_t)std::max(1, N));
  ws.fw.ensure(sizeof(T) * (size_t)std::max(1, N) * 9);
  const int* off = ws.nbr_off.get<int>();
  const Rec<T>* rec = ws.rec.get<Rec<T>>();
  T* h = ws.h.get<T>();
  if (N > 0) {
    const size_t lm = moments_lds<T, LL, KK>(m.M), lf = force_lds<T, LL, KK>(m.M);
    if (m.gated) {
      hipLaunchKernelGGL((k_moments<T, LL, KK, true>), dim3(N), dim3(WAVE), lm, st, m, b.types, off, rec, h, h,
                         ws.gbasic.get<T>(), ws.site.get<T>(), ws.abar.get<T>());
      SUS2_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL((k_moments<T, LL, KK, false>), dim3(N), dim3(WAVE), lm, st, m, b.types, off, rec, h, h,
                       ws.gbasic.get<T>(), ws.site.get<T>(), ws.abar.get<T>());
    SUS2_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_force<T, LL, KK, false>), dim3(N), dim3(WAVE), lf, st, m, b.types, off, rec, h,
                       ws.abar.get<T>(), ws.gbasic.get<T>(), ws.gbar.get<T>(), ws.fw.get<T>());
    SUS2_CHECK(hipGetLastError());
    if (m.gated) {
      hipLaunchKernelGGL((k_force<T, LL, KK, true>), dim3(N), dim3(WAVE), lf, st, m, b.types, off, rec, h,
                         ws.gbar.get<T>(), ws.gbasic.get<T>(), ws.gbar.get<T>(), ws.fw.get<T>());
      SUS2_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_atom_out<T>, dim3((N + 255) / 256), dim3(256), 0, st, N, ws.site.get<T>(), ws.fw.get<T>(),
                       out.force, out.site);
    SUS2_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finalize<T>, dim3(b.nframes), dim3(256), 0, st, b.nframes, b.frame_off, ws.site.get<T>(),
                     ws.fw.get<T>(), out.energy, out.virial);
  SUS2_CHECK(hipGetLastError());
}

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
}

}  // namespace sus2
