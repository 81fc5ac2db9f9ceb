// Shape-templated EFS kernels and their per-shape driver; instantiated per
// (L, K) in sus2_efs_l*k*.hip so the shapes compile in parallel.
//
// Two kinds of kernels:
//   edge kernels   one 64-lane workgroup (one wavefront) per centre atom,
//                  lanes over that atom's neighbour edges;
//   graph kernels  one lane per atom: the K->rank contraction and the product
//                  graph (same schedule for every atom, so schedule reads are
//                  wave-uniform scalar loads) over node-major arrays
//                  V[node][atom] that are coalesced across the wavefront.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/sus2amd.h"
#include "sus2_common.hpp"

namespace sus2 {

template <class T, int LL, int KK>
struct Shape {
  static constexpr int L1 = LL + 1, H = L1 * L1, F = KK * L1, KH = KK * H, NM = (KH + WAVE - 1) / WAVE;
  static constexpr int LR = WAVE + 1;  // padded LDS staging row
};

__device__ __forceinline__ int l_of(int h) {
  int l = 0;
  while ((l + 1) * (l + 1) <= h) ++l;
  return l;
}

// ------------------------------------------------------------------ edge: moments
// Basic moments of one pass for centre atom i (workgroup i):
//   mom[(h,k)][i] = sum_j w_mu(t_i, t_j, h_j) R_mu(r_ij) Y_h(d_ij),  mu = (k, l(h)).
// Edge phase: lane = edge, stages w R and Y in LDS; moment phase: lane =
// moment, sums the staged edges in edge order.
template <class T, int LL, int KK, bool GATE>
__global__ __launch_bounds__(WAVE) void k_edge_moments(DModel<T> m, int N, const int32_t* __restrict__ types,
                                                       const int* __restrict__ nbr_off,
                                                       const Rec<T>* __restrict__ rec, const T* __restrict__ h,
                                                       T* __restrict__ mom) {
  using S = Shape<T, LL, KK>;
  constexpr int L1 = S::L1, H = S::H, F = S::F, KH = S::KH, NM = S::NM, LR = S::LR;
  __shared__ T sR[F * LR];
  __shared__ T sY[H * LR];
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
        const T* pr = sR + (k * L1 + l_of(hh)) * LR;
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
    if (c < KH) mom[(size_t)c * N + i] = acc[t];
  }
}

// ------------------------------------------------------------------ graph: forward
// lane = atom.  Node values of the pass live in V[(node - lo)][N].
template <class T, int LL, int KK>
__device__ __forceinline__ void graph_forward_lane(const DModel<T>& m, const DPass<T>& p, bool gate, int N, int i,
                                                   const T* __restrict__ mom, T* __restrict__ V) {
  using S = Shape<T, LL, KK>;
  constexpr int H = S::H;
  const int lo = p.node_lo, cnt = p.node_cnt;
  // basic nodes that the pass reads directly (unranked graphs)
  for (int c = 0; c < KK * H; ++c) {
    const int node = m.basic_node[c];
    if (node >= lo && node < lo + cnt) V[(size_t)(node - lo) * N + i] = mom[(size_t)c * N + i];
  }
  // K -> rank contraction
  for (int a = 0; a < p.rank; ++a)
    for (int hh = 0; hh < H; ++hh) {
      const int l = l_of(hh), mm = hh - l * l - l;
      T s = T(0);
#pragma unroll
      for (int k = 0; k < KK; ++k)
        if (m.basic_node[hh * KK + k] >= 0) s += m.rankw(gate, l, a, k) * mom[(size_t)(hh * KK + k) * N + i];
      V[(size_t)(rank_node(p.begin, p.rank, H, l, mm, a) - lo) * N + i] = s;
    }
  // products in layer order: out = sum_rows c v[a] v[b]
  for (int o = 0; o < p.nfout; ++o) {
    T s = T(0);
    const int r1 = p.f_roff[o + 1];
    for (int r = p.f_roff[o]; r < r1; ++r)
      s += p.f_c[r] * V[(size_t)(p.f_a[r] - lo) * N + i] * V[(size_t)(p.f_b[r] - lo) * N + i];
    V[(size_t)(p.f_out[o] - lo) * N + i] = s;
  }
}

// Reverse: A[n] = seed[n] + sum over consumer rows c A[out] V[partner], nodes in
// reverse-layer order, then the transposed contraction to canonical basic
// adjoints abar[i][(h,k)] (atom-major, read per neighbour by the edge kernels).
template <class T, int LL, int KK>
__device__ __forceinline__ void graph_reverse_lane(const DModel<T>& m, const DPass<T>& p, bool gate, int N, int i,
                                                   const T* __restrict__ V, T* __restrict__ A,
                                                   T* __restrict__ abar) {
  using S = Shape<T, LL, KK>;
  constexpr int H = S::H, KH = S::KH;
  const int lo = p.node_lo, cnt = p.node_cnt;
  for (int k = 0; k < p.nrnode; ++k) {
    const int n = p.r_node[k];
    T s = A[(size_t)(n - lo) * N + i];
    const int r1 = p.r_roff[k + 1];
    for (int r = p.r_roff[k]; r < r1; ++r)
      s += p.r_c[r] * A[(size_t)(p.r_out[r] - lo) * N + i] * V[(size_t)(p.r_partner[r] - lo) * N + i];
    A[(size_t)(n - lo) * N + i] = s;
  }
  T* out = abar + (size_t)i * KH;
  for (int hh = 0; hh < H; ++hh) {
    const int l = l_of(hh), mm = hh - l * l - l;
#pragma unroll
    for (int k = 0; k < KK; ++k) {
      const int node = m.basic_node[hh * KK + k];
      T s = T(0);
      if (node >= 0) {
        if (node >= lo && node < lo + cnt) s = A[(size_t)(node - lo) * N + i];
        for (int a = 0; a < p.rank; ++a)
          s += m.rankw(gate, l, a, k) * A[(size_t)(rank_node(p.begin, p.rank, H, l, mm, a) - lo) * N + i];
      }
      out[hh * KK + k] = s;
    }
  }
}

// gate forward: V_g, then h_i = W S_i
template <class T, int LL, int KK>
__global__ __launch_bounds__(WAVE) void k_graph_gate_fwd(DModel<T> m, int N, const T* __restrict__ mom,
                                                         T* __restrict__ V, T* __restrict__ h) {
  const int i = blockIdx.x * WAVE + threadIdx.x;
  if (i >= N) return;
  graph_forward_lane<T, LL, KK>(m, m.gp, true, N, i, mom, V);
  constexpr int F = Shape<T, LL, KK>::F;
  const int lo = m.gp.node_lo;
  T acc[F];
#pragma unroll
  for (int mu = 0; mu < F; ++mu) acc[mu] = T(0);
  for (int q = 0; q < m.Qg; ++q) {
    const T sq = V[(size_t)(m.gate_nodes[q] - lo) * N + i];
#pragma unroll
    for (int mu = 0; mu < F; ++mu) acc[mu] += m.W(mu, q) * sq;
  }
#pragma unroll
  for (int mu = 0; mu < F; ++mu) h[(size_t)i * F + mu] = acc[mu];
}

// outer forward + readout + reverse
template <class T, int LL, int KK>
__global__ __launch_bounds__(WAVE) void k_graph_outer(DModel<T> m, int N, const int32_t* __restrict__ types,
                                                      const T* __restrict__ mom, T* __restrict__ V,
                                                      T* __restrict__ A, T* __restrict__ site,
                                                      T* __restrict__ abar) {
  const int i = blockIdx.x * WAVE + threadIdx.x;
  if (i >= N) return;
  const DPass<T>& p = m.op;
  graph_forward_lane<T, LL, KK>(m, p, false, N, i, mom, V);
  const int ti = types[i];
  const int lo = p.node_lo;
  T E = m.site0(ti);
  for (int n = 0; n < p.node_cnt; ++n) A[(size_t)n * N + i] = T(0);
  for (int q = 0; q < m.Q; ++q) {
    const int node = m.scalar_map[q] - lo;
    const T w = m.readout(ti, q);
    E += w * V[(size_t)node * N + i];
    A[(size_t)node * N + i] += w;
  }
  site[i] = E;
  graph_reverse_lane<T, LL, KK>(m, p, false, N, i, V, A, abar);
}

// gate reverse: seeds S_bar = W^T h_bar_i, graph reverse over the stored V_g
template <class T, int LL, int KK>
__global__ __launch_bounds__(WAVE) void k_graph_gate_rev(DModel<T> m, int N, const T* __restrict__ hbar,
                                                         const T* __restrict__ V, T* __restrict__ A,
                                                         T* __restrict__ gbar) {
  const int i = blockIdx.x * WAVE + threadIdx.x;
  if (i >= N) return;
  constexpr int F = Shape<T, LL, KK>::F;
  const DPass<T>& p = m.gp;
  const int lo = p.node_lo;
  T hb[F];
#pragma unroll
  for (int mu = 0; mu < F; ++mu) hb[mu] = hbar[(size_t)i * F + mu];
  for (int n = 0; n < p.node_cnt; ++n) A[(size_t)n * N + i] = T(0);
  for (int q = 0; q < m.Qg; ++q) {
    T s = T(0);
#pragma unroll
    for (int mu = 0; mu < F; ++mu) s += hb[mu] * m.W(mu, q);
    A[(size_t)(m.gate_nodes[q] - lo) * N + i] += s;
  }
  graph_reverse_lane<T, LL, KK>(m, p, true, N, i, V, A, gbar);
}

// ------------------------------------------------------------------ edge: forces
// One directed edge d (r, u = d/r) against basic adjoints `a` (KH, (h,k)
// order, scaled by `par` = (-1)^l when the edge is evaluated reversed):
//   sY = sum a dR Y,  g = sum a R grad Y,  and optionally hb_mu += c_mu sum_m a Y.
template <class T, int LL, int KK, bool REV, bool HB, class Adj>
__device__ __forceinline__ void edge_reverse(T ux, T uy, T uz, T inv, const T* Rw, const T* dRw, const T* cr,
                                             Adj&& adj, T& sY, T& gx, T& gy, T& gz, T* hb) {
  constexpr int L1 = LL + 1;
  sh_stream_grad<T, LL>(ux, uy, uz, inv, [&](int l, int mm, T Y, T yx, T yy, T yz) {
    const int hh = hidx(l, mm);
    const T par = (REV && (l & 1)) ? T(-1) : T(1);
#pragma unroll
    for (int k = 0; k < KK; ++k) {
      const int mu = k * L1 + l;
      const T a = adj(hh * KK + k) * par;
      sY += a * dRw[mu] * Y;
      const T g = a * Rw[mu];
      gx += g * yx; gy += g * yy; gz += g * yz;
      if (HB) hb[mu] += cr[mu] * a * Y;
    }
  });
}

// Gather-based backward of one pass for centre i: for every edge i -> j,
//   g_own = dE_i/dd_ij (centre adjoints), g_rev = dE_j/dd_ji (neighbour's),
//   F_i = sum (g_own - g_rev),  W_i = -sum d (x) g_own,
// and for the outer pass of an Effect model the gate adjoint
//   h_bar_i,mu = alpha_{t_i,mu} sum_j tau_{t_j,mu} R_mu(r; t_j,t_i) sum_m a_j Y(d_ji).
template <class T, int LL, int KK, bool GATE>
__global__ __launch_bounds__(WAVE) void k_edge_force(DModel<T> m, const int32_t* __restrict__ types,
                                                     const int* __restrict__ nbr_off,
                                                     const Rec<T>* __restrict__ rec, const T* __restrict__ h,
                                                     const T* __restrict__ adj, T* __restrict__ fw,
                                                     T* __restrict__ hbar) {
  using S = Shape<T, LL, KK>;
  constexpr int F = S::F, KH = S::KH;
  __shared__ T sA[KH];
  __shared__ T x3i[F];
  const int i = blockIdx.x, lane = threadIdx.x;
  const int ti = types[i];
  const bool hgate = !GATE && m.gated;
  for (int c = lane; c < KH; c += WAVE) sA[c] = adj[(size_t)i * KH + c];
  for (int mu = lane; mu < F; mu += WAVE)
    x3i[mu] = GATE ? m.gtype(ti, mu) : m.tau(ti, mu) + (m.gated ? m.alpha(ti, mu) * h[(size_t)i * F + mu] : T(0));
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
    T R[F], dR[F], cr[F];
    // ---- own direction: E_i through d_ij, pair (t_i, t_j)
    radial<T, LL, KK, true>(m, GATE, ti * m.C + tj, r, R, dR);
#pragma unroll
    for (int mu = 0; mu < F; ++mu) {
      T w;
      if (GATE) w = x3i[mu] * m.gtype(tj, mu);
      else {
        T x3j = m.tau(tj, mu);
        if (m.gated) x3j += m.alpha(tj, mu) * h[(size_t)j * F + mu];
        w = m.tau(ti, mu) * x3j;
      }
      R[mu] *= w; dR[mu] *= w;
    }
    T sY = T(0), gx = T(0), gy = T(0), gz = T(0);
    edge_reverse<T, LL, KK, false, false>(ux, uy, uz, inv, R, dR, cr, [&](int c) { return sA[c]; }, sY, gx, gy, gz,
                                          hb);
    const T gx0 = ux * sY + gx, gy0 = uy * sY + gy, gz0 = uz * sY + gz;
    // ---- reverse direction: E_j through d_ji = -d_ij, pair (t_j, t_i)
    radial<T, LL, KK, true>(m, GATE, tj * m.C + ti, r, R, dR);
#pragma unroll
    for (int mu = 0; mu < F; ++mu) {
      T w;
      if (GATE) w = x3i[mu] * m.gtype(tj, mu);
      else {
        const T tauj = m.tau(tj, mu);
        w = tauj * x3i[mu];
        if (hgate) cr[mu] = tauj * m.alpha(ti, mu) * R[mu];
      }
      R[mu] *= w; dR[mu] *= w;
    }
    const T* aj = adj + (size_t)j * KH;
    sY = gx = gy = gz = T(0);
    if (hgate)
      edge_reverse<T, LL, KK, true, true>(ux, uy, uz, inv, R, dR, cr, [&](int c) { return aj[c]; }, sY, gx, gy, gz,
                                          hb);
    else
      edge_reverse<T, LL, KK, true, false>(ux, uy, uz, inv, R, dR, cr, [&](int c) { return aj[c]; }, sY, gx, gy,
                                           gz, hb);
    // g_rev = dE_j/dd_ji: d(r)/d(d_ji) = -u, grad Y(-d) = -(-1)^l grad Y(d)
    const T gx1 = -ux * sY - gx, gy1 = -uy * sY - gy, gz1 = -uz * sY - gz;
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
  for (int mu = 0; mu < F; ++mu) {
    const T s = wave_sum(hb[mu]);
    if (lane == 0) hbar[(size_t)i * F + mu] = s;
  }
}

// ------------------------------------------------------------------ driver
template <class T, int LL, int KK>
std::string kname(const char* base) {
  return std::string(base) + "<" + tname<T>() + ",L" + std::to_string(LL) + "K" + std::to_string(KK) + ">";
}

template <class T, int LL, int KK>
void run_efs_shape(const DevModelSet& dm, Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st) {
  using S = Shape<T, LL, KK>;
  const DModel<T>& m = dm.get<T>();
  const int N = b.natoms;
  build_neighbors<T>(ws, b, dm.host->rc, st);
  const size_t n1 = std::max(1, N);
  ws.h.ensure(sizeof(T) * n1 * S::F);
  ws.hbar.ensure(sizeof(T) * n1 * S::F);
  ws.mom.ensure(sizeof(T) * n1 * S::KH);
  ws.abar.ensure(sizeof(T) * n1 * S::KH);
  ws.gbar.ensure(sizeof(T) * n1 * S::KH);
  ws.vg.ensure(sizeof(T) * n1 * std::max(1, m.gp.node_cnt));
  ws.vo.ensure(sizeof(T) * n1 * std::max(1, m.op.node_cnt));
  ws.ag.ensure(sizeof(T) * n1 * std::max(1, std::max(m.gp.node_cnt, m.op.node_cnt)));
  ws.site.ensure(sizeof(T) * n1);
  ws.fw.ensure(sizeof(T) * n1 * 9);
  const int* off = ws.nbr_off.get<int>();
  const Rec<T>* rec = ws.rec.get<Rec<T>>();
  T* h = ws.h.get<T>();
  const int gl = (N + WAVE - 1) / WAVE;
  if (N > 0) {
    if (m.gated) {
      SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("edge_moments_gate")), (k_edge_moments<T, LL, KK, true>), dim3(N),
                  dim3(WAVE), 0, m, N, b.types, off, rec, h, ws.mom.get<T>());
      SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("graph_gate_fwd")), (k_graph_gate_fwd<T, LL, KK>), dim3(gl), dim3(WAVE),
                  0, m, N, ws.mom.get<T>(), ws.vg.get<T>(), h);
    }
    SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("edge_moments_outer")), (k_edge_moments<T, LL, KK, false>), dim3(N),
                dim3(WAVE), 0, m, N, b.types, off, rec, h, ws.mom.get<T>());
    SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("graph_outer")), (k_graph_outer<T, LL, KK>), dim3(gl), dim3(WAVE), 0, m, N,
                b.types, ws.mom.get<T>(), ws.vo.get<T>(), ws.ag.get<T>(), ws.site.get<T>(), ws.abar.get<T>());
    SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("edge_force_outer")), (k_edge_force<T, LL, KK, false>), dim3(N),
                dim3(WAVE), 0, m, b.types, off, rec, h, ws.abar.get<T>(), ws.fw.get<T>(), ws.hbar.get<T>());
    if (m.gated) {
      SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("graph_gate_rev")), (k_graph_gate_rev<T, LL, KK>), dim3(gl), dim3(WAVE),
                  0, m, N, ws.hbar.get<T>(), ws.vg.get<T>(), ws.ag.get<T>(), ws.gbar.get<T>());
      SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("edge_force_gate")), (k_edge_force<T, LL, KK, true>), dim3(N),
                  dim3(WAVE), 0, m, b.types, off, rec, h, ws.gbar.get<T>(), ws.fw.get<T>(), ws.hbar.get<T>());
    }
  }
  finish_efs<T>(ws, b, out, st);
}

}  // namespace sus2
