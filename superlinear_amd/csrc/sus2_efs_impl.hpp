// Shape-templated EFS kernels and their per-shape driver; instantiated per
// (L, K) in sus2_efs_l*k*.hip so the shapes compile in parallel.
//
// Every kernel here runs one 64-lane workgroup (one wavefront) per centre
// atom:
//   k_pass_forward<gate>  edge phase (lane = edge: radial functions and
//                         harmonics staged in LDS), moment phase (lane = basic
//                         moment, sum over the staged edges in edge order),
//                         then the K->rank contraction and the product graph
//                         with product rows spread over the lanes.  gate:
//                         h_i = W S_i and the gate node values for the reverse;
//                         outer: site energy and the basic-moment adjoints.
//   k_edge_backward<gate> lane = edge: g_e = dE_i/dd_e of the centre's own
//                         energy for each directed edge; the centre's sum over
//                         its outgoing edges and its virial are reduced in the
//                         workgroup, g_e is stored at the edge's slot in its
//                         target's incoming list, and (outer pass of an
//                         Effect model) so is the gate-adjoint message u_e of
//                         the neighbour's gate signal.
//   k_gather<gate>        F_i -= sum_{in edges} g_e' over the contiguous
//                         incoming slots (no atomics), and for the outer pass
//                         of an Effect model h_bar_i followed by the gate graph
//                         reverse to the gate basic adjoints.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/sus2amd.h"
#include "sus2_common.hpp"

namespace sus2 {

constexpr int LR = WAVE + 1;  // padded LDS staging row

__host__ __device__ __forceinline__ int l_of(int h) {
  int l = 0;
  while ((l + 1) * (l + 1) <= h) ++l;
  return l;
}

// ------------------------------------------------------------------ staged sums
// sum_{e < cnt} a[e] b[e] over staged LDS rows (four independent partial sums
// in a fixed pattern, so that the row loads overlap).
template <class T>
__device__ __forceinline__ T lds_dot(const T* __restrict__ a, const T* __restrict__ b, int cnt) {
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
  int e = 0;
  for (; e + 4 <= cnt; e += 4) {
    s0 += a[e] * b[e];
    s1 += a[e + 1] * b[e + 1];
    s2 += a[e + 2] * b[e + 2];
    s3 += a[e + 3] * b[e + 3];
  }
  for (; e < cnt; ++e) s0 += a[e] * b[e];
  return (s0 + s1) + (s2 + s3);
}

// sum_{e < cnt, key[e] == t} v[e] in FP64 over staged LDS rows.
template <class T>
__device__ __forceinline__ double lds_sum_where(const T* __restrict__ v, const T* __restrict__ key, T t, int cnt) {
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int e = 0;
  for (; e + 4 <= cnt; e += 4) {
    s0 += key[e] == t ? (double)v[e] : 0.0;
    s1 += key[e + 1] == t ? (double)v[e + 1] : 0.0;
    s2 += key[e + 2] == t ? (double)v[e + 2] : 0.0;
    s3 += key[e + 3] == t ? (double)v[e + 3] : 0.0;
  }
  for (; e < cnt; ++e) s0 += key[e] == t ? (double)v[e] : 0.0;
  return (s0 + s1) + (s2 + s3);
}

// ------------------------------------------------------------------ wave graph walks
// Basic nodes take the moments; rank nodes the K -> rank contraction
// (sus2_sl_kernel.cpp:104-105, 217-221).  v: LDS window of nc node values
// starting at node lo (node n at v[n - lo]; lo = 0, nc = M for all nodes);
// basic nodes outside the window are not used by the pass and are skipped.
template <class T, int KK>
__device__ void wave_contract(const DModel<T>& m, const DPass<T>& p, bool gate, const T* mom, T* v, int lo = 0,
                              int nc = 1 << 30) {
  const int lane = threadIdx.x, H = m.H, KH = KK * H;
  for (int c = lane; c < KH; c += WAVE) {
    const int node = m.basic_node[c] - lo;
    if (node >= 0 && node < nc) v[node] = mom[c];
  }
  const int nr = p.rank * H;
  for (int idx = lane; idx < nr; idx += WAVE) {
    const int a = idx / H, hh = idx % H;
    const int l = l_of(hh), mm = hh - l * l - l;
    T s = T(0);
#pragma unroll
    for (int k = 0; k < KK; ++k)
      if (m.basic_node[hh * KK + k] >= 0) s += m.rankw(gate, l, a, k) * mom[hh * KK + k];
    v[rank_node(p.begin, p.rank, H, l, mm, a) - lo] = s;
  }
  __syncthreads();
}

// Row loop of a graph layer: fn(r) for r = r0 + lane, r0 + lane + WAVE, ...
// below r1, four rows at a time so that their schedule loads are in flight
// together (the row data are the same for every atom, L2-resident).
template <class Fn>
__device__ __forceinline__ void wave_rows(int r0, int r1, Fn&& fn) {
  int r = r0 + (int)threadIdx.x;
  for (; r + 3 * WAVE < r1; r += 4 * WAVE) {
    fn(r);
    fn(r + WAVE);
    fn(r + 2 * WAVE);
    fn(r + 3 * WAVE);
  }
  for (; r < r1; r += WAVE) fn(r);
}

// Product layers with rows spread over the lanes, then per-output sums of the
// staged row products in serialised row order (sus2_sl_kernel.cpp:308-313).
template <class T>
__device__ void wave_graph_forward(const DPass<T>& p, T* v, T* scratch, int lo = 0) {
  const int lane = threadIdx.x;
  for (int l = 0; l < p.nfl; ++l) {
    const int4 bd = p.fl_bnd[l];
    const int o0 = bd.x, o1 = bd.y, r0 = bd.z, r1 = bd.w;
    // this lane's first output of the layer, loaded before the row products
    const int oa = o0 + lane;
    int ob = 0, oe = 0, on = 0;
    if (oa < o1) {
      ob = p.f_roff[oa];
      oe = p.f_roff[oa + 1];
      on = p.f_out[oa] - lo;
    }
    wave_rows(r0, r1, [&](int r) {
      const GRow<T> g = p.f_row[r];
      scratch[r - r0] = g.c * v[g.x - lo] * v[g.y - lo];
    });
    __syncthreads();
    if (oa < o1) {
      T s = T(0);
      for (int r = ob - r0; r < oe - r0; ++r) s += scratch[r];
      v[on] = s;
    }
    for (int o = oa + WAVE; o < o1; o += WAVE) {
      T s = T(0);
      const int e = p.f_roff[o + 1] - r0;
      for (int r = p.f_roff[o] - r0; r < e; ++r) s += scratch[r];
      v[p.f_out[o] - lo] = s;
    }
    __syncthreads();
  }
}

// Reverse: a[n] += sum over consumer rows c a[out] v[partner], nodes in
// reverse-layer order; a holds the seeds on entry.
template <class T>
__device__ void wave_graph_reverse(const DPass<T>& p, const T* v, T* a, T* scratch, int lo = 0) {
  const int lane = threadIdx.x;
  for (int l = 0; l < p.nrl; ++l) {
    const int4 bd = p.rl_bnd[l];
    const int n0 = bd.x, n1 = bd.y, q0 = bd.z, q1 = bd.w;
    // this lane's first node of the layer, loaded before the row products
    const int ka = n0 + lane;
    int kb = 0, ke = 0, kn = 0;
    if (ka < n1) {
      kb = p.r_roff[ka];
      ke = p.r_roff[ka + 1];
      kn = p.r_node[ka] - lo;
    }
    wave_rows(q0, q1, [&](int q) {
      const GRow<T> g = p.r_row[q];
      scratch[q - q0] = g.c * a[g.x - lo] * v[g.y - lo];
    });
    __syncthreads();
    if (ka < n1) {
      T s = a[kn];
      for (int q = kb - q0; q < ke - q0; ++q) s += scratch[q];
      a[kn] = s;
    }
    for (int k = ka + WAVE; k < n1; k += WAVE) {
      const int n = p.r_node[k] - lo;
      T s = a[n];
      const int e = p.r_roff[k + 1] - q0;
      for (int q = p.r_roff[k] - q0; q < e; ++q) s += scratch[q];
      a[n] = s;
    }
    __syncthreads();
  }
}

// Transposed contraction to canonical basic adjoints out[(h,k)].
template <class T, int KK>
__device__ void wave_adjoint_to_basic(const DModel<T>& m, const DPass<T>& p, bool gate, const T* a, T* out,
                                      int lo = 0, int nc = 1 << 30) {
  const int lane = threadIdx.x, H = m.H, KH = KK * H;
  for (int c = lane; c < KH; c += WAVE) {
    const int hh = c / KK, k = c % KK;
    const int l = l_of(hh), mm = hh - l * l - l;
    const int node = m.basic_node[c];
    T s = T(0);
    if (node >= 0) {
      // (a basic node outside the window has no adjoint in this pass)
      if (node - lo >= 0 && node - lo < nc) s = a[node - lo];
      for (int aa = 0; aa < p.rank; ++aa)
        s += m.rankw(gate, l, aa, k) * a[rank_node(p.begin, p.rank, H, l, mm, aa) - lo];
    }
    out[c] = s;
  }
}

template <class T>
__host__ __device__ inline int graph_scratch(const DPass<T>& p) {
  return p.max_fw_rows > p.max_rv_rows ? p.max_fw_rows : p.max_rv_rows;
}

// ------------------------------------------------------------------ pass forward
// LDS of k_pass_forward: the species tables, then one region used first for
// the edge staging [F + H][row] and afterwards (the staging is dead once the
// moments are in registers) for moments [KH], the pass's node values and
// adjoints [2 NC] (NC = the larger pass's contiguous node range) and the graph
// scratch -- the overlay keeps the footprint small enough for every workgroup
// of a 4096-atom frame to be resident at once.  Wide shapes (F > 16) stage a
// batch of 64 edges in two halves of 32 rows (the second half's radial values
// wait in registers, its harmonics are formed again), halving the staging.
template <class Cf>
struct FwdStage {
  static constexpr bool HALF = Cf::F > 16;
  // radial coefficient rows staged in LDS (narrow shapes); wide shapes read
  // them from the coefficient vector: the LDS rows and the registers their
  // reads hold would cost these passes a workgroup per SIMD
  static constexpr bool CL = !HALF;
  static constexpr int ROW = HALF ? WAVE / 2 + 1 : LR;  // padded staging row
  static constexpr int COLS = HALF ? WAVE / 2 : WAVE;   // edges per staged batch
};

template <class Cf>
__host__ __device__ inline size_t forward_region(int nc, int sc) {
  const size_t stage = (size_t)(Cf::F + Cf::H) * FwdStage<Cf>::ROW;
  const size_t graph = (size_t)Cf::KH + 2 * (size_t)nc + (size_t)(sc > 1 ? sc : 1);
  return stage > graph ? stage : graph;
}

// LDS node windows: the gate pass's range, and the larger of the two passes'.
template <class T>
__host__ __device__ inline int gate_nodes(const DModel<T>& m) {
  return m.gp.node_cnt > 1 ? m.gp.node_cnt : 1;
}

template <class T>
__host__ __device__ inline int pass_nodes(const DModel<T>& m) {
  const int nc = m.gp.node_cnt > m.op.node_cnt ? m.gp.node_cnt : m.op.node_cnt;
  return nc > 1 ? nc : 1;
}

template <class Cf>
size_t forward_lds(const DModel<typename Cf::T>& m) {
  // the scratch also holds the gate scalars S [Qg] of the gate pass
  const int sc = std::max({graph_scratch(m.gp), graph_scratch(m.op), m.Qg});
  return sizeof(typename Cf::T) *
         (species_tables_size<Cf, FwdStage<Cf>::CL>(m.C) + forward_region<Cf>(pass_nodes(m), sc));
}

// Moments of the staged columns [c0, c1) of a batch: lane = basic moment (k, h).
template <class Cf, class T = typename Cf::T>
__device__ __forceinline__ void stage_moments(const T* sR, const T* sY, int c0, int c1, T* acc) {
  constexpr int KK = Cf::KK, L1 = Cf::L1, KH = Cf::KH, NM = Cf::NM, ROW = FwdStage<Cf>::ROW;
  const int lane = threadIdx.x;
#pragma unroll
  for (int t = 0; t < NM; ++t) {
    const int c = lane + WAVE * t;
    if (c < KH) {
      const int hh = c / KK, k = c % KK;
      acc[t] += lds_dot(sR + (k * L1 + l_of(hh)) * ROW + c0, sY + hh * ROW + c0, c1 - c0);
    }
  }
}

// Moments of each of the NA centres from the staged columns [0, cnt) that
// hold edges [first, first + cnt): centre a owns edges [eb[a], eb[a + 1]).
template <class Cf, int NA, class T = typename Cf::T>
__device__ __forceinline__ void stage_moments_split(const T* sR, const T* sY, int first, int cnt, const int* eb,
                                                    T (&acc)[NA][Cf::NM]) {
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int c0 = max(0, min(cnt, eb[a] - first)), c1 = max(0, min(cnt, eb[a + 1] - first));
    if (c1 > c0) stage_moments<Cf>(sR, sY, c0, c1, acc[a]);
  }
}

// Edge phase of a forward pass for NA consecutive centres i0 .. i0 + na - 1
// (their edge lists are contiguous, so the staged batches of two centres'
// ~90 edges each fill the lanes better than one centre's): the moments of
// centre a into acc[a] (the workgroup's LDS: species tables already loaded
// into tab, then `region`).
template <class Cf, bool GATE, int NA, class T = typename Cf::T>
__device__ __forceinline__ void forward_edges(const DModel<T>& m, const SpeciesTables<T>& tab, T* region, int i0,
                                              int na, const int32_t* __restrict__ types,
                                              const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                              const T* __restrict__ h, T (&acc)[NA][Cf::NM]) {
  constexpr int LL = Cf::LL, H = Cf::H, F = Cf::F, NM = Cf::NM;
  constexpr bool HALF = FwdStage<Cf>::HALF;
  constexpr int ROW = FwdStage<Cf>::ROW, HW = WAVE / 2;
  T* sR = region;        // [F][ROW]
  T* sY = sR + F * ROW;  // [H][ROW]
  const int lane = threadIdx.x;
  int eb[NA + 1], ta[NA];
#pragma unroll
  for (int a = 0; a <= NA; ++a) eb[a] = nbr_off[i0 + min(a, na)];
#pragma unroll
  for (int a = 0; a < NA; ++a) ta[a] = types[i0 + min(a, na - 1)];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int t = 0; t < NM; ++t) acc[a][t] = T(0);
  const int e0 = eb[0], e1 = eb[NA];
  // the next batch's record loaded while this one is worked on
  Rec<T> qn = e0 + lane < e1 ? rec[e0 + lane] : Rec<T>{T(1), T(0), T(0), T(0)};
  for (int base = e0; base < e1; base += WAVE) {
    const int e = base + lane;
    const Rec<T> q = qn;
    if (e + WAVE < e1) qn = rec[e + WAVE];
    // the centre of the lane's edge
    int ti = ta[0];
#pragma unroll
    for (int a = 1; a < NA; ++a)
      if (e >= eb[a]) ti = ta[a];
    if constexpr (HALF) {
      // wide shapes: lanes [0, 32) stage their edges at once, lanes [32, 64)
      // keep their weighted radial values and stage them (and their
      // harmonics, formed again) after the first half's moments
      const bool first = lane < HW;
      const int col = lane & (HW - 1);
      T keep[F];
      T ux = T(0), uy = T(0), uz = T(0);
      if (e < e1) {
        const int j = rec_j(q);
        T r, inv;
        edge_length(q.x * q.x + q.y * q.y + q.z * q.z, r, inv);
        const int tj = rec_t(q);
        const T* x0i = tab.x0 + ti * F;
        const T* x0j = tab.x0 + tj * F;
        const T* alj = tab.alpha + tj * F;
        T hj[F];
        if (!GATE && m.gated) load_row<T, F>(h + (size_t)j * F, hj);  // issued before the basis work
        radial_each<Cf, false, 0, Cf::NBLK, FwdStage<Cf>::CL>(m, GATE, tab, ti * m.C + tj, r, [&](int mu, T R, T) {
          T x1 = x0j[mu];
          if (!GATE && m.gated) x1 += alj[mu] * hj[mu];
          keep[mu] = x0i[mu] * x1 * R;
          if (first) sR[mu * ROW + col] = keep[mu];
        });
        ux = q.x * inv;
        uy = q.y * inv;
        uz = q.z * inv;
        if (first) sh_stream_values<T, LL>(ux, uy, uz, [&](int l, int mm, T Y) { sY[hidx(l, mm) * ROW + col] = Y; });
      }
      __syncthreads();
      stage_moments_split<Cf, NA>(sR, sY, base, min(HW, e1 - base), eb, acc);
      __syncthreads();
      const int cnt2 = min(HW, e1 - base - HW);
      if (cnt2 > 0) {
        if (!first && e < e1) {
#pragma unroll
          for (int mu = 0; mu < F; ++mu) sR[mu * ROW + col] = keep[mu];
          sh_stream_values<T, LL>(ux, uy, uz, [&](int l, int mm, T Y) { sY[hidx(l, mm) * ROW + col] = Y; });
        }
        __syncthreads();
        stage_moments_split<Cf, NA>(sR, sY, base + HW, cnt2, eb, acc);
        __syncthreads();
      }
    } else {
      if (e < e1) {
        const int j = rec_j(q);
        T r, inv;
        edge_length(q.x * q.x + q.y * q.y + q.z * q.z, r, inv);
        const int tj = rec_t(q);
        const T* x0i = tab.x0 + ti * F;
        const T* x0j = tab.x0 + tj * F;
        const T* alj = tab.alpha + tj * F;
        T Rv[F], Y[H], hj[F];
        if (!GATE && m.gated) load_row<T, F>(h + (size_t)j * F, hj);  // issued before the basis work
        radial<Cf, false, FwdStage<Cf>::CL>(m, GATE, tab, ti * m.C + tj, r, Rv, nullptr);
        sh_values<T, LL>(q.x * inv, q.y * inv, q.z * inv, Y);
#pragma unroll
        for (int mu = 0; mu < F; ++mu) {
          T x1 = x0j[mu];
          if (!GATE && m.gated) x1 += alj[mu] * hj[mu];
          sR[mu * ROW + lane] = x0i[mu] * x1 * Rv[mu];
        }
#pragma unroll
        for (int hh = 0; hh < H; ++hh) sY[hh * ROW + lane] = Y[hh];
      }
      __syncthreads();
      stage_moments_split<Cf, NA>(sR, sY, base, min(WAVE, e1 - base), eb, acc);
      __syncthreads();
    }
  }
  // (the edge loop ends with a barrier, so the staging region is free after it)
}

// Graph phase of a forward pass for centre i of species ti from its moments
// acc: rank contraction, product graph, then the gate signals h_i and node
// values (gate pass) or the site energy and the moment adjoints abar_i
// (outer pass).  The staging region is free on entry.
template <class Cf, bool GATE, class T = typename Cf::T>
__device__ __forceinline__ void forward_graph(const DModel<T>& m, T* region, int i, int ti, const T* acc,
                                              T* __restrict__ h, T* __restrict__ vg, T* __restrict__ site,
                                              T* __restrict__ abar) {
  constexpr int KK = Cf::KK, F = Cf::F, KH = Cf::KH, NM = Cf::NM;
  const DPass<T>& p = GATE ? m.gp : m.op;
  const int lo = p.node_lo, nc = p.node_cnt;  // the pass's node window
  T* mom = region;                            // [KH]
  T* v = mom + KH;                            // [nc]      node n at v[n - lo]
  T* a = v + pass_nodes(m);                   // [nc]
  T* scratch = a + pass_nodes(m);
  const int lane = threadIdx.x;
  for (int n = lane; n < nc; n += WAVE) v[n] = T(0);
#pragma unroll
  for (int t = 0; t < NM; ++t) {
    const int c = lane + WAVE * t;
    if (c < KH) mom[c] = acc[t];
  }
  __syncthreads();
  wave_contract<T, KK>(m, p, GATE, mom, v, lo, nc);
  wave_graph_forward(p, v, scratch, lo);
  if (GATE) {
    // h_i = W S_i (S gathered into the free scratch first; narrow shapes
    // spread each channel's sum over several lanes), node values of the gate
    // range for the reverse
    T* sS = scratch;
    for (int q = lane; q < m.Qg; q += WAVE) sS[q] = v[m.gate_nodes[q] - lo];
    __syncthreads();
    if constexpr (F <= WAVE / 2) {
      // HP lanes per channel, each over a contiguous part of the scalars; the
      // parts added in order by the part-0 lane
      constexpr int HP = F <= WAVE / 8 ? 8 : F <= WAVE / 4 ? 4 : 2;
      const int mu = lane % F, part = lane / F;
      const int qp = (m.Qg + HP - 1) / HP, q0 = min(m.Qg, part * qp), q1 = min(m.Qg, q0 + qp);
      T s0 = T(0), s1 = T(0);
      if (part < HP) {
        int q = q0;
        for (; q + 1 < q1; q += 2) {
          s0 += m.W(mu, q) * sS[q];
          s1 += m.W(mu, q + 1) * sS[q + 1];
        }
        if (q < q1) s0 += m.W(mu, q) * sS[q];
      }
      T s = s0 + s1;
      T tot = s;
#pragma unroll
      for (int k = 1; k < HP; ++k) tot += __shfl(s, min(lane + k * F, WAVE - 1), WAVE);
      if (lane < F) h[(size_t)i * F + mu] = tot;
    } else {
      for (int mu = lane; mu < F; mu += WAVE) {
        T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
        int q = 0;
        for (; q + 3 < m.Qg; q += 4) {
          s0 += m.W(mu, q) * sS[q];
          s1 += m.W(mu, q + 1) * sS[q + 1];
          s2 += m.W(mu, q + 2) * sS[q + 2];
          s3 += m.W(mu, q + 3) * sS[q + 3];
        }
        for (; q < m.Qg; ++q) s0 += m.W(mu, q) * sS[q];
        h[(size_t)i * F + mu] = (s0 + s1) + (s2 + s3);
      }
    }
    T* out = vg + (size_t)i * p.node_cnt;
    for (int n = lane; n < p.node_cnt; n += WAVE) out[n] = v[n];
    return;
  }
  T part = T(0);
  for (int q = lane; q < m.Q; q += WAVE) part += m.readout(ti, q) * v[m.scalar_map[q] - lo];
  // the per-species constant (species + shift coefficients) is added in FP64
  // when the site energies are written out
  const T E = wave_sum(part);
  if (lane == 0) site[i] = E;
  for (int n = lane; n < nc; n += WAVE) a[n] = T(0);
  __syncthreads();
  if (m.scalar_unique) {  // distinct nodes: one lane per scalar
    for (int q = lane; q < m.Q; q += WAVE) a[m.scalar_map[q] - lo] += m.readout(ti, q);
  } else if (lane == 0) {  // repeated nodes: in scalar order
    for (int q = 0; q < m.Q; ++q) a[m.scalar_map[q] - lo] += m.readout(ti, q);
  }
  __syncthreads();
  wave_graph_reverse(p, v, a, scratch, lo);
  wave_adjoint_to_basic<T, KK>(m, p, false, a, abar + (size_t)i * KH, lo, nc);
}

// NA consecutive centres from i0 (fewer past natoms): their edge phase
// together, then each centre's graph phase in turn.
template <class Cf, bool GATE, int NA, class T = typename Cf::T>
__device__ __forceinline__ void pass_forward_group(const DModel<T>& m, const SpeciesTables<T>& tab, T* region, int i0,
                                                   int natoms, const int32_t* __restrict__ types,
                                                   const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                                   T* __restrict__ h, T* __restrict__ vg, T* __restrict__ site,
                                                   T* __restrict__ abar) {
  const int na = min(NA, natoms - i0);
  T acc[NA][Cf::NM];
  forward_edges<Cf, GATE, NA>(m, tab, region, i0, na, types, nbr_off, rec, h, acc);
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    if (a >= na) break;
    if (a > 0) __syncthreads();  // the previous centre's graph phase used the region
    forward_graph<Cf, GATE>(m, region, i0 + a, types[i0 + a], acc[a], h, vg, site, abar);
  }
}

// Forward pass over atoms [0, natoms) in groups of NA consecutive centres: one
// group per workgroup (grid = natoms / NA), or (PERSIST) a persistent grid
// striding over the groups with the species tables loaded once per
// workgroup.  The two are separate instances so that the one-group form keeps
// its straight-line register allocation.
template <class Cf, int NA>
constexpr int forward_waves() { return (sizeof(typename Cf::T) == 4 && NA > 1) ? 4 : 1; }

template <class Cf, bool GATE, bool PERSIST, int NA = 1, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE, (forward_waves<Cf, NA>())) void k_pass_forward(DModel<T> m, int natoms, const int32_t* __restrict__ types,
                                                       const int* __restrict__ nbr_off,
                                                       const Rec<T>* __restrict__ rec, T* __restrict__ h,
                                                       T* __restrict__ vg, T* __restrict__ site,
                                                       T* __restrict__ abar) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tabs_base = reinterpret_cast<T*>(smem_raw);
  T* region = tabs_base + species_tables_size<Cf, FwdStage<Cf>::CL>(m.C);
  SpeciesTables<T> tab;
  load_species_tables<Cf, FwdStage<Cf>::CL>(m, GATE, tabs_base, tab);
  __syncthreads();
  if constexpr (!PERSIST) {
    pass_forward_group<Cf, GATE, NA>(m, tab, region, blockIdx.x * NA, natoms, types, nbr_off, rec, h, vg, site, abar);
  } else {
    for (int i0 = blockIdx.x * NA; i0 < natoms; i0 += gridDim.x * NA) {
      pass_forward_group<Cf, GATE, NA>(m, tab, region, i0, natoms, types, nbr_off, rec, h, vg, site, abar);
      __syncthreads();  // the next group overwrites the region
    }
  }
}

// ------------------------------------------------------------------ edge backward
// One directed edge e = i -> j of the centre's own energy:
//   g_e = sum_b abar_i[b] w_mu (R'_mu u Y_b + R_mu grad Y_b),
// summed over the centre's edges into F_i (with the virial -sum d (x) g) and
// written to gedge_in at the edge's incoming slot (each pass overwrites it
// after the previous pass's gather consumed it), and for the
// outer pass of an Effect model the gate-adjoint message
//   u_e,mu = tau_{t_i,mu} R_mu(r; t_i, t_j) sum_m abar_i[k,l,m] Y_lm(d_e),
// which the neighbour j gathers into h_bar_j (times alpha_{t_j,mu}).

// FP32, L <= 3: at most 168 VGPRs, so three waves share each SIMD.
template <class Cf, bool GATE>
constexpr int edge_backward_waves() { return (sizeof(typename Cf::T) == 4 && Cf::LL <= 3) ? 3 : 1; }

// NA centres i0 .. i0 + NA - 1 of an edge-backward pass (their edge lists
// are contiguous, so two or four atoms' edges fill the lanes better than one
// atom's ~90).  sA: [NA][KH] adjoints, sV: [NA][9][LR] lane partials, sXR:
// [F][LR] gate-message columns; tab already loaded.
// The NA centres' moment adjoints [na][KH] into LDS: NA KH is known at
// compile time, so each lane issues its loads U at a time before storing them
// (a load-then-store loop would pay one memory round trip per 64 values,
// seven at L5 with two centres); U is kept small so that the staging adds
// few registers to the kernels' peak.
template <class Cf, int NA, class T = typename Cf::T>
__device__ __forceinline__ void stage_adjoints(T* sA, const T* __restrict__ adj, int i0, int na) {
  constexpr int NAK = NA * Cf::KH, IT = (NAK + WAVE - 1) / WAVE;
  constexpr int U = IT < 16 / (int)sizeof(T) ? IT : 16 / (int)sizeof(T);
  const int n = na * Cf::KH;
  const T* src = adj + (size_t)i0 * Cf::KH;
#pragma unroll
  for (int it0 = 0; it0 < IT; it0 += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = (int)threadIdx.x + (it0 + u) * WAVE;
      v[u] = c < n ? src[c] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = (int)threadIdx.x + (it0 + u) * WAVE;
      if (c < n) sA[c] = v[u];
    }
  }
}

template <class Cf, bool GATE, int NA, bool MSG = true, class T = typename Cf::T>
__device__ __forceinline__ void edge_backward_group(const DModel<T>& m, const SpeciesTables<T>& tab, T* sA, T* sV,
                                                    T* sXR, int i0, int natoms, const int32_t* __restrict__ types,
                                                    const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                                    const T* __restrict__ h, const T* __restrict__ adj,
                                                    const int* __restrict__ in_pos, T* __restrict__ gedge_in,
                                                    T* __restrict__ uedge, T* __restrict__ fw) {
  constexpr int LL = Cf::LL, KK = Cf::KK, L1 = Cf::L1, F = Cf::F, KH = Cf::KH;
  const int lane = threadIdx.x;
  // gated outer pass: the neighbour's gate signal enters the radial weight;
  // with MSG it also writes the gate-adjoint messages u_e (else the gather
  // recomputes them)
  const bool hgate = !GATE && m.gated;
  const bool hmsg = MSG && hgate;
  const int na = min(NA, natoms - i0);
  stage_adjoints<Cf, NA>(sA, adj, i0, na);
#pragma unroll
  for (int c = 0; c < NA * 9; ++c) sV[c * LR + lane] = T(0);
  // the centres' edge ranges: [eb[a], eb[a + 1]) for atom i0 + a
  int eb[NA + 1], ta[NA];
#pragma unroll
  for (int a = 0; a <= NA; ++a) eb[a] = nbr_off[i0 + min(a, na)];
#pragma unroll
  for (int a = 0; a < NA; ++a) ta[a] = types[i0 + min(a, na - 1)];
  __syncthreads();
  // outgoing force sum and virial of the lane's current centre; a lane's edges
  // visit the centres in order, so on a change of centre the partial goes to
  // the lane's LDS slot
  T w[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) w[k] = T(0);
  int acur = 0;
  for (int e = eb[0] + lane; e < eb[NA]; e += WAVE) {
    int ai = 0;  // the edge's centre among the workgroup's atoms
#pragma unroll
    for (int a = 1; a < NA; ++a) ai += e >= eb[a];
    if (ai != acur) {
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        sV[(acur * 9 + k) * LR + lane] = w[k];
        w[k] = T(0);
      }
      acur = ai;
    }
    int ti = ta[0];
#pragma unroll
    for (int a = 1; a < NA; ++a)
      if (ai == a) ti = ta[a];
    const T* sAi = sA + ai * KH;
    const T* x0 = tab.x0 + ti * F;
    const Rec<T> q = rec[e];
    const int j = rec_j(q);
    const T dx = q.x, dy = q.y, dz = q.z;
    T r, inv;
    edge_length(dx * dx + dy * dy + dz * dz, r, inv);
    const T ux = dx * inv, uy = dy * inv, uz = dz * inv;
    const int tj = rec_t(q);
    // the neighbour's gate-signal row: one batch of vector loads, issued
    // before the basis work (not one dependent load per channel)
    T hj[F];
    if (hgate) load_row<T, F>(h + (size_t)j * F, hj);
    // weighted radial values x0 x1 R and x0 x1 R' in registers; x0 R (the gate
    // message factor) parked in this lane's own LDS column
    T R[F], dR[F], us[F];
    radial<Cf, true>(m, GATE, tab, ti * m.C + tj, r, R, dR);
    const T* x0j = tab.x0 + tj * F;
    const T* alj = tab.alpha + tj * F;
#pragma unroll
    for (int mu = 0; mu < F; ++mu) {
      T x1 = x0j[mu];
      if (hgate) x1 += alj[mu] * hj[mu];
      if (hmsg) sXR[mu * LR + lane] = x0[mu] * R[mu];
      const T wt = x0[mu] * x1;
      R[mu] *= wt;
      dR[mu] *= wt;
      us[mu] = T(0);
    }
    // sY = sum_h (sum_k abar R') Y_h, g = sum_h (sum_k abar R) G_h and, per
    // degree l, ls[l] = sum (sum_k abar R) Y_h: the harmonics' Cartesian
    // gradients (G - l Y u) / r are assembled once after the stream
    T sY = T(0), gx = T(0), gy = T(0), gz = T(0), lsum = T(0);
    sh_stream_solid<T, LL>(ux, uy, uz, [&](int l, int mm, T Y, T Gx, T Gy, T Gz) {
      const int hh = hidx(l, mm);
      // the moment adjoints of this (l, m) are read here: their address takes
      // a zero from an empty asm of Y (hoisted over the whole unrolled stream
      // they would hold K H registers; the asm touches no memory)
      int z;
      asm("" : "=v"(z) : "0"(0), "v"(Y));
      const T* sAh = sAi + z;
      // channel sums first (sum_k abar R', sum_k abar R), then Y and grad Y once
      T adR = T(0), aR = T(0);
#pragma unroll
      for (int k = 0; k < KK; ++k) {
        const int mu = k * L1 + l;
        const T av = sAh[hh * KK + k];
        adR += av * dR[mu];
        aR += av * R[mu];
        if (hmsg) us[mu] = fma(av, Y, us[mu]);
      }
      sY += adR * Y;
      gx += aR * Gx; gy += aR * Gy; gz += aR * Gz;
      if (l > 0) lsum += (T(l) * aR) * Y;
    });
    const T sr = sY - lsum * inv;
    const T g0 = ux * sr + gx * inv, g1 = uy * sr + gy * inv, g2 = uz * sr + gz * inv;
    const int slot = in_pos[e];  // the edge's position in its target's incoming list
    T* gi = gedge_in + 3 * (size_t)slot;
    gi[0] = g0; gi[1] = g1; gi[2] = g2;
    w[0] += g0; w[1] += g1; w[2] += g2;
    w[3] -= dx * g0; w[4] -= dy * g1; w[5] -= dz * g2;
    w[6] -= dx * g1; w[7] -= dx * g2; w[8] -= dy * g2;
    if (hmsg) {
      T* ue = uedge + (size_t)slot * F;
#pragma unroll
      for (int mu = 0; mu < F; ++mu) ue[mu] = sXR[mu * LR + lane] * us[mu];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) sV[(acur * 9 + k) * LR + lane] = w[k];
  __syncthreads();
  // one lane per (centre, component): its row of lane partials summed in a
  // fixed order; the outer pass starts the atom's row, the gate pass adds
  static_assert(NA * 9 <= WAVE, "one lane per centre component");
  if (lane < NA * 9 && lane / 9 < na) {
    const T* row = sV + lane * LR;
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll
    for (int e = 0; e < WAVE; e += 4) {
      s0 += row[e];
      s1 += row[e + 1];
      s2 += row[e + 2];
      s3 += row[e + 3];
    }
    const T tot = (s0 + s1) + (s2 + s3);
    T* o = fw + (size_t)(i0 + lane / 9) * 9 + lane % 9;
    *o = GATE ? *o + tot : tot;
  }
}

// Edge backward over groups of NA centres: one group per workgroup, or
// (PERSIST) a persistent grid striding over the groups.
template <class Cf, bool GATE, int NA, bool PERSIST, bool MSG = true, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE, (edge_backward_waves<Cf, GATE>())) void k_edge_backward(DModel<T> m, int natoms,
                                                        const int32_t* __restrict__ types,
                                                        const int* __restrict__ nbr_off,
                                                        const Rec<T>* __restrict__ rec, const T* __restrict__ h,
                                                        const T* __restrict__ adj, const int* __restrict__ in_pos,
                                                        T* __restrict__ gedge_in, T* __restrict__ uedge,
                                                        T* __restrict__ fw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sA = reinterpret_cast<T*>(smem_raw);  // [NA][KH]
  // [NA][9][LR] per-lane partials of each centre: outgoing force sum, virial
  // (padded rows, so that the row sums at the end read distinct banks)
  T* sV = sA + NA * Cf::KH + species_tables_size<Cf>(m.C);
  T* sXR = sV + NA * 9 * LR;  // [F][LR] lane-private columns: x0 R (gate message)
  SpeciesTables<T> tab;
  load_species_tables<Cf>(m, GATE, sA + NA * Cf::KH, tab);
  if constexpr (!PERSIST) {
    edge_backward_group<Cf, GATE, NA, MSG>(m, tab, sA, sV, sXR, blockIdx.x * NA, natoms, types, nbr_off, rec, h,
                                           adj, in_pos, gedge_in, uedge, fw);
  } else {
    for (int i0 = blockIdx.x * NA; i0 < natoms; i0 += gridDim.x * NA) {
      edge_backward_group<Cf, GATE, NA, MSG>(m, tab, sA, sV, sXR, i0, natoms, types, nbr_off, rec, h, adj, in_pos,
                                             gedge_in, uedge, fw);
      __syncthreads();  // the next group overwrites sA and sV
    }
  }
}

// Gate-adjoint messages of the outer pass on their own (split form): for every
// edge e = i -> j of the NA centres, u_e,mu = tau_{t_i,mu} R_mu(r; t_i, t_j)
// sum_m abar_i[k,l,m] Y_lm(d_e) at the edge's incoming slot -- the values the
// combined pass writes, formed by the same operations in the same order, but
// from radial and harmonic values only (no derivatives, no forces), so the
// force pass and this one each hold far fewer registers than the combined one.
template <class Cf, int NA, class T = typename Cf::T>
__device__ __forceinline__ void edge_messages_group(const DModel<T>& m, const SpeciesTables<T>& tab, T* sA, int i0,
                                                    int natoms, const int32_t* __restrict__ types,
                                                    const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                                    const T* __restrict__ adj, const int* __restrict__ in_pos,
                                                    T* __restrict__ uedge) {
  constexpr int LL = Cf::LL, KK = Cf::KK, L1 = Cf::L1, F = Cf::F, KH = Cf::KH;
  const int lane = threadIdx.x;
  const int na = min(NA, natoms - i0);
  for (int c = lane; c < na * KH; c += WAVE) sA[c] = adj[(size_t)i0 * KH + c];
  int eb[NA + 1], ta[NA];
#pragma unroll
  for (int a = 0; a <= NA; ++a) eb[a] = nbr_off[i0 + min(a, na)];
#pragma unroll
  for (int a = 0; a < NA; ++a) ta[a] = types[i0 + min(a, na - 1)];
  __syncthreads();
  for (int e = eb[0] + lane; e < eb[NA]; e += WAVE) {
    int ai = 0;
#pragma unroll
    for (int a = 1; a < NA; ++a) ai += e >= eb[a];
    int ti = ta[0];
#pragma unroll
    for (int a = 1; a < NA; ++a)
      if (ai == a) ti = ta[a];
    const T* sAi = sA + ai * KH;
    const T* x0 = tab.x0 + ti * F;
    const Rec<T> q = rec[e];
    const int tj = rec_t(q);
    const T dx = q.x, dy = q.y, dz = q.z;
    T r, inv;
    edge_length(dx * dx + dy * dy + dz * dz, r, inv);
    // x0 R per channel first, then the harmonics degree by degree: the K
    // messages of degree l are finished and stored before degree l + 1 is
    // formed (the m terms summed in the order of the combined pass, so the
    // values are the same bit for bit)
    T w[F];
    radial_each<Cf, false>(m, false, tab, ti * m.C + tj, r, [&](int mu, T R, T) { w[mu] = x0[mu] * R; });
    T* ue = uedge + (size_t)in_pos[e] * F;
    T dep = r;
    sh_values_by_l<T, LL>(dx * inv, dy * inv, dz * inv, [&](int l, const T* Yl) {
      // the adjoint reads of degree l wait for the previous degree's sums (an
      // empty asm's zero): read all at once they would hold K H registers
      int z;
      asm("" : "=v"(z) : "0"(0), "v"(dep), "v"(Yl[l]));
      const T* sAh = sAi + z;
#pragma unroll
      for (int k = 0; k < KK; ++k) {
        T u = T(0);
        u = fma(sAh[hidx(l, 0) * KK + k], Yl[l], u);
#pragma unroll
        for (int mm = 1; mm <= l; ++mm) {
          u = fma(sAh[hidx(l, mm) * KK + k], Yl[l + mm], u);
          u = fma(sAh[hidx(l, -mm) * KK + k], Yl[l - mm], u);
        }
        ue[k * L1 + l] = w[k * L1 + l] * u;
        dep = u;
      }
    });
  }
}

template <class Cf, int NA, bool PERSIST, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE) void k_edge_messages(DModel<T> m, int natoms, const int32_t* __restrict__ types,
                                                        const int* __restrict__ nbr_off,
                                                        const Rec<T>* __restrict__ rec, const T* __restrict__ adj,
                                                        const int* __restrict__ in_pos, T* __restrict__ uedge) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sA = reinterpret_cast<T*>(smem_raw);  // [NA][KH], then the outer pass's species tables
  SpeciesTables<T> tab;
  load_species_tables<Cf>(m, false, sA + NA * Cf::KH, tab);
  if constexpr (!PERSIST) {
    edge_messages_group<Cf, NA>(m, tab, sA, blockIdx.x * NA, natoms, types, nbr_off, rec, adj, in_pos, uedge);
  } else {
    for (int i0 = blockIdx.x * NA; i0 < natoms; i0 += gridDim.x * NA) {
      edge_messages_group<Cf, NA>(m, tab, sA, i0, natoms, types, nbr_off, rec, adj, in_pos, uedge);
      __syncthreads();  // the next group overwrites sA
    }
  }
}

// ------------------------------------------------------------------ gather
// Gate graph reverse of one atom from its gate-signal adjoint h_bar (LDS hbs,
// alpha already applied): seeds S_bar_q = sum_mu h_bar_mu W_mu,q, reverse over
// the stored gate node values, transposed contraction -> gate basic adjoints.
// v, a: LDS windows over the gate pass's node range (node n at [n - node_lo]).
template <class Cf, class T = typename Cf::T>
__device__ void gate_reverse_atom(const DModel<T>& m, int i, const T* __restrict__ vg, T* __restrict__ gbar,
                                  const T* hbs, T* v, T* a, T* scratch) {
  constexpr int KK = Cf::KK, F = Cf::F;
  const int lane = threadIdx.x;
  const DPass<T>& p = m.gp;
  const int lo = p.node_lo, nc = p.node_cnt;
  const T* vi = vg + (size_t)i * nc;
  wave_rows(0, nc, [&](int n) {  // four node loads in flight per lane
    v[n] = vi[n];
    a[n] = T(0);
  });
  __syncthreads();
  for (int q = lane; q < m.Qg; q += WAVE) {
    T s0 = T(0), s1 = T(0);
#pragma unroll
    for (int mu = 0; mu + 1 < F; mu += 2) {
      s0 += hbs[mu] * m.W(mu, q);
      s1 += hbs[mu + 1] * m.W(mu + 1, q);
    }
    if (F % 2) s0 += hbs[F - 1] * m.W(F - 1, q);
    a[m.gate_nodes[q] - lo] = s0 + s1;
  }
  __syncthreads();
  wave_graph_reverse(p, v, a, scratch, lo);
  wave_adjoint_to_basic<T, KK>(m, p, true, a, gbar + (size_t)i * Cf::KH, lo, nc);
}

// F_a -= sum_{e' in in(a)} g_e' over the atom's incoming-list slots (written
// there by the edge kernels, which also added the atom's outgoing sum; no
// atomics, contiguous reads).  The outer pass of an
// Effect model also forms h_bar_a = alpha_{t_a} sum_{e' in in(a)} u_e' and
// either runs the gate graph reverse at once (hbar_out == null) or stores
// h_bar for a later reverse (the pair-style path, whose ghost adjoints are
// communicated in between).  Ghost atoms (a >= nowned) start from a zeroed row.
// RECOMP (batch path of an Effect model): the messages are not read from
// uedge but formed here from the atom's own outgoing records -- each is the
// reverse of one incoming edge j -> i, so
//   u_{j->i,mu} = tau_{t_j,mu} R_mu(r; t_j, t_i) sum_m abar_j[k,l,m] Y_lm(x_i - x_j)
// with the same operations, in the same order, as the edge backward of j
// would have used; `tab` holds the outer pass's species tables.
template <class Cf, bool GATE, bool RECOMP = false, class T = typename Cf::T>
__device__ __forceinline__ void gather_atom(const DModel<T>& m, int i, const int32_t* __restrict__ types,
                                            const int* __restrict__ in_off, const T* __restrict__ gedge_in,
                                            const T* __restrict__ uedge, const T* __restrict__ vg,
                                            T* __restrict__ gbar, T* __restrict__ hbar_out, T* __restrict__ fw,
                                            const Rec<T>* __restrict__ rec = nullptr,
                                            const T* __restrict__ abar = nullptr,
                                            const SpeciesTables<T>* tab = nullptr) {
  constexpr int F = Cf::F, LL = Cf::LL, KK = Cf::KK, L1 = Cf::L1, KH = Cf::KH;
  const int lane = threadIdx.x;
  const bool hgate = !GATE && m.gated;
  T fx = T(0), fy = T(0), fz = T(0);
  T hb[F];
#pragma unroll
  for (int mu = 0; mu < F; ++mu) hb[mu] = T(0);
  // incoming slots (contiguous); narrow shapes take two batches per step so
  // that their loads are in flight together
  constexpr int IN_STEP = (F <= 16 && !RECOMP) ? 2 : 1;
  const int k1 = in_off[i + 1];
  const int ti = RECOMP ? types[i] : 0;
  for (int k0 = in_off[i] + lane; k0 < k1; k0 += IN_STEP * WAVE) {
#pragma unroll
    for (int u = 0; u < IN_STEP; ++u) {
      const int k = k0 + u * WAVE;
      if (k >= k1) break;
      const T* gi = gedge_in + 3 * (size_t)k;
      fx -= gi[0]; fy -= gi[1]; fz -= gi[2];
      if constexpr (RECOMP) {
        if (hgate) {
          const Rec<T> q = rec[k];
          const int j = rec_j(q), tj = rec_t(q);
          const T dx = -q.x, dy = -q.y, dz = -q.z;  // j -> i
          T r, inv;
          edge_length(dx * dx + dy * dy + dz * dz, r, inv);
          const T* aj = abar + (size_t)j * KH;
          T S[F];
#pragma unroll
          for (int mu = 0; mu < F; ++mu) S[mu] = T(0);
          sh_stream_values<T, LL>(dx * inv, dy * inv, dz * inv, [&](int l, int mm, T Y) {
            const int hh = hidx(l, mm);
            int z;  // (the neighbour's adjoints of this (l, m) are read here)
            asm("" : "=v"(z) : "0"(0), "v"(Y));
            const T* ajh = aj + z;
#pragma unroll
            for (int c = 0; c < KK; ++c) S[c * L1 + l] += ajh[hh * KK + c] * Y;
          });
          const T* x0 = tab->x0 + tj * F;
          radial_each<Cf, false>(m, false, *tab, tj * m.C + ti, r,
                                 [&](int mu, T R, T) { hb[mu] += (x0[mu] * R) * S[mu]; });
        }
      } else if (hgate) {
        const T* ue = uedge + (size_t)k * F;
#pragma unroll
        for (int mu = 0; mu < F; ++mu) hb[mu] += ue[mu];
      }
    }
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int gn = gate_nodes(m);
  T* v = reinterpret_cast<T*>(smem_raw);  // [gn] gate-pass node window
  T* a = v + gn;                           // [gn]
  T* hbs = a + gn;                         // [F]
  T* scratch = hbs + F;
  if constexpr (F >= 16) {
    if (hgate) {
      // F >= 16: the 3 + F lane sums through an LDS transpose (lane
      // columns, then one row per lane summed in order), before v / a are used
      T* tr = v;  // [3 + F][LR]
      tr[lane] = fx;
      tr[LR + lane] = fy;
      tr[2 * LR + lane] = fz;
#pragma unroll
      for (int mu = 0; mu < F; ++mu) tr[(3 + mu) * LR + lane] = hb[mu];
      __syncthreads();
      T tot = T(0);
      if (lane < 3 + F) {
        const T* row = tr + lane * LR;
        T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll 4
        for (int e = 0; e < WAVE; e += 4) {
          s0 += row[e];
          s1 += row[e + 1];
          s2 += row[e + 2];
          s3 += row[e + 3];
        }
        tot = (s0 + s1) + (s2 + s3);
      }
      __syncthreads();
      fx = __shfl(tot, 0, WAVE);
      fy = __shfl(tot, 1, WAVE);
      fz = __shfl(tot, 2, WAVE);
      if (lane == 0) {
        T* o = fw + (size_t)i * 9;
        o[0] += fx; o[1] += fy; o[2] += fz;
      }
      const int ta = types[i];
      if (lane >= 3 && lane < 3 + F) {
        const int mu = lane - 3;
        const T s = tot * m.alpha(ta, mu);
        if (hbar_out) hbar_out[(size_t)i * F + mu] = s;
        else hbs[mu] = s;
      }
      if (hbar_out) return;
      gate_reverse_atom<Cf>(m, i, vg, gbar, hbs, v, a, scratch);
      return;
    }
  }
  if (!hgate) {
    T f3[3] = {fx, fy, fz};
    wave_sum_n(f3);
    fx = f3[0]; fy = f3[1]; fz = f3[2];
  } else {  // the force and message sums together
    T fh[3 + F];
    fh[0] = fx; fh[1] = fy; fh[2] = fz;
#pragma unroll
    for (int mu = 0; mu < F; ++mu) fh[3 + mu] = hb[mu];
    wave_sum_n(fh);
    fx = fh[0]; fy = fh[1]; fz = fh[2];
#pragma unroll
    for (int mu = 0; mu < F; ++mu) hb[mu] = fh[3 + mu];
  }
  if (lane == 0) {
    T* o = fw + (size_t)i * 9;
    o[0] += fx; o[1] += fy; o[2] += fz;
  }
  if (!hgate) return;
  const int ta = types[i];
#pragma unroll
  for (int mu = 0; mu < F; ++mu) {
    const T s = hb[mu] * m.alpha(ta, mu);
    if (hbar_out) {
      if (lane == 0) hbar_out[(size_t)i * F + mu] = s;
    } else if (lane == 0) {
      hbs[mu] = s;
    }
  }
  if (hbar_out) return;
  gate_reverse_atom<Cf>(m, i, vg, gbar, hbs, v, a, scratch);
}

// Gather over atoms [0, natoms): one atom per workgroup, or (PERSIST) a
// persistent grid striding over the atoms.
template <class Cf, bool GATE, bool PERSIST, bool RECOMP = false, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE) void k_gather(DModel<T> m, int natoms, const int32_t* __restrict__ types,
                                                 const int* __restrict__ in_off, const T* __restrict__ gedge_in,
                                                 const T* __restrict__ uedge, const T* __restrict__ vg,
                                                 T* __restrict__ gbar, T* __restrict__ hbar_out,
                                                 T* __restrict__ fw, const Rec<T>* __restrict__ rec,
                                                 const T* __restrict__ abar) {
  SpeciesTables<T> tab{};
  if constexpr (RECOMP) {  // the outer pass's tables after the walk / transpose region
    extern __shared__ __align__(16) unsigned char smem_raw[];
    load_species_tables<Cf>(m, false, reinterpret_cast<T*>(smem_raw) + gather_region<Cf>(m), tab);
    __syncthreads();
  }
  if constexpr (!PERSIST) {
    gather_atom<Cf, GATE, RECOMP>(m, blockIdx.x, types, in_off, gedge_in, uedge, vg, gbar, hbar_out, fw, rec, abar,
                                  &tab);
  } else {
    for (int i = blockIdx.x; i < natoms; i += gridDim.x) {
      gather_atom<Cf, GATE, RECOMP>(m, i, types, in_off, gedge_in, uedge, vg, gbar, hbar_out, fw, rec, abar, &tab);
      __syncthreads();  // the next atom reuses the LDS
    }
  }
}

// Gate reverse of owned atoms from stored h_bar (pair-style path).
template <class Cf, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE) void k_gate_reverse(DModel<T> m, const T* __restrict__ hbar,
                                                       const T* __restrict__ vg, T* __restrict__ gbar) {
  constexpr int F = Cf::F;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int gn = gate_nodes(m);
  T* v = reinterpret_cast<T*>(smem_raw);
  T* a = v + gn;
  T* hbs = a + gn;
  T* scratch = hbs + F;
  const int i = blockIdx.x;
  for (int mu = threadIdx.x; mu < F; mu += WAVE) hbs[mu] = hbar[(size_t)i * F + mu];
  gate_reverse_atom<Cf>(m, i, vg, gbar, hbs, v, a, scratch);
}

// Elements of the gather's walk / transpose region: the gate graph walk, or
// for F >= 16 the 3 + F lane sums transposed through the same region.
template <class Cf>
__host__ __device__ inline int gather_region(const DModel<typename Cf::T>& m) {
  const int sc = graph_scratch(m.gp);
  const int walk = 2 * gate_nodes(m) + Cf::F + (sc > 1 ? sc : 1);
  const int sums = Cf::F >= 16 ? (3 + Cf::F) * LR : 0;
  return walk > sums ? walk : sums;
}

template <class Cf>
size_t gather_lds(const DModel<typename Cf::T>& m, bool recomp = false) {
  return sizeof(typename Cf::T) *
         ((size_t)gather_region<Cf>(m) + (recomp ? (size_t)species_tables_size<Cf>(m.C) : 0));
}

template <class Cf>
size_t backward_lds(const DModel<typename Cf::T>& m, int na) {
  return sizeof(typename Cf::T) *
         ((size_t)na * (Cf::KH + 9 * LR) + species_tables_size<Cf>(m.C) + (m.gated ? (size_t)Cf::F * LR : 0));
}

// ------------------------------------------------------------------ driver
template <class Cf>
std::string kname(const char* base) {
  return std::string(base) + "<" + tname<typename Cf::T>() + ",L" + std::to_string(Cf::LL) + "K" +
         std::to_string(Cf::KK) + "B" + std::to_string(Cf::BB) + ">";
}

constexpr size_t LDS_MAX = 160 * 1024;  // LDS of one CDNA4 compute unit

template <class K>
void allow_lds(K kern, size_t bytes) {
  if (bytes > LDS_MAX)
    throw Error(SUS2_E_UNSUPPORTED, "model needs " + std::to_string(bytes) + " bytes of LDS per workgroup (limit " +
                                        std::to_string(LDS_MAX) + ")");
  if (bytes > 64 * 1024)
    SUS2_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)bytes));
}

// Workgroups of one wavefront of `kern` that fit on the whole device at once.
// With whole_simds the count per CU is rounded down to a multiple of its four
// SIMDs: grid-stride loops give every group the same share of atoms, and for
// an issue-bound kernel a fifth wave on a CU shares a SIMD with another and
// finishes last.
template <class K>
int resident_groups(K kern, size_t lds, bool whole_simds) {
  constexpr int SIMDS = 4;
  allow_lds(kern, lds);
  int per_cu = 0, dev = 0, cus = 0;
  SUS2_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), WAVE, lds));
  SUS2_CHECK(hipGetDevice(&dev));
  SUS2_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (whole_simds && per_cu > SIMDS) per_cu -= per_cu % SIMDS;
  return std::max(1, per_cu) * std::max(1, cus);
}

// Grid of a per-atom (per-group) kernel over `units` work items under the
// launch shape: one workgroup per item, or grid_mult times the resident
// workgroups striding over them (persistent).
template <class K>
unsigned tuned_grid(const EfsTune& t, K kern, size_t lds, long units) {
  if (units <= 0) return 0;
  if (t.grid_mult <= 0) return (unsigned)units;
  return (unsigned)std::min<long>(units, (long)t.grid_mult * resident_groups(kern, lds, false));
}

// Batches of at least this many atoms run the forward passes with two centres
// per workgroup (two centres' ~90 edges each fill the lanes of the 64-edge
// batches better than one centre's); smaller ones keep one centre per
// workgroup, so that every SIMD still has several workgroups in flight.
constexpr int kFwdPairAtoms = 32768;

template <class Cf, bool GATE, int NA>
void launch_pass_forward_na(const DevModelSet& dm, Workspace& ws, int n, const int32_t* types, const int* off,
                            const Rec<typename Cf::T>* rec, typename Cf::T* h, typename Cf::T* vg,
                            typename Cf::T* site, typename Cf::T* abar, hipStream_t st) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  const size_t lf = forward_lds<Cf>(m);
  const char* name = GATE ? "pass_forward_gate" : "pass_forward_outer";
  const long groups = (n + NA - 1) / NA;
  if (dm.tune_of<T>().grid_mult <= 0) {
    allow_lds(k_pass_forward<Cf, GATE, false, NA>, lf);
    SUS2_LAUNCH(ws, st, kname<Cf>(name), (k_pass_forward<Cf, GATE, false, NA>), dim3((unsigned)groups), dim3(WAVE),
                lf, m, n, types, off, rec, h, vg, site, abar);
    return;
  }
  const unsigned g = tuned_grid(dm.tune_of<T>(), k_pass_forward<Cf, GATE, true, NA>, lf, groups);
  SUS2_LAUNCH(ws, st, kname<Cf>(name), (k_pass_forward<Cf, GATE, true, NA>), dim3(g), dim3(WAVE), lf, m, n, types,
              off, rec, h, vg, site, abar);
}

template <class Cf, bool GATE>
void launch_pass_forward(const DevModelSet& dm, Workspace& ws, int n, const int32_t* types, const int* off,
                         const Rec<typename Cf::T>* rec, typename Cf::T* h, typename Cf::T* vg, typename Cf::T* site,
                         typename Cf::T* abar, hipStream_t st) {
  if (n <= 0) return;
  if (n >= kFwdPairAtoms)
    launch_pass_forward_na<Cf, GATE, 2>(dm, ws, n, types, off, rec, h, vg, site, abar, st);
  else
    launch_pass_forward_na<Cf, GATE, 1>(dm, ws, n, types, off, rec, h, vg, site, abar, st);
}

template <class Cf, bool GATE, int NA, bool MSG>
void launch_edge_backward_na(const DevModelSet& dm, Workspace& ws, int n, const int32_t* types, const int* off,
                             const Rec<typename Cf::T>* rec, const typename Cf::T* h, const typename Cf::T* adj,
                             const int* in_pos, typename Cf::T* gedge_in, typename Cf::T* uedge,
                             typename Cf::T* fw, hipStream_t st) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  const size_t lb = backward_lds<Cf>(m, NA);
  if (n <= 0) return;
  const char* name = GATE ? "edge_backward_gate" : MSG ? "edge_backward_outer" : "edge_backward_outer_nomsg";
  const int groups = (n + NA - 1) / NA;
  if (dm.tune_of<T>().grid_mult <= 0) {
    allow_lds(k_edge_backward<Cf, GATE, NA, false, MSG>, lb);
    SUS2_LAUNCH(ws, st, kname<Cf>(name), (k_edge_backward<Cf, GATE, NA, false, MSG>), dim3(groups), dim3(WAVE), lb,
                m, n, types, off, rec, h, adj, in_pos, gedge_in, uedge, fw);
    return;
  }
  const unsigned g = tuned_grid(dm.tune_of<T>(), k_edge_backward<Cf, GATE, NA, true, MSG>, lb, groups);
  SUS2_LAUNCH(ws, st, kname<Cf>(name), (k_edge_backward<Cf, GATE, NA, true, MSG>), dim3(g), dim3(WAVE), lb, m, n,
              types, off, rec, h, adj, in_pos, gedge_in, uedge, fw);
}

// Edge backward with the tuned centres per group (1, 2 or 4); MSG: the outer
// pass writes the gate-adjoint messages (else the gather recomputes them).
template <class Cf, bool GATE, bool MSG = true>
void launch_edge_backward(const DevModelSet& dm, Workspace& ws, int n, const int32_t* types, const int* off,
                          const Rec<typename Cf::T>* rec, const typename Cf::T* h, const typename Cf::T* adj,
                          const int* in_pos, typename Cf::T* gedge_in, typename Cf::T* uedge, typename Cf::T* fw,
                          hipStream_t st) {
  switch (dm.tune_of<typename Cf::T>().eb_atoms) {
    case 1: return launch_edge_backward_na<Cf, GATE, 1, MSG>(dm, ws, n, types, off, rec, h, adj, in_pos, gedge_in, uedge, fw, st);
    case 4: return launch_edge_backward_na<Cf, GATE, 4, MSG>(dm, ws, n, types, off, rec, h, adj, in_pos, gedge_in, uedge, fw, st);
    default: return launch_edge_backward_na<Cf, GATE, 2, MSG>(dm, ws, n, types, off, rec, h, adj, in_pos, gedge_in, uedge, fw, st);
  }
}

template <class Cf, int NA>
void launch_edge_messages_na(const DevModelSet& dm, Workspace& ws, int n, const int32_t* types, const int* off,
                             const Rec<typename Cf::T>* rec, const typename Cf::T* adj, const int* in_pos,
                             typename Cf::T* uedge, hipStream_t st) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  const size_t lm = sizeof(T) * ((size_t)NA * Cf::KH + species_tables_size<Cf>(m.C));
  if (n <= 0) return;
  const int groups = (n + NA - 1) / NA;
  if (dm.tune_of<T>().grid_mult <= 0) {
    allow_lds(k_edge_messages<Cf, NA, false>, lm);
    SUS2_LAUNCH(ws, st, kname<Cf>("edge_messages"), (k_edge_messages<Cf, NA, false>), dim3(groups), dim3(WAVE), lm,
                m, n, types, off, rec, adj, in_pos, uedge);
    return;
  }
  const unsigned g = tuned_grid(dm.tune_of<T>(), k_edge_messages<Cf, NA, true>, lm, groups);
  SUS2_LAUNCH(ws, st, kname<Cf>("edge_messages"), (k_edge_messages<Cf, NA, true>), dim3(g), dim3(WAVE), lm, m, n,
              types, off, rec, adj, in_pos, uedge);
}

// Messages-only pass with the tuned centres per group.
template <class Cf>
void launch_edge_messages(const DevModelSet& dm, Workspace& ws, int n, const int32_t* types, const int* off,
                          const Rec<typename Cf::T>* rec, const typename Cf::T* adj, const int* in_pos,
                          typename Cf::T* uedge, hipStream_t st) {
  switch (dm.tune_of<typename Cf::T>().eb_atoms) {
    case 1: return launch_edge_messages_na<Cf, 1>(dm, ws, n, types, off, rec, adj, in_pos, uedge, st);
    case 4: return launch_edge_messages_na<Cf, 4>(dm, ws, n, types, off, rec, adj, in_pos, uedge, st);
    default: return launch_edge_messages_na<Cf, 2>(dm, ws, n, types, off, rec, adj, in_pos, uedge, st);
  }
}

// RECOMP (batch path, outer pass of an Effect model): the messages are formed
// from rec / abar instead of read from uedge; in_off is then the outgoing CSR.
template <class Cf, bool GATE, bool RECOMP = false>
void launch_gather(const DevModelSet& dm, Workspace& ws, int nall, const int32_t* types, const int* in_off,
                   const typename Cf::T* gedge_in, const typename Cf::T* uedge, const typename Cf::T* vg,
                   typename Cf::T* gbar, typename Cf::T* hbar_out, typename Cf::T* fw, hipStream_t st,
                   const Rec<typename Cf::T>* rec = nullptr, const typename Cf::T* abar = nullptr) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  // gated outer gathers with F >= 16 sum their messages through LDS; the gate
  // graph reverse of the outer gather needs its walk LDS
  const bool outer_gated = !GATE && m.gated;
  const size_t lg = outer_gated ? gather_lds<Cf>(m, RECOMP) : 0;
  if (nall <= 0) return;
  const char* name = GATE ? "gather_gate" : RECOMP ? "gather_outer_recomp" : "gather_outer";
  if (dm.tune_of<T>().grid_mult <= 0) {
    allow_lds(k_gather<Cf, GATE, false, RECOMP>, lg);
    SUS2_LAUNCH(ws, st, kname<Cf>(name), (k_gather<Cf, GATE, false, RECOMP>), dim3(nall), dim3(WAVE), lg, m, nall,
                types, in_off, gedge_in, uedge, vg, gbar, hbar_out, fw, rec, abar);
    return;
  }
  const unsigned g = tuned_grid(dm.tune_of<T>(), k_gather<Cf, GATE, true, RECOMP>, lg, nall);
  SUS2_LAUNCH(ws, st, kname<Cf>(name), (k_gather<Cf, GATE, true, RECOMP>), dim3(g), dim3(WAVE), lg, m, nall, types,
              in_off, gedge_in, uedge, vg, gbar, hbar_out, fw, rec, abar);
}

template <class Cf>
void run_efs_shape(const DevModelSet& dm, Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  const int N = b.natoms;
  build_neighbors<T>(ws, b, dm.host->rc, st);
  const size_t n1 = std::max(1, N), ne = std::max<long>(1, ws.nedges);
  ws.h.ensure(sizeof(T) * n1 * Cf::F);
  ws.abar.ensure(sizeof(T) * n1 * Cf::KH);
  ws.gbar.ensure(sizeof(T) * n1 * Cf::KH);
  ws.vg.ensure(sizeof(T) * n1 * std::max(1, m.gp.node_cnt));
  ws.site.ensure(sizeof(T) * n1);
  ws.fw.ensure(sizeof(T) * n1 * 9);
  ws.gedge_in.ensure(sizeof(T) * ne * 3);
  // gate-adjoint messages: written by the outer edge backward (0), by a
  // messages-only pass after a forces-only one (2), or recomputed by the
  // gather (1, no per-edge message buffer)
  const int mset = dm.tune_of<T>().msg_recompute;
  const int mmode = !m.gated ? 0 : mset >= 0 ? mset : (Cf::LL >= 4 ? 2 : 0);
  const bool recomp = mmode == 1;
  if (m.gated && !recomp) ws.uedge.ensure(sizeof(T) * ne * Cf::F);
  const int* off = ws.nbr_off.get<int>();
  const Rec<T>* rec = ws.rec.get<Rec<T>>();
  T* h = ws.h.get<T>();
  if (N > 0) {
    T* vg = ws.vg.get<T>();
    T* abar = ws.abar.get<T>();
    T* gbar = ws.gbar.get<T>();
    T* fw = ws.fw.get<T>();
    T* gi = ws.gedge_in.get<T>();
    T* ue = ws.uedge.get<T>();
    const int* ip = ws.in_pos.get<int>();
    if (m.gated) launch_pass_forward<Cf, true>(dm, ws, N, b.types, off, rec, h, vg, ws.site.get<T>(), abar, st);
    launch_pass_forward<Cf, false>(dm, ws, N, b.types, off, rec, h, vg, ws.site.get<T>(), abar, st);
    if (recomp) {
      launch_edge_backward<Cf, false, false>(dm, ws, N, b.types, off, rec, h, abar, ip, gi, ue, fw, st);
      launch_gather<Cf, false, true>(dm, ws, N, b.types, off, gi, ue, vg, gbar, (T*)nullptr, fw, st, rec, abar);
    } else if (mmode == 2) {
      launch_edge_backward<Cf, false, false>(dm, ws, N, b.types, off, rec, h, abar, ip, gi, ue, fw, st);
      launch_edge_messages<Cf>(dm, ws, N, b.types, off, rec, abar, ip, ue, st);
      launch_gather<Cf, false>(dm, ws, N, b.types, off, gi, ue, vg, gbar, (T*)nullptr, fw, st);
    } else {
      launch_edge_backward<Cf, false>(dm, ws, N, b.types, off, rec, h, abar, ip, gi, ue, fw, st);
      launch_gather<Cf, false>(dm, ws, N, b.types, off, gi, ue, vg, gbar, (T*)nullptr, fw, st);
    }
    if (m.gated) {
      launch_edge_backward<Cf, true>(dm, ws, N, b.types, off, rec, h, gbar, ip, gi, ue, fw, st);
      launch_gather<Cf, true>(dm, ws, N, b.types, off, gi, ue, vg, gbar, (T*)nullptr, fw, st);
    }
  }
  finish_efs<T>(ws, b, out, SiteConst{dm.d.view.theta, m.o_species, m.o_shift, m.C}, st);
}

// ------------------------------------------------------------------ pair-style stages
// Owned atoms [0, nowned) are centres; ghosts [nowned, nall) appear only as
// neighbours.  ws.nbr_off / ws.rec hold the owned atoms' edges (CSR), pv.in_off
// / ws.in_edges the incoming lists of all atoms.  The stages follow
// Kernel::begin / outer / finish (interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp:
// 385-464, 487-546, 549-562).
template <class Cf>
void pair_stage_shape(int stage, const DevModelSet& dm, Workspace& ws, const PairView& pv, hipStream_t st) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  const int n = pv.nowned, all = pv.nall;
  const size_t a1 = std::max(1, all), ne = std::max<long>(1, ws.nedges);
  const int* off = ws.nbr_off.get<int>();
  const Rec<T>* rec = ws.rec.get<Rec<T>>();
  if (stage == 0) {  // begin: gate pass over the owned atoms; ghost gate signals arrive by forward comm
    ws.h.ensure(sizeof(T) * a1 * Cf::F);
    ws.hbar.ensure(sizeof(T) * a1 * Cf::F);
    ws.abar.ensure(sizeof(T) * a1 * Cf::KH);
    ws.gbar.ensure(sizeof(T) * a1 * Cf::KH);
    ws.vg.ensure(sizeof(T) * a1 * std::max(1, m.gp.node_cnt));
    ws.site.ensure(sizeof(T) * a1);
    ws.fw.ensure(sizeof(T) * a1 * 9);
    ws.gedge_in.ensure(sizeof(T) * ne * 3);
    if (m.gated) ws.uedge.ensure(sizeof(T) * ne * Cf::F);
    if (m.gated && all > 0) {
      fill_zero(ws, ws.h.get<T>(), (size_t)all * Cf::F, st);
      launch_pass_forward<Cf, true>(dm, ws, n, pv.types, off, rec, ws.h.get<T>(), ws.vg.get<T>(), ws.site.get<T>(),
                                    ws.abar.get<T>(), st);
    }
    return;
  }
  if (stage == 1) {  // outer: outer pass, its coordinate reverse, forces and gate adjoints of all atoms
    if (all > 0) fill_zero(ws, ws.fw.get<T>(), (size_t)all * 9, st);
    launch_pass_forward<Cf, false>(dm, ws, n, pv.types, off, rec, ws.h.get<T>(), ws.vg.get<T>(), ws.site.get<T>(),
                                   ws.abar.get<T>(), st);
    launch_edge_backward<Cf, false>(dm, ws, n, pv.types, off, rec, ws.h.get<T>(), ws.abar.get<T>(),
                                    ws.in_pos.get<int>(), ws.gedge_in.get<T>(), ws.uedge.get<T>(), ws.fw.get<T>(), st);
    launch_gather<Cf, false>(dm, ws, all, pv.types, pv.in_off, ws.gedge_in.get<T>(), ws.uedge.get<T>(),
                             ws.vg.get<T>(), ws.gbar.get<T>(), m.gated ? ws.hbar.get<T>() : (T*)nullptr,
                             ws.fw.get<T>(), st);
    return;
  }
  // finish: gate reverse of the owned atoms (their adjoints now include the
  // ghosts' contributions), gate coordinate reverse, forces, reductions
  if (m.gated && n > 0) {
    const size_t lg = gather_lds<Cf>(m);
    allow_lds(k_gate_reverse<Cf>, lg);
    SUS2_LAUNCH(ws, st, kname<Cf>("gate_reverse"), (k_gate_reverse<Cf>), dim3(n), dim3(WAVE), lg, m,
                ws.hbar.get<T>(), ws.vg.get<T>(), ws.gbar.get<T>());
    launch_edge_backward<Cf, true>(dm, ws, n, pv.types, off, rec, ws.h.get<T>(), ws.gbar.get<T>(),
                                   ws.in_pos.get<int>(), ws.gedge_in.get<T>(), ws.uedge.get<T>(), ws.fw.get<T>(), st);
    launch_gather<Cf, true>(dm, ws, all, pv.types, pv.in_off, ws.gedge_in.get<T>(), ws.uedge.get<T>(),
                            ws.vg.get<T>(), ws.gbar.get<T>(), (T*)nullptr, ws.fw.get<T>(), st);
  }
  Batch b;
  b.nframes = 1;
  b.natoms = all;
  b.frame_off = pv.frame_off;  // {0, nowned}: energy and virial of the owned atoms
  b.types = pv.types;
  EfsOut out;
  out.energy = pv.energy;
  out.force = pv.force;
  out.virial = pv.virial;
  out.site = pv.site;
  ws.max_frame_atoms = n;
  finish_efs<T>(ws, b, out, SiteConst{dm.d.view.theta, m.o_species, m.o_shift, m.C}, st, n);
}

}  // namespace sus2
