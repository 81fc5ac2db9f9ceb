// Shape-templated parameter-gradient kernels of the SUS2-SH path; instantiated
// per shape in sus2_efs_l*k*.hip beside the EFS kernels.
//
// What is computed (AccumulateEFSCombinationGrad, dev_src/mtpr.cpp:9188 and
// src/basic_mlip.cpp:421-446):
//   grad += d/dtheta [ sum_i Ebar_i E_i  +  D_w E ],
// where D_w E = sum_e w_e . dE/dd_e is the derivative of the total energy
// along the edge displacement field w_e = frc_w_i - frc_w_j + str_w(frame) d_e
// for every directed edge e = (i -> j).  This is forward-over-reverse: every
// pass carries primal values and their tangent along w (moments A and A_dot,
// graph nodes v and v_dot, gate signal h and h_dot), and the reverse sweep
// carries two adjoints per quantity, P = dPhi/dX and Q = dPhi/dX_dot.  Per-edge
// quantities are recomputed in every sweep; nothing is cached per edge except
// the gate messages of the outer pass (2F values per edge), which the gate
// reverse gathers over the sorted incoming-edge lists (no atomics).
//
// Kernels (one 64-lane workgroup = one wavefront, LDS-staged edges):
//   k_grad_gate_forward   per atom: gate moments G, G_dot; gate graph forward +
//                         tangent; h = W S, h_dot = W S_dot; stores the gate
//                         node values / tangents and the type-free moments.
//   k_grad_outer          grid-stride over atoms: outer moments + tangent,
//                         graph forward + tangent, readout, graph reverse of
//                         the tangent program (P, Q), then a second edge sweep
//                         producing the radial / type / warp / alpha gradients
//                         and the gate messages (h_bar, h_dot_bar parts).
//   k_grad_gate_reverse   grid-stride: gathers h_bar, h_dot_bar, gate graph
//                         reverse, gate edge sweep (gate radial / type / warp).
//   k_grad_W              W_bar = sum_j h_bar_j S_j^T + h_dot_bar_j S_dot_j^T
//                         (split over atom chunks, FP64 partials).
//   k_grad_reduce         fixed-order FP64 sum of the per-workgroup partials.
// Parameter gradients are accumulated in FP64 (per-lane registers for the
// radial coefficients, an FP64 LDS vector for the rest); per-edge arithmetic
// runs in the kernel precision T.
#pragma once
#include "sus2_efs_impl.hpp"

namespace sus2 {

// Chebyshev values and xi-derivatives of one warp block, scaling folded in:
// Tb[q] = S T_q(xi), T1[q] = dTb/dxi, T2[q] = d2Tb/dxi2.
template <int BB, class T>
__device__ __forceinline__ void cheb3(T S, T xi, T* Tb, T* T1, T* T2) {
  Tb[0] = S; T1[0] = T(0); T2[0] = T(0);
  if (BB > 1) { Tb[1] = S * xi; T1[1] = S; T2[1] = T(0); }
#pragma unroll
  for (int q = 2; q < BB; ++q) {
    Tb[q] = T(2) * xi * Tb[q - 1] - Tb[q - 2];
    T1[q] = T(2) * (Tb[q - 1] + xi * T1[q - 1]) - T1[q - 2];
    T2[q] = T(4) * T1[q - 1] + T(2) * xi * T2[q - 1] - T2[q - 2];
  }
}

// Index of a coefficient in the per-workgroup partial vector, which skips the
// gate-weight block W (reduced separately by k_grad_W).
__host__ __device__ __forceinline__ int pidx(int p, int oW, int nW) { return (oW >= 0 && p >= oW) ? p - nW : p; }

// Channel halves of the edge sweeps: with one warp block per channel k
// (scaling-map mode 2) and an even K, every sweep runs over an atom's edges
// twice, for the channels k in [0, K/2) and [K/2, K).  That halves the staged
// LDS rows (more workgroups resident per CU) and the live radial values per
// lane, at the price of the harmonics being evaluated twice.
template <class Cf>
struct Halves {
  static constexpr int NH = (Cf::SM == 2 && Cf::KK % 2 == 0) ? 2 : 1;
  static constexpr int KP = Cf::KK / NH;           // channels k per half
  static constexpr int FP = KP * Cf::L1;           // radial functions per half
  static constexpr int BP = NH == 2 ? KP : Cf::NBLK;  // warp blocks per half
};

// LDS layout shared by host (sizes) and device (carving): the species tables,
// the canonical adjoints Pc, Qc, the type-free moments At, Atd and the gate
// adjoint row, then one region used in turn by the edge staging of sweep 1,
// the graph phase (A, Ad, v, vd, P, Q, scratch) and the staging of sweep 2
// (their lifetimes do not overlap), then the FP64 partial vector.
template <class Cf>
struct GradLds {
  int stage_rows, M, sc, tabs, pn, C;
  // per-species edge masks of a staged batch (64-bit words, in units of T)
  __host__ __device__ size_t mask_words() const { return (size_t)C * 8 / sizeof(typename Cf::T) + 2; }
  __host__ __device__ size_t graph_floats() const {
    return 2 * (size_t)Cf::KH + 4 * (size_t)M + 2 * (size_t)sc;
  }
  __host__ __device__ size_t region() const {
    const size_t st = (size_t)stage_rows * LR, gr = graph_floats();
    return st > gr ? st : gr;
  }
  __host__ __device__ size_t floats() const {
    return (size_t)tabs + 4 * (size_t)Cf::KH + 2 * Cf::F + 2 + mask_words() + region();
  }
  __host__ __device__ size_t bytes() const {
    // T area first, then the FP64 partial vector (8-byte aligned)
    const size_t t = floats() * sizeof(typename Cf::T);
    return ((t + 7) & ~(size_t)7) + sizeof(double) * (size_t)pn;
  }
};

template <class Cf>
__host__ __device__ inline int grad_stage_rows(bool sweep2, bool outer_gated) {
  using Hv = Halves<Cf>;
  const int r1 = 2 * Hv::FP + 2 * Cf::H;
  const int r2 = 4 * Hv::BP + (outer_gated ? 4 : 3) * Hv::FP;
  return sweep2 ? (r1 > r2 ? r1 : r2) : r1;
}

template <class Cf>
GradLds<Cf> grad_lds(const DModel<typename Cf::T>& m, bool gate, bool sweep2, int pn) {
  GradLds<Cf> g;
  g.stage_rows = grad_stage_rows<Cf>(sweep2, !gate && m.gated);
  g.M = m.M;
  g.sc = std::max(graph_scratch(gate ? m.gp : m.op), 1);
  g.tabs = species_tables_size<Cf>(m.C);
  g.pn = pn;
  g.C = m.C;
  return g;
}

// Pointers into the LDS block.
template <class T>
struct GradSmem {
  T *stage, *At, *Atd, *A, *Ad, *Pc, *Qc, *v, *vd, *P, *Q, *s1, *s2, *tabs, *hbs;
  unsigned long long* tmask;  // [C]
  double* part;
};

template <class Cf>
__device__ GradSmem<typename Cf::T> carve(const GradLds<Cf>& g, unsigned char* raw) {
  using T = typename Cf::T;
  GradSmem<T> s;
  T* p = reinterpret_cast<T*>(raw);
  s.tabs = p; p += g.tabs;
  s.Pc = p; p += Cf::KH;
  s.Qc = p; p += Cf::KH;
  s.At = p; p += Cf::KH;
  s.Atd = p; p += Cf::KH;
  s.hbs = p; p += 2 * Cf::F + 2;
  {  // 8-byte aligned mask words
    unsigned char* q = reinterpret_cast<unsigned char*>(p);
    const size_t off = (size_t)(q - raw);
    s.tmask = reinterpret_cast<unsigned long long*>(raw + ((off + 7) & ~(size_t)7));
    p += g.mask_words();
  }
  T* region = p;
  s.stage = region;  // edge staging (sweeps 1 and 2)
  s.A = region;      // graph phase
  s.Ad = s.A + Cf::KH;
  s.v = s.Ad + Cf::KH;
  s.vd = s.v + g.M;
  s.P = s.vd + g.M;
  s.Q = s.P + g.M;
  s.s1 = s.Q + g.M;
  s.s2 = s.s1 + g.sc;
  const size_t t = g.floats() * sizeof(T);
  s.part = reinterpret_cast<double*>(raw + ((t + 7) & ~(size_t)7));
  return s;
}

// ------------------------------------------------------------------ graph with tangents
// Contraction of A and A_dot into v and v_dot (both zeroed first).
template <class T, int KK>
__device__ void wave_contract_tan(const DModel<T>& m, const DPass<T>& p, bool gate, const T* A, const T* Ad, T* v,
                                  T* vd) {
  for (int n = threadIdx.x; n < m.M; n += WAVE) { v[n] = T(0); vd[n] = T(0); }
  __syncthreads();
  wave_contract<T, KK>(m, p, gate, A, v);
  wave_contract<T, KK>(m, p, gate, Ad, vd);
}

// Product layers: v[out] = sum c v[a] v[b];  v_dot[out] = sum c (v_dot[a] v[b] + v[a] v_dot[b]).
template <class T>
__device__ void wave_graph_forward_tan(const DPass<T>& p, T* v, T* vd, T* s1, T* s2) {
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
      on = p.f_out[oa];
    }
    wave_rows(r0, r1, [&](int r) {
      const GRow<T> g = p.f_row[r];
      const T c = g.c;
      const int a = g.x, b = g.y;
      s1[r - r0] = c * v[a] * v[b];
      s2[r - r0] = c * (vd[a] * v[b] + v[a] * vd[b]);
    });
    __syncthreads();
    if (oa < o1) {
      T x = T(0), y = T(0);
      for (int r = ob - r0; r < oe - r0; ++r) { x += s1[r]; y += s2[r]; }
      v[on] = x;
      vd[on] = y;
    }
    for (int o = oa + WAVE; o < o1; o += WAVE) {
      T x = T(0), y = T(0);
      const int e = p.f_roff[o + 1] - r0;
      for (int r = p.f_roff[o] - r0; r < e; ++r) { x += s1[r]; y += s2[r]; }
      v[p.f_out[o]] = x;
      vd[p.f_out[o]] = y;
    }
    __syncthreads();
  }
}

// Reverse of the tangent program: for every consumer row (out, partner, c) of
// node n,  Q[n] += c Q[out] v[partner],
//          P[n] += c (P[out] v[partner] + Q[out] v_dot[partner]).
template <class T>
__device__ void wave_graph_reverse_tan(const DPass<T>& p, const T* v, const T* vd, T* P, T* Q, T* s1, T* s2) {
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
      kn = p.r_node[ka];
    }
    wave_rows(q0, q1, [&](int q) {
      const GRow<T> g = p.r_row[q];
      const T c = g.c;
      const int o = g.x, pt = g.y;
      s1[q - q0] = c * (P[o] * v[pt] + Q[o] * vd[pt]);
      s2[q - q0] = c * Q[o] * v[pt];
    });
    __syncthreads();
    if (ka < n1) {
      T x = P[kn], y = Q[kn];
      for (int q = kb - q0; q < ke - q0; ++q) { x += s1[q]; y += s2[q]; }
      P[kn] = x;
      Q[kn] = y;
    }
    for (int k = ka + WAVE; k < n1; k += WAVE) {
      const int n = p.r_node[k];
      T x = P[n], y = Q[n];
      const int e = p.r_roff[k + 1] - q0;
      for (int q = p.r_roff[k] - q0; q < e; ++q) { x += s1[q]; y += s2[q]; }
      P[n] = x;
      Q[n] = y;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ sweep 1
// Moments and tangents of one atom for the channels k of half HALF:
// At = sum_e x1 R Y, Atd = sum_e [x1 (R' w_r Y + R grad Y . w) + x1_dot R Y]
// (the centre factor x0 is applied afterwards: A = x0 At), with x1 =
// neighbour type factor (+ alpha h_j for the outer pass of an Effect model)
// and x1_dot = alpha h_dot_j.  The moment dot products over a staged batch
// use two lanes per moment (one per half of the batch) when a half has at
// most WAVE / 2 moments.
template <class Cf, bool GATE, int HALF>
__device__ void grad_sweep_moments(const DModel<typename Cf::T>& m, const SpeciesTables<typename Cf::T>& tab,
                                   const Rec<typename Cf::T>* __restrict__ rec, const double* __restrict__ frc_w,
                                   const typename Cf::T* Sw, const double* fwi,
                                   const typename Cf::T* __restrict__ h, const typename Cf::T* __restrict__ hd, int ti,
                                   int e0, int e1, typename Cf::T* stage, typename Cf::T* At, typename Cf::T* Atd) {
  using T = typename Cf::T;
  using Hv = Halves<Cf>;
  constexpr int LL = Cf::LL, KK = Cf::KK, L1 = Cf::L1, H = Cf::H, F = Cf::F;
  constexpr int KP = Hv::KP, FP = Hv::FP, K0 = HALF * KP, MU0 = K0 * L1;
  constexpr int B0 = Hv::NH == 2 ? K0 : 0, B1 = Hv::NH == 2 ? K0 + KP : Cf::NBLK;
  constexpr bool ROW_ALIGNED = (MU0 * sizeof(T)) % 16 == 0;
  constexpr int NMH = H * KP;                        // moments of the half
  constexpr int SPLIT = NMH <= WAVE / 2 ? 2 : 1;     // lanes per moment: l and l + WAVE / 2
  constexpr int NMT = SPLIT == 2 ? 1 : (NMH + WAVE - 1) / WAVE;
  const int lane = threadIdx.x;
  const bool hg = !GATE && m.gated;
  T* sS1 = stage;             // [FP][LR]  x1 R
  T* sS2 = sS1 + FP * LR;     // [FP][LR]  x1 R' w_r + x1_dot R
  T* sY = sS2 + FP * LR;      // [H][LR]
  T* sYw = sY + H * LR;       // [H][LR]  grad Y . w
  T acc[NMT], accd[NMT];
#pragma unroll
  for (int t = 0; t < NMT; ++t) { acc[t] = T(0); accd[t] = T(0); }
  for (int base = e0; base < e1; base += WAVE) {
    const int e = base + lane;
    if (e < e1) {
      const Rec<T> q = rec[e];
      const int j = rec_j(q);
      const T dx = q.x, dy = q.y, dz = q.z;
      T r, inv;
      edge_length(dx * dx + dy * dy + dz * dz, r, inv);
      const T ux = dx * inv, uy = dy * inv, uz = dz * inv;
      const int tj = rec_t(q);
      const double* fj = frc_w + 3 * (size_t)j;
      const T wx = T(fwi[0] - fj[0]) + Sw[0] * dx + Sw[1] * dy + Sw[2] * dz;
      const T wy = T(fwi[1] - fj[1]) + Sw[3] * dx + Sw[4] * dy + Sw[5] * dz;
      const T wz = T(fwi[2] - fj[2]) + Sw[6] * dx + Sw[7] * dy + Sw[8] * dz;
      const T wr = ux * wx + uy * wy + uz * wz;
      T R[FP], dR[FP], hj[FP], hdj[FP];
      if (hg) {  // issued before the basis work
        load_row<T, FP, ROW_ALIGNED>(h + (size_t)j * F + MU0, hj);
        load_row<T, FP, ROW_ALIGNED>(hd + (size_t)j * F + MU0, hdj);
      }
      radial_each<Cf, true, B0, B1>(m, GATE, tab, ti * m.C + tj, r, [&](int mu, T v, T dv) {
        if (mu < MU0 || mu >= MU0 + FP) return;
        R[mu - MU0] = v;
        dR[mu - MU0] = dv;
      });
      const T* x0j = tab.x0 + tj * F + MU0;
      const T* alj = tab.alpha + tj * F + MU0;
#pragma unroll
      for (int c = 0; c < FP; ++c) {
        T x1 = x0j[c], xd = T(0);
        if (hg) {
          x1 += alj[c] * hj[c];
          xd = alj[c] * hdj[c];
        }
        sS1[c * LR + lane] = x1 * R[c];
        sS2[c * LR + lane] = x1 * dR[c] * wr + xd * R[c];
      }
      // w . grad_d Y = (w . G - l Y (u . w)) / r
      sh_stream_solid<T, LL>(ux, uy, uz, [&](int l, int mm, T Y, T Gx, T Gy, T Gz) {
        const int hh = hidx(l, mm);
        sY[hh * LR + lane] = Y;
        sYw[hh * LR + lane] = (Gx * wx + Gy * wy + Gz * wz - T(l) * wr * Y) * inv;
      });
    }
    __syncthreads();
    const int cnt = min(WAVE, e1 - base);
#pragma unroll
    for (int t = 0; t < NMT; ++t) {
      const int idx = lane + WAVE * t;
      const int mloc = SPLIT == 2 ? lane % (WAVE / 2) : idx, prt = SPLIT == 2 ? lane / (WAVE / 2) : 0;
      if (mloc < NMH) {
        const int hh = mloc / KP, kk = mloc % KP;
        const int c = kk * L1 + l_of(hh);
        const int lo = (SPLIT == 2 && prt) ? (cnt + 1) / 2 : 0;
        const int hi = (SPLIT == 2 && !prt) ? (cnt + 1) / 2 : cnt;
        const T* p1 = sS1 + c * LR + lo;
        const T* p2 = sS2 + c * LR + lo;
        const T* py = sY + hh * LR + lo;
        const T* pw = sYw + hh * LR + lo;
        acc[t] += lds_dot(p1, py, hi - lo);
        accd[t] += lds_dot(p2, py, hi - lo) + lds_dot(p1, pw, hi - lo);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < NMT; ++t) {
    T a = acc[t], ad = accd[t];
    if constexpr (SPLIT == 2) {  // the two halves of the batches (lanes idx and idx + WAVE / 2)
      a += __shfl_xor(a, WAVE / 2, WAVE);
      ad += __shfl_xor(ad, WAVE / 2, WAVE);
    }
    const int idx = lane + WAVE * t;  // with SPLIT == 2 the lanes below WAVE / 2 write
    if (idx < NMH) {
      const int hh = idx / KP, kk = idx % KP;
      At[hh * KK + K0 + kk] = a;
      Atd[hh * KK + K0 + kk] = ad;
    }
  }
}

// ------------------------------------------------------------------ sweep 2
// Parameter sweep over the centre's edges with the canonical adjoints Pc, Qc:
// per edge  p_mu = sum_m Pc Y, q_mu = sum_m Qc Y, q'_mu = sum_m Qc grad Y . w,
//   Z = R (p + q') + R' w_r q,   Gv = x0 (x1 (p + q') + x1_dot q),  Gd = x0 x1 w_r q
// and the radial coefficient, type, warp (slope/shift) and alpha gradients
// follow from R = f C0, R' = g C0 + f C1 xi_r with C_k = sum_b c_b d^k Tb/dxi^k.
// Outer pass of an Effect model: messages (alpha x0 Z, alpha x0 R q) per edge.
template <class Cf, bool GATE, int HALF>
__device__ void grad_sweep_params(const DModel<typename Cf::T>& m, const SpeciesTables<typename Cf::T>& tab,
                                  const Rec<typename Cf::T>* __restrict__ rec,
                                  const double* __restrict__ frc_w, const typename Cf::T* Sw, const double* fwi,
                                  const typename Cf::T* __restrict__ h, const typename Cf::T* __restrict__ hd,
                                  typename Cf::T* __restrict__ msg, const int* __restrict__ in_pos, int ti,
                                  int e0, int e1,
                                  const GradSmem<typename Cf::T>& s, double* creg, int pn_oW, int pn_nW) {
  using T = typename Cf::T;
  using Hv = Halves<Cf>;
  constexpr int LL = Cf::LL, KK = Cf::KK, L1 = Cf::L1, F = Cf::F, BB = Cf::BB, NBLK = Cf::NBLK;
  constexpr int KP = Hv::KP, FP = Hv::FP, BP = Hv::BP, K0 = HALF * KP, MU0 = K0 * L1;
  constexpr int B0 = Hv::NH == 2 ? K0 : 0;  // first warp block of the half
  const int lane = threadIdx.x, C = m.C, CC = C * C;
  const bool hg = !GATE && m.gated;
  // staged rows of the half: blocks b - B0, channels mu - MU0
  T* sXi = s.stage;             // [BP][LR]
  T* sXr = sXi + BP * LR;       // [BP][LR]
  T* sDs = sXr + BP * LR;       // [BP][LR]
  T* sDt = sDs + BP * LR;       // [BP][LR]
  T* sCv = sDt + BP * LR;       // [FP][LR]
  T* sCd = sCv + FP * LR;       // [FP][LR]
  T* sX1 = sCd + FP * LR;       // [FP][LR]  x0 Z  (d Phi / d x1)
  T* sAl = sX1 + FP * LR;       // [FP][LR]  x0 (Z h_j + R q h_dot_j)  (outer, gated)
  constexpr int CR = coef_row<Cf>();
  const int so = (GATE && m.o_gscal >= 0) ? m.o_gscal : m.o_scal;
  const T* x0 = tab.x0 + ti * F;
  // c-gradient lanes: (channel of the half, subset)
  constexpr int NSUB = WAVE / FP;
  const int cmu = lane % FP, csub = lane / FP;
  const bool cact = csub < NSUB;
  const int cblk = Hv::NH == 2 ? cmu / L1 : Cf::blk(cmu / L1, cmu % L1);
  for (int base = e0; base < e1; base += WAVE) {
    const int e = base + lane;
    int lane_t = -1;  // species of this lane's edge (none past the list)
    if (e < e1) {
      const Rec<T> q = rec[e];
      const int j = rec_j(q);
      const T dx = q.x, dy = q.y, dz = q.z;
      T r, inv;
      edge_length(dx * dx + dy * dy + dz * dz, r, inv);
      const T ux = dx * inv, uy = dy * inv, uz = dz * inv;
      const int tj = rec_t(q);
      lane_t = tj;
      const double* fj = frc_w + 3 * (size_t)j;
      const T wx = T(fwi[0] - fj[0]) + Sw[0] * dx + Sw[1] * dy + Sw[2] * dz;
      const T wy = T(fwi[1] - fj[1]) + Sw[3] * dx + Sw[4] * dy + Sw[5] * dz;
      const T wz = T(fwi[2] - fj[2]) + Sw[6] * dx + Sw[7] * dy + Sw[8] * dz;
      const T wr = ux * wx + uy * wy + uz * wz;
      const T d0 = r - m.rc, f0 = d0 * d0, g0 = T(2) * d0;
      const bool act = r < m.rca;
      const T d1 = r - m.rca, f1 = act ? d1 * d1 : T(0), g1 = act ? T(2) * d1 : T(0);
      const int pair = ti * C + tj;
      const T* x0j = tab.x0 + tj * F;
      const T* alj = tab.alpha + tj * F;
      // the radial coefficient rows come from LDS; a zero the compiler cannot
      // see through (an empty asm of the edge's distance) keeps their reads in
      // the edge loop instead of holding all of them in registers
      int zc;
      asm("" : "=v"(zc) : "0"(0), "v"(r));
      // per-channel work of one radial function mu = (k, l): p, q, q' from the
      // harmonics, the radial values and their coefficient / warp gradients;
      // hj, hdj: the neighbour's gate signal and its tangent for mu (gated)
      auto channel = [&](int mu, int k, int l, const T* Yv, const T* Ywv, const T* Tb, const T* T1, const T* T2,
                         T xr, T xs, T xt, T xrs, T xrt, T hj, T hdj, T& ds, T& dt) {
        T pp = T(0), qv = T(0), qw = T(0);
#pragma unroll
        for (int mm = -l; mm <= l; ++mm) {
          const int hh = hidx(l, mm);
          const T pc = s.Pc[hh * KK + k], qc = s.Qc[hh * KK + k];
          pp += pc * Yv[hh];
          qv += qc * Yv[hh];
          qw += qc * Ywv[hh];
        }
        T c[CR];
        load_row<T, CR>(tab.coef + (zc + mu * CR), c);
        T C0 = T(0), C1 = T(0), C2 = T(0);
#pragma unroll
        for (int qq = 0; qq < BB; ++qq) {
          C0 += c[qq] * Tb[qq];
          C1 += c[qq] * T1[qq];
          C2 += c[qq] * T2[qq];
        }
        const T f = l == 0 ? f0 : f1, g = l == 0 ? g0 : g1;
        const T R = f * C0, Rp = g * C0 + f * C1 * xr;
        T x1 = x0j[mu], xd = T(0);
        if (hg) {
          x1 += alj[mu] * hj;
          xd = alj[mu] * hdj;
        }
        const T pq = pp + qw;
        const T Z = R * pq + Rp * wr * qv;
        const T Rq = R * qv;
        const T Gv = x0[mu] * (x1 * pq + xd * qv);
        const T Gd = x0[mu] * x1 * wr * qv;
        const T Cv = Gv * f + Gd * g, Cd = Gd * f;
        ds += xs * C1 * Cv + Cd * (C2 * xs * xr + C1 * xrs);
        dt += xt * C1 * Cv + Cd * (C2 * xt * xr + C1 * xrt);
        sCv[(mu - MU0) * LR + lane] = Cv;
        sCd[(mu - MU0) * LR + lane] = Cd;
        sX1[(mu - MU0) * LR + lane] = x0[mu] * Z;
        if (hg) {
          sAl[(mu - MU0) * LR + lane] = x0[mu] * (Z * hj + Rq * hdj);
          T* mo = msg + (size_t)in_pos[e] * 2 * F;
          mo[mu] = alj[mu] * x0[mu] * Z;
          mo[F + mu] = alj[mu] * x0[mu] * Rq;
        }
      };
      auto block_basis = [&](int b, T* Tb, T* T1, T* T2, T& xi, T& xr, T& xs, T& xt, T& xrs, T& xrt) {
        const T sl = tab.slope[b * CC + pair], sh = tab.shift[b * CC + pair];
        xi = dtanh(sl * (r - sh) * T(0.5));
        const T xx = T(1) - xi * xi;
        xr = xx * sl * T(0.5); xs = xx * (r - sh) * T(0.5); xt = -xr;
        xrs = T(0.5) * xx - sl * xi * xs; xrt = -sl * xi * xt;
        cheb3<BB>(m.scaling, xi, Tb, T1, T2);
      };
      T Yv[Cf::H], Ywv[Cf::H];
      // w . grad_d Y = (w . G - l Y (u . w)) / r
      sh_stream_solid<T, LL>(ux, uy, uz, [&](int l, int mm, T Y, T Gx, T Gy, T Gz) {
        const int hh = hidx(l, mm);
        Yv[hh] = Y;
        Ywv[hh] = (Gx * wx + Gy * wy + Gz * wz - T(l) * wr * Y) * inv;
      });
      if constexpr (Cf::SM == 2) {
        // blocks are the channels k of the half: one block's basis live at a time
#pragma unroll 1
        for (int k = K0; k < K0 + KP; ++k) {
          // the block's channels of the neighbour's gate row, loaded together
          // before the basis work
          T hk[L1], hdk[L1];
#pragma unroll
          for (int l = 0; l < L1; ++l) {
            hk[l] = hg ? h[(size_t)j * F + k * L1 + l] : T(0);
            hdk[l] = hg ? hd[(size_t)j * F + k * L1 + l] : T(0);
          }
          T Tb[BB], T1[BB], T2[BB], xi, xr, xs, xt, xrs, xrt;
          block_basis(k, Tb, T1, T2, xi, xr, xs, xt, xrs, xrt);
          T ds = T(0), dt = T(0);
#pragma unroll
          for (int l = 0; l < L1; ++l)
            channel(k * L1 + l, k, l, Yv, Ywv, Tb, T1, T2, xr, xs, xt, xrs, xrt, hk[l], hdk[l], ds, dt);
          sXi[(k - B0) * LR + lane] = xi;
          sXr[(k - B0) * LR + lane] = xr;
          sDs[(k - B0) * LR + lane] = ds;
          sDt[(k - B0) * LR + lane] = dt;
        }
      } else {
#pragma unroll
        for (int b = 0; b < NBLK; ++b) {
          T Tb[BB], T1[BB], T2[BB], xi, xr, xs, xt, xrs, xrt;
          block_basis(b, Tb, T1, T2, xi, xr, xs, xt, xrs, xrt);
          T ds = T(0), dt = T(0);
#pragma unroll
          for (int k = 0; k < KK; ++k)
#pragma unroll
            for (int l = 0; l < L1; ++l)
              if (Cf::blk(k, l) == b)
                channel(k * L1 + l, k, l, Yv, Ywv, Tb, T1, T2, xr, xs, xt, xrs, xrt,
                        hg ? h[(size_t)j * F + k * L1 + l] : T(0), hg ? hd[(size_t)j * F + k * L1 + l] : T(0), ds,
                        dt);
          sXi[b * LR + lane] = xi;
          sXr[b * LR + lane] = xr;
          sDs[b * LR + lane] = ds;
          sDt[b * LR + lane] = dt;
        }
      }
    }
    for (int t = 0; t < C; ++t) {  // which staged edges have species t
      const unsigned long long mk = __ballot(lane_t == t);
      if (lane == 0) s.tmask[t] = mk;
    }
    __syncthreads();
    const int cnt = min(WAVE, e1 - base);
    // radial coefficients: lane (mu, subset) re-runs the Chebyshev recurrence
    // over its share of the batch (at most WAVE / NSUB edges, summed in T),
    // then adds the batch sums to its FP64 registers
    if (cact) {
      const T S = m.scaling;
      T cb[BB];
#pragma unroll
      for (int qq = 0; qq < BB; ++qq) cb[qq] = T(0);
      for (int ee = csub; ee < cnt; ee += NSUB) {
        const T xi = sXi[cblk * LR + ee], cv = sCv[cmu * LR + ee], cd = sCd[cmu * LR + ee] * sXr[cblk * LR + ee];
        T ta = S, tb = S * xi, da = T(0), db = S;
        cb[0] += cv * ta;
        if (BB > 1) cb[1] += cv * tb + cd * db;
#pragma unroll
        for (int qq = 2; qq < BB; ++qq) {
          const T tn = T(2) * xi * tb - ta, dn = T(2) * (tb + xi * db) - da;
          cb[qq] += cv * tn + cd * dn;
          ta = tb; tb = tn; da = db; db = dn;
        }
      }
#pragma unroll
      for (int qq = 0; qq < BB; ++qq) creg[qq] += (double)cb[qq];
    }
    // type-bucketed sums: neighbour type factor, alpha, warp slope/shift; each
    // lane walks the staged edges of its species in edge order (mask bits)
    const int n1 = C * FP, n2 = hg ? C * FP : 0, n3 = 2 * BP * C;
    for (int idx = lane; idx < n1 + n2 + n3; idx += WAVE) {
      const T* src;
      int p, t;
      if (idx < n1 + n2) {
        const int k2 = idx < n1 ? idx : idx - n1;
        t = k2 / FP;
        const int c = k2 % FP, mu = MU0 + c;
        src = (idx < n1 ? sX1 : sAl) + c * LR;
        if (idx < n1) p = GATE ? m.o_gtype + t * F + mu : m.o_rad + mu * (BB + C) + BB + t;
        else p = m.o_alpha + t * F + mu;
      } else {
        const int k3 = idx - n1 - n2;
        const int bl = k3 / (2 * C), w = (k3 / C) % 2, b = B0 + bl;
        t = k3 % C;
        src = (w ? sDt : sDs) + bl * LR;
        p = so + (2 * b + w) * CC + ti * C + t;
      }
      unsigned long long mk = s.tmask[t];
      double acc = 0;
      while (mk) {
        const int k = __ffsll((long long)mk) - 1;
        acc += (double)src[k];
        mk &= mk - 1;
      }
      s.part[pidx(p, pn_oW, pn_nW)] += acc;
    }
    __syncthreads();
  }
}

// Per-atom, moment-space gradients after the reverse: rank weights
// (sum_m P[rank node] A + Q[rank node] A_dot) and the centre type factor
// (sum_m Pc At + Qc Atd).  Also scales At -> A beforehand (done by caller).
template <class Cf, bool GATE>
__device__ void grad_moment_space(const DModel<typename Cf::T>& m, const DPass<typename Cf::T>& p,
                                  const GradSmem<typename Cf::T>& s, int ti, int pn_oW, int pn_nW) {
  using T = typename Cf::T;
  constexpr int KK = Cf::KK, L1 = Cf::L1, F = Cf::F, BB = Cf::BB;
  const int lane = threadIdx.x, H = Cf::H, C = m.C;
  const int rank = p.rank;
  const int nr = L1 * rank * KK;
  for (int idx = lane; idx < nr; idx += WAVE) {
    const int l = idx / (rank * KK), a = (idx / KK) % rank, k = idx % KK;
    double acc = 0;
    for (int mm = -l; mm <= l; ++mm) {
      const int c = hidx(l, mm) * KK + k;
      if (m.basic_node[c] < 0) continue;
      const int node = rank_node(p.begin, rank, H, l, mm, a);
      acc += (double)(s.P[node] * s.A[c] + s.Q[node] * s.Ad[c]);
    }
    const int pp = (GATE ? m.o_gw : m.o_ow) + (l * rank + a) * KK + k;
    s.part[pidx(pp, pn_oW, pn_nW)] += acc;
  }
  for (int mu = lane; mu < F; mu += WAVE) {
    const int k = mu / L1, l = mu % L1;
    double acc = 0;
    for (int mm = -l; mm <= l; ++mm) {
      const int c = hidx(l, mm) * KK + k;
      acc += (double)(s.Pc[c] * s.At[c] + s.Qc[c] * s.Atd[c]);
    }
    const int pp = GATE ? m.o_gtype + ti * F + mu : m.o_rad + mu * (BB + C) + BB + ti;
    s.part[pidx(pp, pn_oW, pn_nW)] += acc;
  }
}

// A = x0 At, A_dot = x0 Atd (centre type factor per channel).
template <class Cf>
__device__ void scale_moments(const SpeciesTables<typename Cf::T>& tab, int ti, const GradSmem<typename Cf::T>& s) {
  constexpr int KK = Cf::KK, L1 = Cf::L1, F = Cf::F, KH = Cf::KH;
  for (int c = threadIdx.x; c < KH; c += WAVE) {
    const int hh = c / KK, k = c % KK;
    const auto x = tab.x0[ti * F + k * L1 + l_of(hh)];
    s.A[c] = x * s.At[c];
    s.Ad[c] = x * s.Atd[c];
  }
}

// Per-atom frame weights: str_w of the atom's frame (T) and frc_w of the atom.
template <class T>
__device__ void load_atom_weights(const GradWeights& gw, int i, T* Sw, double* fwi) {
  const int f = gw.atom_frame[i];
#pragma unroll
  for (int q = 0; q < 9; ++q) Sw[q] = T(gw.str_w[9 * (size_t)f + q]);
#pragma unroll
  for (int q = 0; q < 3; ++q) fwi[q] = gw.frc_w[3 * (size_t)i + q];
}

// ------------------------------------------------------------------ kernels
template <class Cf, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE) void k_grad_gate_forward(DModel<T> m, GradLds<Cf> L, GradWeights gw,
                                                            const int32_t* __restrict__ types,
                                                            const int* __restrict__ nbr_off,
                                                            const Rec<T>* __restrict__ rec, T* __restrict__ h,
                                                            T* __restrict__ hd, T* __restrict__ vg,
                                                            T* __restrict__ vgd, T* __restrict__ gmom,
                                                            T* __restrict__ gmomd) {
  constexpr int F = Cf::F, KH = Cf::KH;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const GradSmem<T> s = carve(L, smem_raw);
  const int i = blockIdx.x, lane = threadIdx.x;
  const int ti = types[i];
  SpeciesTables<T> tab;
  load_species_tables<Cf>(m, true, s.tabs, tab);
  T Sw[9];
  double fwi[3];
  load_atom_weights(gw, i, Sw, fwi);
  __syncthreads();
  grad_sweep_moments<Cf, true, 0>(m, tab, rec, gw.frc_w, Sw, fwi, nullptr, nullptr, ti, nbr_off[i], nbr_off[i + 1],
                                  s.stage, s.At, s.Atd);
  if constexpr (Halves<Cf>::NH > 1)
    grad_sweep_moments<Cf, true, 1>(m, tab, rec, gw.frc_w, Sw, fwi, nullptr, nullptr, ti, nbr_off[i],
                                    nbr_off[i + 1], s.stage, s.At, s.Atd);
  __syncthreads();
  for (int c = lane; c < KH; c += WAVE) {
    gmom[(size_t)i * KH + c] = s.At[c];
    gmomd[(size_t)i * KH + c] = s.Atd[c];
  }
  scale_moments<Cf>(tab, ti, s);
  __syncthreads();
  const DPass<T>& p = m.gp;
  wave_contract_tan<T, Cf::KK>(m, p, true, s.A, s.Ad, s.v, s.vd);
  __syncthreads();
  wave_graph_forward_tan(p, s.v, s.vd, s.s1, s.s2);
  for (int mu = lane; mu < F; mu += WAVE) {
    T a = T(0), b = T(0);
    for (int q = 0; q < m.Qg; ++q) {
      const T w = m.W(mu, q);
      a += w * s.v[m.gate_nodes[q]];
      b += w * s.vd[m.gate_nodes[q]];
    }
    h[(size_t)i * F + mu] = a;
    hd[(size_t)i * F + mu] = b;
  }
  for (int n = lane; n < p.node_cnt; n += WAVE) {
    vg[(size_t)i * p.node_cnt + n] = s.v[p.node_lo + n];
    vgd[(size_t)i * p.node_cnt + n] = s.vd[p.node_lo + n];
  }
}

// Zero the workgroup's FP64 partial vector.
__device__ __forceinline__ void zero_part(double* part, int pn) {
  for (int k = threadIdx.x; k < pn; k += WAVE) part[k] = 0.0;
}

// Fold the per-lane radial-coefficient registers (creg[half][b]) into the
// partial vector (subsets one after another), then write the partial row.
template <class Cf>
__device__ void flush_partial(const DModel<typename Cf::T>& m, bool gate, const double* creg, double* part, int pn,
                              int oW, int nW, double* __restrict__ out_row) {
  using Hv = Halves<Cf>;
  constexpr int FP = Hv::FP, BB = Cf::BB, NSUB = WAVE / FP;
  const int lane = threadIdx.x, cmu = lane % FP, csub = lane / FP;
  const bool own_rad = gate && m.o_grad >= 0;
  const int base = own_rad ? m.o_grad : m.o_rad, stride = own_rad ? BB : BB + m.C;
#pragma unroll
  for (int hf = 0; hf < Hv::NH; ++hf) {
    const int mu = hf * FP + cmu;
    for (int sub = 0; sub < NSUB; ++sub) {
      if (csub == sub)
#pragma unroll
        for (int b = 0; b < BB; ++b) part[pidx(base + mu * stride + b, oW, nW)] += creg[hf * BB + b];
      __syncthreads();
    }
  }
  for (int k = lane; k < pn; k += WAVE) out_row[k] = part[k];
}

template <class Cf, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE) void k_grad_outer(DModel<T> m, GradLds<Cf> L, GradWeights gw, int natoms,
                                                     const int32_t* __restrict__ types,
                                                     const int* __restrict__ nbr_off, const Rec<T>* __restrict__ rec,
                                                     const T* __restrict__ h, const T* __restrict__ hd,
                                                     const int* __restrict__ in_pos, T* __restrict__ msg,
                                                     double* __restrict__ partial, int oW, int nW) {
  constexpr int KK = Cf::KK, BB = Cf::BB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const GradSmem<T> s = carve(L, smem_raw);
  const int lane = threadIdx.x;
  SpeciesTables<T> tab;
  load_species_tables<Cf>(m, false, s.tabs, tab);
  zero_part(s.part, L.pn);
  constexpr int NH = Halves<Cf>::NH;
  double creg[NH * BB];
#pragma unroll
  for (int b = 0; b < NH * BB; ++b) creg[b] = 0.0;
  const DPass<T>& p = m.op;
  __syncthreads();
  for (int i = blockIdx.x; i < natoms; i += gridDim.x) {
    const int ti = types[i];
    T Sw[9];
    double fwi[3];
    load_atom_weights(gw, i, Sw, fwi);
    const int e0 = nbr_off[i], e1 = nbr_off[i + 1];
    grad_sweep_moments<Cf, false, 0>(m, tab, rec, gw.frc_w, Sw, fwi, h, hd, ti, e0, e1, s.stage, s.At, s.Atd);
    if constexpr (NH > 1)
      grad_sweep_moments<Cf, false, 1>(m, tab, rec, gw.frc_w, Sw, fwi, h, hd, ti, e0, e1, s.stage, s.At, s.Atd);
    __syncthreads();
    scale_moments<Cf>(tab, ti, s);
    __syncthreads();
    wave_contract_tan<T, KK>(m, p, false, s.A, s.Ad, s.v, s.vd);
    __syncthreads();
    wave_graph_forward_tan(p, s.v, s.vd, s.s1, s.s2);
    // readout and its gradients; seeds of the reverse
    const double Eb = gw.ene_w[i];
    double gsp = 0;  // d/d species[ti] through a shared readout
    for (int q = lane; q < m.Q; q += WAVE) {
      const int node = m.scalar_map[q];
      const double g = Eb * (double)s.v[node] + (double)s.vd[node];
      if (m.species_readout) {
        s.part[pidx(m.o_read + ti * m.Q + q, oW, nW)] += g;
      } else {
        s.part[pidx(m.o_read + q, oW, nW)] += (double)m.theta[m.o_species + ti] * g;
        gsp += (double)m.theta[m.o_read + q] * g;
      }
    }
    gsp = wave_sum(gsp);
    for (int n = lane; n < m.M; n += WAVE) { s.P[n] = T(0); s.Q[n] = T(0); }
    __syncthreads();
    if (lane == 0) {
      s.part[pidx(m.o_species + ti, oW, nW)] += Eb + gsp;
      s.part[pidx(m.o_shift + ti, oW, nW)] += Eb;
    }
    auto seed = [&](int q) {
      const T rho = m.readout(ti, q);
      s.P[m.scalar_map[q]] += T(Eb) * rho;
      s.Q[m.scalar_map[q]] += rho;
    };
    if (m.scalar_unique) {  // distinct nodes: one lane per scalar
      for (int q = lane; q < m.Q; q += WAVE) seed(q);
    } else if (lane == 0) {  // repeated nodes: in scalar order
      for (int q = 0; q < m.Q; ++q) seed(q);
    }
    __syncthreads();
    wave_graph_reverse_tan(p, s.v, s.vd, s.P, s.Q, s.s1, s.s2);
    wave_adjoint_to_basic<T, KK>(m, p, false, s.P, s.Pc);
    wave_adjoint_to_basic<T, KK>(m, p, false, s.Q, s.Qc);
    __syncthreads();
    grad_moment_space<Cf, false>(m, p, s, ti, oW, nW);
    __syncthreads();
    grad_sweep_params<Cf, false, 0>(m, tab, rec, gw.frc_w, Sw, fwi, h, hd, msg, in_pos, ti, e0, e1, s, creg, oW,
                                    nW);
    if constexpr (NH > 1)
      grad_sweep_params<Cf, false, 1>(m, tab, rec, gw.frc_w, Sw, fwi, h, hd, msg, in_pos, ti, e0, e1, s, creg + BB,
                                      oW, nW);
  }
  __syncthreads();
  flush_partial<Cf>(m, false, creg, s.part, L.pn, oW, nW, partial + (size_t)blockIdx.x * L.pn);
}

template <class Cf, class T = typename Cf::T>
__global__ __launch_bounds__(WAVE) void k_grad_gate_reverse(DModel<T> m, GradLds<Cf> L, GradWeights gw, int natoms,
                                                            const int32_t* __restrict__ types,
                                                            const int* __restrict__ nbr_off,
                                                            const Rec<T>* __restrict__ rec,
                                                            const T* __restrict__ vg, const T* __restrict__ vgd,
                                                            const T* __restrict__ gmom, const T* __restrict__ gmomd,
                                                            const T* __restrict__ msg, T* __restrict__ hbar,
                                                            double* __restrict__ partial, int oW, int nW) {
  constexpr int KK = Cf::KK, BB = Cf::BB, F = Cf::F, KH = Cf::KH;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const GradSmem<T> s = carve(L, smem_raw);
  const int lane = threadIdx.x;
  SpeciesTables<T> tab;
  load_species_tables<Cf>(m, true, s.tabs, tab);
  zero_part(s.part, L.pn);
  constexpr int NH = Halves<Cf>::NH;
  double creg[NH * BB];
#pragma unroll
  for (int b = 0; b < NH * BB; ++b) creg[b] = 0.0;
  const DPass<T>& p = m.gp;
  __syncthreads();
  for (int i = blockIdx.x; i < natoms; i += gridDim.x) {
    const int ti = types[i];
    T Sw[9];
    double fwi[3];
    load_atom_weights(gw, i, Sw, fwi);
    const int e0 = nbr_off[i], e1 = nbr_off[i + 1];
    // gather the gate messages of the incoming edges: column c of the message
    // rows (h_bar part, then h_dot_bar part) summed over the atom's incoming
    // slots, coalesced along the rows; with 2F <= 64 the lanes form WAVE / 2F
    // groups that take every group-th slot, and the group sums are added in
    // group order
    constexpr int NC = 2 * F;
    auto column_sum = [&](int col, int k0, int step) {
      T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
      int k = k0;
      for (; k + 3 * step < e1; k += 4 * step) {
        a0 += msg[(size_t)k * NC + col];
        a1 += msg[(size_t)(k + step) * NC + col];
        a2 += msg[(size_t)(k + 2 * step) * NC + col];
        a3 += msg[(size_t)(k + 3 * step) * NC + col];
      }
      for (; k < e1; k += step) a0 += msg[(size_t)k * NC + col];
      return (a0 + a1) + (a2 + a3);
    };
    if constexpr (NC <= WAVE) {
      constexpr int NG = WAVE / NC;
      const int col = lane % NC, grp = lane / NC;
      if (grp < NG) s.stage[grp * NC + col] = column_sum(col, e0 + grp, NG);
      __syncthreads();
      if (lane < NC) {
        T t = T(0);
        for (int q = 0; q < NG; ++q) t += s.stage[q * NC + lane];
        s.hbs[lane] = t;
      }
    } else {
      for (int col = lane; col < NC; col += WAVE) s.hbs[col] = column_sum(col, e0, 1);
    }
    __syncthreads();  // the group sums in the staging region are consumed
    // the atom's gate node values, tangents and moments: rows four at a time,
    // so that their loads are in flight together (a load-then-store loop
    // would wait for each row's loads before issuing the next)
    const T* __restrict__ vi = vg + (size_t)i * p.node_cnt;
    const T* __restrict__ vdi = vgd + (size_t)i * p.node_cnt;
    wave_rows(0, m.M, [&](int n) {
      const bool in = n >= p.node_lo && n < p.node_lo + p.node_cnt;
      s.v[n] = in ? vi[n - p.node_lo] : T(0);
      s.vd[n] = in ? vdi[n - p.node_lo] : T(0);
      s.P[n] = T(0);
      s.Q[n] = T(0);
    });
    const T* __restrict__ ai = gmom + (size_t)i * KH;
    const T* __restrict__ adi = gmomd + (size_t)i * KH;
    wave_rows(0, KH, [&](int c) {
      s.At[c] = ai[c];
      s.Atd[c] = adi[c];
    });
    __syncthreads();
    for (int mu = lane; mu < 2 * F; mu += WAVE) hbar[(size_t)i * 2 * F + mu] = s.hbs[mu];
    for (int q = lane; q < m.Qg; q += WAVE) {
      T a = T(0), b = T(0);
#pragma unroll
      for (int mu = 0; mu < F; ++mu) {
        const T w = m.W(mu, q);
        a += w * s.hbs[mu];
        b += w * s.hbs[F + mu];
      }
      s.P[m.gate_nodes[q]] = a;
      s.Q[m.gate_nodes[q]] = b;
    }
    scale_moments<Cf>(tab, ti, s);
    __syncthreads();
    wave_graph_reverse_tan(p, s.v, s.vd, s.P, s.Q, s.s1, s.s2);
    wave_adjoint_to_basic<T, KK>(m, p, true, s.P, s.Pc);
    wave_adjoint_to_basic<T, KK>(m, p, true, s.Q, s.Qc);
    __syncthreads();
    grad_moment_space<Cf, true>(m, p, s, ti, oW, nW);
    __syncthreads();
    grad_sweep_params<Cf, true, 0>(m, tab, rec, gw.frc_w, Sw, fwi, nullptr, nullptr, nullptr, nullptr, ti, e0, e1,
                                   s, creg, oW, nW);
    if constexpr (NH > 1)
      grad_sweep_params<Cf, true, 1>(m, tab, rec, gw.frc_w, Sw, fwi, nullptr, nullptr, nullptr, nullptr, ti, e0, e1,
                                     s, creg + BB, oW, nW);
  }
  __syncthreads();
  flush_partial<Cf>(m, true, creg, s.part, L.pn, oW, nW, partial + (size_t)blockIdx.x * L.pn);
}

// W_bar[mu][q] over a chunk of atoms: thread = one (mu, q) entry, FP64 sums.
template <class T>
__global__ __launch_bounds__(256) void k_grad_W(int natoms, int F, int Qg, int chunk, const T* __restrict__ hbar,
                                                const T* __restrict__ vg, const T* __restrict__ vgd, int node_cnt,
                                                int node_lo, const int* __restrict__ gate_nodes,
                                                double* __restrict__ wpart) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= F * Qg) return;
  const int mu = idx / Qg, q = idx % Qg;
  const int col = gate_nodes[q] - node_lo;
  const int a0 = blockIdx.y * chunk, a1 = min(natoms, a0 + chunk);
  double acc = 0;
  for (int a = a0; a < a1; ++a) {
    const T* hb = hbar + (size_t)a * 2 * F;
    acc += (double)hb[mu] * (double)vg[(size_t)a * node_cnt + col] +
           (double)hb[F + mu] * (double)vgd[(size_t)a * node_cnt + col];
  }
  wpart[(size_t)blockIdx.y * F * Qg + idx] = acc;
}

// grad[p] (=|+=) scale * sum of the partial rows, in a fixed order and in two
// stages.  k_grad_fold: block (bx, s) covers 64 coefficients and row fold s
// (of RED_S contiguous row ranges); its 16 thread rows sum contiguous
// sub-ranges, added in thread-row order into fold[s][p].  k_grad_reduce adds
// the RED_S folds in fold order.
constexpr int RED_X = 64, RED_Y = 16, RED_S = 16;

static __global__ __launch_bounds__(RED_X * RED_Y) void k_grad_fold(int ncoeff, int rows, int pn,
                                                                    const double* __restrict__ part, int oW, int nW,
                                                                    int wrows, const double* __restrict__ wpart,
                                                                    double* __restrict__ fold) {
  __shared__ double red[RED_Y][RED_X];
  const int x = threadIdx.x, y = threadIdx.y, s = blockIdx.y;
  const int p = blockIdx.x * RED_X + x;
  double acc = 0;
  if (p < ncoeff) {
    const bool isw = oW >= 0 && p >= oW && p < oW + nW;
    const int nr = isw ? wrows : rows;
    const int fchunk = (nr + RED_S - 1) / RED_S;
    const int f0 = s * fchunk, f1 = min(nr, f0 + fchunk);
    const int chunk = (fchunk + RED_Y - 1) / RED_Y;
    const int r0 = f0 + y * chunk, r1 = min(f1, r0 + chunk);
    if (isw) {
      for (int r = r0; r < r1; ++r) acc += wpart[(size_t)r * nW + (p - oW)];
    } else {
      const int k = pidx(p, oW, nW);
      for (int r = r0; r < r1; ++r) acc += part[(size_t)r * pn + k];
    }
  }
  red[y][x] = acc;
  __syncthreads();
  if (y == 0 && p < ncoeff) {
    double t = 0;
    for (int q = 0; q < RED_Y; ++q) t += red[q][x];
    fold[(size_t)s * ncoeff + p] = t;
  }
}

static __global__ __launch_bounds__(256) void k_grad_reduce(int ncoeff, const double* __restrict__ fold, double scale,
                                                            int add, double* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncoeff) return;
  double t = 0;
  for (int s = 0; s < RED_S; ++s) t += fold[(size_t)s * ncoeff + p];
  grad[p] = (add ? grad[p] : 0.0) + scale * t;
}

// ------------------------------------------------------------------ driver
template <class Cf>
void run_grad_shape(const DevModelSet& dm, Workspace& ws, const Batch& b, const GradWeights& gw_in, double* grad,
                    double scale, bool add, bool reuse_neighbors, hipStream_t st) {
  using T = typename Cf::T;
  const DModel<T>& m = dm.get<T>();
  const int N = b.natoms;
  if (!reuse_neighbors) build_neighbors<T>(ws, b, dm.host->rc, st);
  GradWeights gw = gw_in;
  gw.atom_frame = ws.atom_frame.get<int>();  // written by the neighbour build of this batch
  const int ncoeff = dm.host->ncoeff;
  const int oW = m.gated ? m.o_W : -1, nW = m.gated ? Cf::F * m.Qg : 0;
  const int pn = ncoeff - nW;
  const size_t n1 = std::max(1, N), ne = std::max<long>(1, ws.nedges);
  // grid-stride groups: as many workgroups as can be resident at once (every
  // group holds its own FP64 partial row, so more would only add rows); the
  // outer pass owns partial rows [0, Go), the gate reverse rows [Go, Go + Gg)
  const GradLds<Cf> lo_ = grad_lds<Cf>(m, false, true, pn), l3_ = grad_lds<Cf>(m, true, true, pn);
  const int Go = std::max(1, std::min(N, resident_groups(k_grad_outer<Cf>, lo_.bytes(), true)));
  const int Gg = m.gated ? std::max(1, std::min(N, resident_groups(k_grad_gate_reverse<Cf>, l3_.bytes(), false))) : 0;
  ws.gpart.ensure(sizeof(double) * (size_t)(Go + Gg) * pn);
  const int* off = ws.nbr_off.get<int>();
  const Rec<T>* rec = ws.rec.get<Rec<T>>();
  int wrows = 0;
  if (N > 0) {
    const GradLds<Cf> lo = grad_lds<Cf>(m, false, true, pn);
    if (m.gated) {
      const int nc = std::max(1, m.gp.node_cnt);
      ws.h.ensure(sizeof(T) * n1 * Cf::F);
      ws.hdot.ensure(sizeof(T) * n1 * Cf::F);
      ws.vg.ensure(sizeof(T) * n1 * nc);
      ws.vgd.ensure(sizeof(T) * n1 * nc);
      ws.gmom.ensure(sizeof(T) * n1 * Cf::KH);
      ws.gmomd.ensure(sizeof(T) * n1 * Cf::KH);
      ws.msg.ensure(sizeof(T) * ne * 2 * Cf::F);
      ws.hbar.ensure(sizeof(T) * n1 * 2 * Cf::F);
      const GradLds<Cf> l1 = grad_lds<Cf>(m, true, false, 0);
      allow_lds(k_grad_gate_forward<Cf>, l1.bytes());
      SUS2_LAUNCH(ws, st, kname<Cf>("grad_gate_forward"), (k_grad_gate_forward<Cf>), dim3(N), dim3(WAVE), l1.bytes(),
                  m, l1, gw, b.types, off, rec, ws.h.get<T>(), ws.hdot.get<T>(), ws.vg.get<T>(), ws.vgd.get<T>(),
                  ws.gmom.get<T>(), ws.gmomd.get<T>());
    }
    allow_lds(k_grad_outer<Cf>, lo.bytes());
    SUS2_LAUNCH(ws, st, kname<Cf>("grad_outer"), (k_grad_outer<Cf>), dim3(Go), dim3(WAVE), lo.bytes(), m, lo, gw, N,
                b.types, off, rec, ws.h.get<T>(), ws.hdot.get<T>(), ws.in_pos.get<int>(), ws.msg.get<T>(),
                ws.gpart.get<double>(), oW, nW);
    if (m.gated) {
      const GradLds<Cf> l3 = grad_lds<Cf>(m, true, true, pn);
      allow_lds(k_grad_gate_reverse<Cf>, l3.bytes());
      SUS2_LAUNCH(ws, st, kname<Cf>("grad_gate_reverse"), (k_grad_gate_reverse<Cf>), dim3(Gg), dim3(WAVE),
                  l3.bytes(), m, l3, gw, N, b.types, off, rec, ws.vg.get<T>(), ws.vgd.get<T>(), ws.gmom.get<T>(),
                  ws.gmomd.get<T>(), ws.msg.get<T>(), ws.hbar.get<T>(), ws.gpart.get<double>() + (size_t)Go * pn, oW,
                  nW);
      const int chunk = 64;
      wrows = (N + chunk - 1) / chunk;
      ws.wpart.ensure(sizeof(double) * (size_t)wrows * nW);
      SUS2_LAUNCH(ws, st, kname<Cf>("grad_W"), k_grad_W<T>, dim3((nW + 255) / 256, wrows), dim3(256), 0, N, Cf::F,
                  m.Qg, chunk, ws.hbar.get<T>(), ws.vg.get<T>(), ws.vgd.get<T>(), m.gp.node_cnt, m.gp.node_lo,
                  m.gate_nodes, ws.wpart.get<double>());
    }
  }
  ws.gfold.ensure(sizeof(double) * (size_t)RED_S * ncoeff);
  SUS2_LAUNCH(ws, st, "grad_fold", k_grad_fold, dim3((ncoeff + RED_X - 1) / RED_X, RED_S), dim3(RED_X, RED_Y), 0,
              ncoeff, N > 0 ? Go + Gg : 0, pn, ws.gpart.get<double>(), oW, nW, wrows, ws.wpart.get<double>(),
              ws.gfold.get<double>());
  SUS2_LAUNCH(ws, st, "grad_reduce", k_grad_reduce, dim3((ncoeff + 255) / 256), dim3(256), 0, ncoeff,
              ws.gfold.get<double>(), scale, add ? 1 : 0, grad);
}

}  // namespace sus2
