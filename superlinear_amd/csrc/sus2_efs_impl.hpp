// Shape-templated EFS kernels (moments, forces) and their per-shape driver;
// instantiated per (L, K) in sus2_efs_l*k*.hip so the shapes compile in
// parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/sus2amd.h"
#include "sus2_common.hpp"

namespace sus2 {

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
    if (m.gated)
      SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("moments_gate")), (k_moments<T, LL, KK, true>), dim3(N),
                  dim3(WAVE), lm, m, b.types, off, rec, h, h, ws.gbasic.get<T>(), ws.site.get<T>(), ws.abar.get<T>());
    SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("moments_outer")), (k_moments<T, LL, KK, false>), dim3(N),
                dim3(WAVE), lm, m, b.types, off, rec, h, h, ws.gbasic.get<T>(), ws.site.get<T>(), ws.abar.get<T>());
    SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("force_outer")), (k_force<T, LL, KK, false>), dim3(N), dim3(WAVE),
                lf, m, b.types, off, rec, h, ws.abar.get<T>(), ws.gbasic.get<T>(), ws.gbar.get<T>(), ws.fw.get<T>());
    if (m.gated)
      SUS2_LAUNCH(ws, st, (kname<T, LL, KK>("force_gate")), (k_force<T, LL, KK, true>), dim3(N), dim3(WAVE),
                  lf, m, b.types, off, rec, h, ws.gbar.get<T>(), ws.gbasic.get<T>(), ws.gbar.get<T>(), ws.fw.get<T>());
  }
  finish_efs<T>(ws, b, out, st);
}

}  // namespace sus2
