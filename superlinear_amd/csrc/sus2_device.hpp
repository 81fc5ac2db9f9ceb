// Device-side model view and per-edge math shared by the SUS2-SH kernels:
// real spherical harmonics (values, or values + Cartesian gradient streamed in
// (m, l) order), the Chebyshev-tanh radial basis with the (r - rc)^2 envelope,
// and the one-wavefront-per-workgroup helpers (wave reductions, layered
// product-graph walks over LDS).
#pragma once
#include <hip/hip_runtime.h>

namespace sus2 {

constexpr int WAVE = 64;
constexpr int BMAX = 16;  // largest radial basis size the kernels are built for

// ------------------------------------------------------------------ model view
// One product-graph row as a 16-byte record, so that a row is one vector
// load: forward (x, y) = (a, b), the row adds c v[a] v[b] to its output;
// reverse (x, y) = (out, partner), the row adds c a[out] v[partner] to its
// node.
template <class T>
struct __align__(16) GRow {
  int x, y;
  T c;
};

template <class T>
struct DPass {  // layered product-graph schedule of one pass (see PassPlan)
  int nfl = 0, nrl = 0, rank = 0, begin = 0, node_lo = 0, node_cnt = 0;
  int nfout = 0, nrnode = 0;  // total forward outs / reverse nodes
  int max_fw_rows = 0, max_rv_rows = 0;
  const int *fl_off = nullptr, *f_out = nullptr, *f_roff = nullptr;
  const GRow<T>* f_row = nullptr;
  const int *rl_off = nullptr, *r_node = nullptr, *r_roff = nullptr;
  const GRow<T>* r_row = nullptr;
  // per layer {first out / node, end, first row, end row}: one 16-byte load
  const int4 *fl_bnd = nullptr, *rl_bnd = nullptr;
};

template <class T>
struct DModel {
  int C, L, H, K, F, B, M, Q, Kw, smode, R, O, Qg;
  T rc, rca, scaling;
  int gated, species_readout;
  int scalar_unique;  // no node appears twice in scalar_map
  const T* theta;
  int o_shift, o_scal, o_rad, o_grad, o_gscal, o_alpha, o_gtype, o_W, o_gw, o_ow, o_species, o_read;
  const int* basic_node;  // [H*K], canonical moment (h, k) -> node or -1
  const int* scalar_map;  // [Q]
  const int* gate_nodes;  // [Qg]
  DPass<T> gp, op;

  __device__ __forceinline__ int block(int k, int l) const {
    return smode == 2 ? k : smode == 1 ? l : smode == 0 ? 0 : k * (L + 1) + l;
  }
  __device__ __forceinline__ T tau(int t, int mu) const { return theta[o_rad + mu * (B + C) + B + t]; }
  __device__ __forceinline__ T alpha(int t, int mu) const { return theta[o_alpha + t * F + mu]; }
  __device__ __forceinline__ T gtype(int t, int mu) const { return theta[o_gtype + t * F + mu]; }
  __device__ __forceinline__ T W(int mu, int q) const { return theta[o_W + mu * Qg + q]; }
  __device__ __forceinline__ T readout(int t, int q) const {
    return species_readout ? theta[o_read + t * Q + q] : theta[o_species + t] * theta[o_read + q];
  }
  __device__ __forceinline__ T rankw(bool gate, int l, int a, int k) const {
    return gate ? theta[o_gw + (l * R + a) * K + k] : theta[o_ow + (l * O + a) * K + k];
  }
};

// ------------------------------------------------------------------ constants
constexpr double csqrt_iter(double x, double g, int n) { return n == 0 ? g : csqrt_iter(x, 0.5 * (g + x / g), n - 1); }
constexpr double csqrt(double x) { return x <= 0 ? 0.0 : csqrt_iter(x, x > 1 ? x : 1.0, 80); }
constexpr double cfact_ratio(int l, int m) {  // (l-m)! / (l+m)!
  double r = 1;
  for (int a = l - m + 1; a <= l + m; ++a) r /= a;
  return r;
}
constexpr double kPi = 3.14159265358979323846;
// normalisation of the real harmonic (sqrt 2 folded in for m > 0)
constexpr double ynorm(int l, int m) {
  return csqrt((2 * l + 1) * cfact_ratio(l, m) / (4 * kPi)) * (m > 0 ? 1.4142135623730950488 : 1.0);
}
constexpr double dfact(int m) {  // (2m-1)!!
  double r = 1;
  for (int a = 1; a <= 2 * m - 1; a += 2) r *= a;
  return r;
}
__host__ __device__ constexpr int hidx(int l, int m) { return l * l + l + m; }

// Normalised associated-Legendre recurrences for l <= 6, evaluated at compile
// time.  With P_l = n_lm p_l (p_l the Legendre factor in u_z and rho = |u|^2,
// n_lm the harmonic's normalisation), the three-term recurrence
//   p_l = ((2l - 1) u_z p_{l-1} - (l + m - 1) rho p_{l-2}) / (l - m)
// becomes P_l = a_lm u_z P_{l-1} - b_lm rho P_{l-2}, so no normalisation is
// multiplied in per harmonic; P_m = n_mm (2m-1)!!.
struct YRec {
  double a[49], b[49], p0[7];
};
constexpr YRec make_yrec() {
  YRec t{};
  for (int m = 0; m <= 6; ++m) {
    t.p0[m] = ynorm(m, m) * dfact(m);
    for (int l = m + 1; l <= 6; ++l) {
      t.a[l * l + l + m] = (2 * l - 1) * ynorm(l, m) / ((l - m) * ynorm(l - 1, m));
      t.b[l * l + l + m] = l >= m + 2 ? (l + m - 1) * ynorm(l, m) / ((l - m) * ynorm(l - 2, m)) : 0.0;
    }
  }
  return t;
}
inline constexpr YRec kYR = make_yrec();

// Compile-time shape of a model: max l, channels K, radial basis size B and
// warp-sharing mode SM (0 pair, 1 L, 2 K, 3 LK).
template <class T_, int L_, int K_, int B_, int SM_>
struct Cfg {
  using T = T_;
  static constexpr int LL = L_, KK = K_, BB = B_, SM = SM_;
  static constexpr int L1 = LL + 1, H = L1 * L1, F = KK * L1, KH = KK * H, NM = (KH + 63) / 64;
  static constexpr int NBLK = SM == 0 ? 1 : SM == 1 ? L1 : SM == 2 ? KK : F;
  static constexpr int blk(int k, int l) { return SM == 2 ? k : SM == 1 ? l : SM == 0 ? 0 : k * L1 + l; }
};

// ------------------------------------------------------------------ harmonics
// One step of the normalised Legendre recurrence, P_l from P_{l-1}, P_{l-2}
// (rho = 1).  Every harmonic stream uses it, and the steps shared by the
// streams (this one and the azimuthal recurrence) are written as explicit
// fused multiply-adds, so the streams' values agree bit for bit whatever
// contractions the compiler would choose in each kernel.
template <class T>
__device__ __forceinline__ T legendre_step(int l, int m, T uz, T p1, T p2) {
  const T a = T(kYR.a[hidx(l, m)]), b = T(kYR.b[hidx(l, m)]);
  return l == m + 1 ? a * (uz * p1) : fma(a, uz * p1, -(b * p2));
}

// Real Y_lm(u) for a unit vector u, l <= LL, flat index l^2+l+m; Condon-
// Shortley phase cancelled, Y_1{-1,0,1} ~ (y, z, x).  Legendre part by the
// normalised recurrence in l; azimuthal part (ux + i uy)^m.  Values streamed
// to fn(l, m, Y) in (m, l) order, so a caller can consume each one at once
// instead of holding all H.
template <class T, int LL, class Fn>
__device__ __forceinline__ void sh_stream_values(T ux, T uy, T uz, Fn&& fn) {
  T cm = T(1), sm = T(0);
#pragma unroll
  for (int m = 0; m <= LL; ++m) {
    if (m > 0) {
      const T c2 = fma(ux, cm, -(uy * sm));
      sm = fma(ux, sm, uy * cm);
      cm = c2;
    }
    T p2 = T(0), p1 = T(0);
#pragma unroll
    for (int l = m; l <= LL; ++l) {
      const T p = l == m ? T(kYR.p0[m]) : legendre_step(l, m, uz, p1, p2);
      if (m == 0) {
        fn(l, 0, p);
      } else {
        fn(l, m, p * cm);
        fn(l, -m, p * sm);
      }
      p2 = p1;
      p1 = p;
    }
  }
}

// The same Y_lm (the same operations, so bit for bit) in (l, m) order: fn(l, Yl)
// once per degree l with Yl[l + m] = Y_lm, m = -l..l.  The recurrence state
// of every m (2 (LL + 1) values) and the azimuthal factors are carried
// instead of one m's, so a caller can finish everything of degree l before
// the next degree is formed.
template <class T, int LL, class Fn>
__device__ __forceinline__ void sh_values_by_l(T ux, T uy, T uz, Fn&& fn) {
  T cm[LL + 1], sm[LL + 1], p1[LL + 1], p2[LL + 1];
  cm[0] = T(1);
  sm[0] = T(0);
#pragma unroll
  for (int m = 1; m <= LL; ++m) {
    cm[m] = fma(ux, cm[m - 1], -(uy * sm[m - 1]));
    sm[m] = fma(ux, sm[m - 1], uy * cm[m - 1]);
  }
#pragma unroll
  for (int l = 0; l <= LL; ++l) {
    T Yl[2 * LL + 1];
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      const T p = l == m ? T(kYR.p0[m]) : legendre_step(l, m, uz, p1[m], l == m + 1 ? T(0) : p2[m]);
      p2[m] = l == m ? T(0) : p1[m];
      p1[m] = p;
      if (m == 0) {
        Yl[l] = p;
      } else {
        Yl[l + m] = p * cm[m];
        Yl[l - m] = p * sm[m];
      }
    }
    fn(l, (const T*)Yl);
  }
}

template <class T, int LL>
__device__ __forceinline__ void sh_values(T ux, T uy, T uz, T* Y) {
  sh_stream_values<T, LL>(ux, uy, uz, [&](int l, int m, T v) { Y[hidx(l, m)] = v; });
}

// Y_lm(u) and the gradient G = grad_u S_lm of the solid harmonic
// S_lm(x) = |x|^l Y_lm(x/|x|) at x = u (|u| = 1), streamed to
// fn(l, m, Y, Gx, Gy, Gz) in (m, l) order.  The Legendre factor is carried as
// a polynomial in (u_z, rho = |u|^2) with both partial derivatives, so S and
// G are exact polynomials.  For an edge d = r u the Cartesian gradient of the
// harmonic is grad_d Y_lm = (G - l Y u) / r: a caller summing
// sum_h c_h grad_d Y_h forms (sum c G - u sum c l Y) / r once per edge
// instead of per harmonic.
template <class T, int LL, class Fn>
__device__ __forceinline__ void sh_stream_solid(T ux, T uy, T uz, Fn&& fn) {
  T cm = T(1), sm = T(0);
  T cx = T(0), cy = T(0), sx = T(0), sy = T(0);  // d(cm)/dux, d(cm)/duy, d(sm)/dux, d(sm)/duy
  const T ux2 = T(2) * ux, uy2 = T(2) * uy, uz2 = T(2) * uz;
#pragma unroll
  for (int m = 0; m <= LL; ++m) {
    if (m > 0) {
      cx = T(m) * cm; cy = -T(m) * sm;
      sx = T(m) * sm; sy = T(m) * cm;
      const T c2 = fma(ux, cm, -(uy * sm));
      sm = fma(ux, sm, uy * cm);
      cm = c2;
    }
    // P = n p, Z = n dp/du_z, Q = n dp/drho at rho = 1
    T p2 = T(0), p1 = T(0), z2 = T(0), z1 = T(0), q2 = T(0), q1 = T(0);
#pragma unroll
    for (int l = m; l <= LL; ++l) {
      T p, z, q;
      if (l == m) {
        p = T(kYR.p0[m]); z = T(0); q = T(0);
      } else {
        const T a = T(kYR.a[hidx(l, m)]), b = T(kYR.b[hidx(l, m)]);
        p = legendre_step(l, m, uz, p1, p2);
        if (l == m + 1) {
          z = a * p1;
          q = T(0);
        } else {
          z = a * (p1 + uz * z1) - b * z2;
          q = a * (uz * q1) - b * (p2 + q2);
        }
      }
      if (m == 0) {
        // cm = 1, sm = 0: no azimuthal factor
        if (l < 2) fn(l, 0, p, T(0), T(0), z);
        else fn(l, 0, p, ux2 * q, uy2 * q, z + uz2 * q);
      } else {
        const T gx = ux2 * q, gy = uy2 * q, gz = z + uz2 * q;
        fn(l, m, p * cm, gx * cm + p * cx, gy * cm + p * cy, gz * cm);
        fn(l, -m, p * sm, gx * sm + p * sx, gy * sm + p * sy, gz * sm);
      }
      p2 = p1; p1 = p; z2 = z1; z1 = z; q2 = q1; q1 = q;
    }
  }
}

// A row of F values (a neighbour's gate signal): 16-byte vector loads when
// the rows are 16-byte aligned (F * sizeof(T) a multiple of 16 on a 256-byte
// aligned buffer), else scalar loads.
template <class T, int F, bool ALIGNED = true>
__device__ __forceinline__ void load_row(const T* __restrict__ p, T* out) {
  if constexpr (ALIGNED && sizeof(T) == 4 && F % 4 == 0) {
#pragma unroll
    for (int k = 0; k < F / 4; ++k) {
      const float4 v = reinterpret_cast<const float4*>(p)[k];
      out[4 * k] = v.x; out[4 * k + 1] = v.y; out[4 * k + 2] = v.z; out[4 * k + 3] = v.w;
    }
  } else if constexpr (ALIGNED && sizeof(T) == 8 && F % 2 == 0) {
#pragma unroll
    for (int k = 0; k < F / 2; ++k) {
      const double2 v = reinterpret_cast<const double2*>(p)[k];
      out[2 * k] = v.x; out[2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < F; ++k) out[k] = p[k];
  }
}

__device__ __forceinline__ float dtanh(float x) { return tanhf(x); }
__device__ __forceinline__ double dtanh(double x) { return tanh(x); }
// Length r of an edge vector with squared length d2 (> 0) and its reciprocal.
// FP32: one hardware reciprocal square root instead of a square root and a
// division, corrected once with the residual e = 1 - d2 y^2 formed by a fused
// multiply-add, y <- y + y e / 2 (close to the correctly rounded reciprocal);
// r is a product the compiler may not fuse into the differences that use it
// (r - rc, r - shift), so every kernel sees the same r and passes that
// recompute an edge agree bit for bit.  FP64: correctly rounded square root
// and division.
__device__ __forceinline__ void edge_length(float d2, float& r, float& inv) {
  const float y = __builtin_amdgcn_rsqf(d2);
  const float e = fmaf(-(d2 * y), y, 1.0f);
  inv = fmaf(0.5f * y, e, y);
  r = __fmul_rn(d2, inv);
}
__device__ __forceinline__ void edge_length(double d2, double& r, double& inv) {
  r = sqrt(d2);
  inv = 1.0 / r;
}

// Per-workgroup copies (LDS) of the small per-species coefficient tables of
// one pass, so per-edge lookups with the neighbour's species are LDS reads.
// The pass's radial coefficients sit there too, one row of BB values per
// radial function padded to whole 16-byte vectors: every lane of a wave reads
// the same row, so a row is a few broadcast LDS reads straight into the FMA
// operands.  (Read from the coefficient vector instead, these uniform values
// are hoisted into scalar registers -- F x BB of them, far more than a wave
// holds -- and every use then pays a lane read of the spilled copy.)
template <class T>
struct SpeciesTables {
  T* slope;  // [NBLK][C*C]
  T* shift;  // [NBLK][C*C]
  T* x0;     // [C][F] centre factor: tau (outer) or gate type g (gate)
  T* alpha;  // [C][F] (outer pass of an Effect model)
  T* coef;   // [F][coef_row] radial coefficients, 16-byte aligned rows
};

// Radial coefficient row length: BB rounded up to whole 16-byte vectors.
template <class Cf>
__host__ __device__ constexpr int coef_row() {
  constexpr int v = 16 / (int)sizeof(typename Cf::T);
  return (Cf::BB + v - 1) / v * v;
}

// LDS elements of the tables; CL: with the radial coefficient rows (the last
// term is room to align them to 16 bytes).
template <class Cf, bool CL = true>
__host__ __device__ constexpr int species_tables_size(int C) {
  return 2 * Cf::NBLK * C * C + 2 * C * Cf::F + (CL ? Cf::F * coef_row<Cf>() + 16 / (int)sizeof(typename Cf::T) : 0);
}

// First 16-byte aligned element at or after p (p aligned to sizeof(T); an
// LDS pointer stays an LDS pointer).
template <class T>
__device__ __forceinline__ T* align16(T* p) {
  const int mis = (int)(reinterpret_cast<uintptr_t>(p) & 15);
  return mis ? p + (16 - mis) / (int)sizeof(T) : p;
}

// Fill the tables from the coefficient vector (all lanes of the workgroup);
// CL: with the radial coefficient rows.
template <class Cf, bool CL = true>
__device__ void load_species_tables(const DModel<typename Cf::T>& m, bool gate, typename Cf::T* base,
                                    SpeciesTables<typename Cf::T>& t) {
  using T = typename Cf::T;
  constexpr int BB = Cf::BB, CR = coef_row<Cf>();
  const int C = m.C, CC = C * C, F = Cf::F;
  t.slope = base;
  t.shift = base + Cf::NBLK * CC;
  t.x0 = t.shift + Cf::NBLK * CC;
  t.alpha = t.x0 + C * F;
  t.coef = CL ? align16(t.alpha + C * F) : nullptr;
  const int so = (gate && m.o_gscal >= 0) ? m.o_gscal : m.o_scal;
  // radial coefficients: the gate pass's own block when the model has one,
  // else the outer block (rows of BB basis weights, then C type factors).
  // Their count is known at compile time, so every thread's loads are issued
  // first and land while the two small tables below make their round trips
  // (a load-then-store loop would pay one round trip per blockDim values).
  constexpr int NCO = CL ? Cf::F * CR : 0, ITC = (NCO + WAVE - 1) / WAVE;  // blockDim >= WAVE
  T cv[ITC > 0 ? ITC : 1];
  if constexpr (CL) {
    const bool own_rad = gate && m.o_grad >= 0;
    const T* __restrict__ coef = m.theta + (own_rad ? m.o_grad : m.o_rad);
    const int stride = own_rad ? BB : BB + C;
#pragma unroll
    for (int it = 0; it < ITC; ++it) {
      const int k = threadIdx.x + it * (int)blockDim.x, mu = k / CR, q = k % CR;
      cv[it] = (k < NCO && q < BB) ? coef[mu * stride + q] : T(0);
    }
  }
  // the two small tables in one loop: each step's four loads before its stores
  const int n1 = Cf::NBLK * CC, n2 = C * F;
  for (int k = threadIdx.x; k < max(n1, n2); k += blockDim.x) {
    const int b = k / CC, pq = k % CC, s = k / F, mu = k % F;
    T sl = T(0), sh = T(0), x0 = T(0), al = T(0);
    if (k < n1) {
      sl = m.theta[so + 2 * b * CC + pq];
      sh = m.theta[so + (2 * b + 1) * CC + pq];
    }
    if (k < n2) {
      x0 = gate ? m.gtype(s, mu) : m.tau(s, mu);
      if (!gate && m.gated) al = m.alpha(s, mu);
    }
    if (k < n1) {
      t.slope[k] = sl;
      t.shift[k] = sh;
    }
    if (k < n2) {
      t.x0[k] = x0;
      t.alpha[k] = al;
    }
  }
  if constexpr (CL) {
#pragma unroll
    for (int it = 0; it < ITC; ++it) {
      const int k = threadIdx.x + it * (int)blockDim.x;
      if (k < NCO) t.coef[k] = cv[it];
    }
  }
}

// Radial functions R_mu(r) (and dR/dr when DER) of one directed edge with
// ordered species pair `pair` = t_i*C + t_j, for the gate or outer
// coefficients.  phi_b = scaling T_b(tanh(s (r - t)/2)) (r - rc_l)^2, rc_0 =
// rc, rc_{l>0} = rca; channels with l > 0 vanish for r >= rca.  Warp blocks
// and basis size are compile-time, so each block's tanh and Chebyshev
// recurrence is evaluated once and only for the channels that share it.
// Each channel goes to fn(mu, R_mu, dR_mu/dr) as soon as it is formed; only
// the warp blocks [B0, B1) are evaluated.
// Radial functions of warp blocks [B0, B1) in evaluation order (block by
// block, channels (k, l) of a block in order): the i-th one's index mu and
// block, and how many there are.  Constant-folded once the callers' loops are
// unrolled.
template <class Cf, int B0, int B1>
__host__ __device__ constexpr int radial_count() {
  int n = 0;
  for (int b = B0; b < B1; ++b)
    for (int k = 0; k < Cf::KK; ++k)
      for (int l = 0; l < Cf::L1; ++l) n += Cf::blk(k, l) == b;
  return n;
}
template <class Cf, int B0, int B1>
__host__ __device__ constexpr int radial_at(int i, bool want_block) {
  for (int b = B0; b < B1; ++b)
    for (int k = 0; k < Cf::KK; ++k)
      for (int l = 0; l < Cf::L1; ++l)
        if (Cf::blk(k, l) == b && i-- == 0) return want_block ? b : k * Cf::L1 + l;
  return -1;
}

template <class Cf, bool DER, int B0 = 0, int B1 = Cf::NBLK, bool CL = true, class Fn>
__device__ __forceinline__ void radial_each(const DModel<typename Cf::T>& m, bool gate,
                                            const SpeciesTables<typename Cf::T>& t, int pair, typename Cf::T r,
                                            Fn&& fn) {
  using T = typename Cf::T;
  constexpr int L1 = Cf::L1, BB = Cf::BB, CR = coef_row<Cf>(), N = radial_count<Cf, B0, B1>();
  const int CC = m.C * m.C;
  const T d0 = r - m.rc, f0 = d0 * d0, g0 = T(2) * d0;
  const bool act = r < m.rca;
  const T d1 = r - m.rca, f1 = act ? d1 * d1 : T(0), g1 = act ? T(2) * d1 : T(0);
  // Coefficient rows come from LDS two at a time: row i + 1 is read while row
  // i is used.  Each read's address takes a zero from an empty asm whose
  // inputs are the edge's distance (row 0 and 1) or the sums of row i - 1: the
  // reads stay inside the caller's edge loop and wait for the row two back to
  // be formed, so the F x BB coefficients are never all held in registers
  // (read once out of the loop, or all at the start of an edge, they would
  // be).  The asm touches no memory, so loads and stores around it move
  // freely.
  // Without the LDS rows (CL false), rows are read from the coefficient
  // vector, and the compiler keeps them in scalar registers.
  const bool own_rad = gate && m.o_grad >= 0;
  const T* __restrict__ gcoef = m.theta + (own_rad ? m.o_grad : m.o_rad);
  const int gstride = own_rad ? BB : BB + m.C;
  T cbuf[2][CR];
  if constexpr (CL) {
    int z;
    asm("" : "=v"(z) : "0"(0), "v"(r));
    load_row<T, CR>(t.coef + (z + radial_at<Cf, B0, B1>(0, false) * CR), cbuf[0]);
  }
  T dep0 = r, dep1 = r;  // the last row's sums
  T Tb[BB], Db[BB];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int mu = radial_at<Cf, B0, B1>(i, false), b = radial_at<Cf, B0, B1>(i, true);
    const int l = mu % L1;
    if (CL && i + 1 < N) {
      int z;
      asm("" : "=v"(z) : "0"(0), "v"(dep0), "v"(dep1));
      load_row<T, CR>(t.coef + (z + radial_at<Cf, B0, B1>(i + 1, false) * CR), cbuf[(i + 1) & 1]);
    }
    if (i == 0 || radial_at<Cf, B0, B1>(i - 1, true) != b) {
      // the block's tanh and Chebyshev recurrence, once for its channels
      const T slope = t.slope[b * CC + pair], shift = t.shift[b * CC + pair];
      const T xi = dtanh(slope * (r - shift) * T(0.5));
      const T dxi = (T(1) - xi * xi) * slope * T(0.5);
      Tb[0] = m.scaling;
      Db[0] = T(0);
      if (BB > 1) {
        Tb[1] = m.scaling * xi;
        Db[1] = m.scaling * dxi;
      }
#pragma unroll
      for (int q = 2; q < BB; ++q) {
        Tb[q] = T(2) * xi * Tb[q - 1] - Tb[q - 2];
        if (DER) Db[q] = T(2) * (dxi * Tb[q - 1] + xi * Db[q - 1]) - Db[q - 2];
      }
    }
    if constexpr (!CL) {
#pragma unroll
      for (int q = 0; q < BB; ++q) cbuf[i & 1][q] = gcoef[mu * gstride + q];
    }
    const T* c = cbuf[i & 1];
    T s = T(0), sd = T(0);
#pragma unroll
    for (int q = 0; q < BB; ++q) {
      s += c[q] * Tb[q];
      if (DER) sd += c[q] * Db[q];
    }
    dep0 = s;
    dep1 = DER ? sd : s;
    const T f = l == 0 ? f0 : f1, g = l == 0 ? g0 : g1;
    fn(mu, f * s, DER ? g * s + f * sd : T(0));
  }
}

// All channels into R[F] (and dR[F] when DER).
template <class Cf, bool DER, bool CL = true>
__device__ __forceinline__ void radial(const DModel<typename Cf::T>& m, bool gate, const SpeciesTables<typename Cf::T>& t,
                                       int pair, typename Cf::T r, typename Cf::T* R, typename Cf::T* dR) {
  using T = typename Cf::T;
  radial_each<Cf, DER, 0, Cf::NBLK, CL>(m, gate, t, pair, r, [&](int mu, T v, T dv) {
    R[mu] = v;
    if (DER) dR[mu] = dv;
  });
}

// ------------------------------------------------------------------ wave utils
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// N wave sums at once, level by level: each butterfly level issues all N
// shuffles before any is added, so their latencies overlap.  Every lane ends
// with the N totals (the same association as wave_sum).
template <class T, int N>
__device__ __forceinline__ void wave_sum_n(T (&v)[N]) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    T t[N];
#pragma unroll
    for (int k = 0; k < N; ++k) t[k] = __shfl_xor(v[k], o, WAVE);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += t[k];
  }
}

__device__ __forceinline__ int lane_id() { return threadIdx.x; }

// rank node of channel a at (l, m) for a pass
__host__ __device__ __forceinline__ int rank_node(int begin, int rank, int H, int l, int m, int a) {
  return begin + rank * (H - (l + 1) * (l + 1)) + (rank - 1 - a) * (2 * l + 1) + m + l;
}

}  // namespace sus2
