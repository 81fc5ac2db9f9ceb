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
#include <atomic>
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
// Cell-list build of the full directed neighbour list, every periodic image
// with |d| < rc (the semantics of src/neighborhoods.cpp:276-348 and
// interfaces/lammps/tests/embedded_efs.cpp:44-66: strict r < rc, self images
// count, a pair may appear through several images in small cells).
//
// Per frame a grid of n_q bins along each lattice direction in fractional
// coordinates (bin width >= the cutoff's fractional reach, at most one bin per
// atom so the bins of frame f can use ids frame_off[f] ...); molecules use
// their bounding box.  An atom scans bin offsets [-S_q, S_q] with
// S_q = floor(reach_q n_q) + 1, which covers every image in small cells
// (each (bin, shift) pair is a distinct image).  Bin contents are sorted by
// atom index, so the edge order -- and every sum over it -- is deterministic.
struct Grid {
  double inv[9];  // fractional = (cartesian - org) * inv (row vector convention)
  double org[3];  // origin of the open (non-periodic) directions
  double A[9];    // basis rows: lattice rows, open rows completed and scaled to the atoms' extent
  double rbox;    // half the longest diagonal of one bin
  int n[3], S[3];
  int per;  // bit q set: periodic along lattice row q
};

// Basis rows and the bin half-diagonal, once the bin counts are fixed.
__device__ inline void grid_box(Grid& g, const double* a) {
  double rb2 = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) g.A[k] = a[k];
  const double in[3] = {1.0 / g.n[0], 1.0 / g.n[1], 1.0 / g.n[2]};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    double v[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double w = (s >> q) & 1 ? in[q] : -in[q];
#pragma unroll
      for (int k = 0; k < 3; ++k) v[k] += w * a[3 * q + k];
    }
    rb2 = fmax(rb2, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
  g.rbox = 0.5 * sqrt(rb2);
}

// Periodic mask of frame f: lattice rows that are non-zero (periodic frames).
// The nine entries are read together and tested without short-circuits, so
// that the loads are one round trip rather than a chain of them.
__device__ inline int grid_mask(const double* c, int periodic) {
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = c[k];
  int per = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    per |= (int)((v[3 * q] != 0.0) | (v[3 * q + 1] != 0.0) | (v[3 * q + 2] != 0.0)) << q;
  return periodic ? per : 0;
}

__device__ inline void cross3(const double* u, const double* v, double* w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}
__device__ inline void unit3(double* w) {
  const double s = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  w[0] /= s; w[1] /= s; w[2] /= s;
}
// One periodic row P: the open rows P+1, P+2 become unit normals (the first
// perpendicular to row P and its smallest Cartesian axis).
template <int P>
__device__ inline void complete_one(double* a) {
  const double* v = a + 3 * P;
  const int m = fabs(v[0]) <= fabs(v[1]) && fabs(v[0]) <= fabs(v[2]) ? 0 : fabs(v[1]) <= fabs(v[2]) ? 1 : 2;
  const double e[3] = {m == 0 ? 1.0 : 0.0, m == 1 ? 1.0 : 0.0, m == 2 ? 1.0 : 0.0};
  constexpr int Q1 = (P + 1) % 3, Q2 = (P + 2) % 3;
  cross3(v, e, a + 3 * Q1);
  unit3(a + 3 * Q1);
  cross3(v, a + 3 * Q1, a + 3 * Q2);
  unit3(a + 3 * Q2);
}
// Two periodic rows: the open row Q becomes their unit normal.
template <int Q>
__device__ inline void complete_two(double* a) {
  cross3(a + 3 * ((Q + 1) % 3), a + 3 * ((Q + 2) % 3), a + 3 * Q);
  unit3(a + 3 * Q);
}
// Basis rows: lattice rows for periodic directions, unit normals for open
// ones (scaled afterwards to the atoms' extent).  Compile-time indices only.
__device__ inline void grid_basis(const double* c, int per, double* a) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) a[3 * q + k] = (per & (1 << q)) ? c[3 * q + k] : 0.0;
  switch (per) {
    case 0: a[0] = a[4] = a[8] = 1.0; break;
    case 1: complete_one<0>(a); break;
    case 2: complete_one<1>(a); break;
    case 4: complete_one<2>(a); break;
    case 3: complete_two<2>(a); break;
    case 5: complete_two<1>(a); break;
    case 6: complete_two<0>(a); break;
    default: break;
  }
}
// Inverse basis, bins (at least r_c wide, at most one per atom) and image
// reach of a frame with basis rows a (open rows already scaled).
// Bins of about rc / BIN_DIV along each direction, so that a neighbour search
// covers (2 BIN_DIV + 1)^3 bins (far ones culled) of few atoms each.
constexpr double BIN_DIV = 2.0;
// Per-atom table of hits before each bin offset (reverse positions), for
// grids of at most NOF_TAB offsets (search extents up to 2 bins, the usual
// case with bins of r_c / 2; wider searches take a binary search instead).
constexpr int NOF_TAB = 125;

__device__ inline void grid_finish(Grid& g, const double* a, double rc, long na) {
  double b[3][3];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      b[q][p] = a[((p + 1) % 3) * 3 + (q + 1) % 3] * a[((p + 2) % 3) * 3 + (q + 2) % 3] -
                a[((p + 1) % 3) * 3 + (q + 2) % 3] * a[((p + 2) % 3) * 3 + (q + 1) % 3];
  const double det = a[0] * b[0][0] + a[1] * b[1][0] + a[2] * b[2][0], rdet = 1.0 / det;
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) g.inv[3 * p + q] = b[p][q] * rdet;
  double reach[3];
  int n[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    reach[q] = rc * sqrt(g.inv[q] * g.inv[q] + g.inv[3 + q] * g.inv[3 + q] + g.inv[6 + q] * g.inv[6 + q]);
    n[q] = (int)fmin(1024.0, fmax(1.0, floor(BIN_DIV / reach[q])));
  }
  // at most one bin per atom: shrink the largest count (lowest index on ties)
  while ((long)n[0] * n[1] * n[2] > (na > 0 ? na : 1)) {
    if (n[0] >= n[1] && n[0] >= n[2]) n[0] = max(1, n[0] - 1);
    else if (n[1] >= n[2]) n[1] = max(1, n[1] - 1);
    else n[2] = max(1, n[2] - 1);
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    g.n[q] = n[q];
    // bins within reach; along an open direction no offset beyond n - 1 exists
    const int s = (int)floor(reach[q] * n[q]) + 1;
    g.S[q] = (g.per & (1 << q)) ? s : min(s, n[q] - 1);
  }
  grid_box(g, a);
}

// Per frame: lattice rows that are non-zero are periodic (a zero row is a
// non-periodic direction, as in the reference's Configuration); open
// directions get unit normals scaled to the atoms' extent, so fractional
// coordinates along them lie in [0, 1).  periodic == 0 frames are molecules.
__global__ void k_grid_setup(int nframes, int natoms, const int64_t* __restrict__ frame_off,
                             const double* __restrict__ pos, const double* __restrict__ cell,
                             const int32_t* __restrict__ periodic, double rc, Grid* __restrict__ grid,
                             int* __restrict__ z0, int* __restrict__ z1, int* __restrict__ z2) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < natoms) {  // per-atom counters of the bin and incoming-list fills start at zero
    z0[f] = 0;
    z1[f] = 0;
    z2[f] = 0;
  }
  if (f >= nframes) return;
  Grid g;
  const long b0 = frame_off[f], na = frame_off[f + 1] - b0;
  g.per = grid_mask(cell + 9 * f, periodic[f]);
  double a[9];
  grid_basis(cell + 9 * f, g.per, a);
  // open rows: scale the unit normal to the extent of the atoms along it
#pragma unroll
  for (int k = 0; k < 3; ++k) g.org[k] = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (g.per & (1 << q)) continue;
    double lo = 0, hi = 0;
    for (long k = 0; k < na; ++k) {
      const double* x = pos + 3 * (b0 + k);
      const double t = x[0] * a[3 * q] + x[1] * a[3 * q + 1] + x[2] * a[3 * q + 2];
      if (k == 0 || t < lo) lo = t;
      if (k == 0 || t > hi) hi = t;
    }
    const double ext = fmax(hi - lo, 1e-6) * (1 + 1e-9) + 1e-9;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      g.org[k] += lo * a[3 * q + k];
      a[3 * q + k] *= ext;
    }
  }
  grid_finish(g, a, rc, na);
  grid[f] = g;
}

// One workgroup per frame (frames of at most FRAME_CELLS_MAX atoms): the grid
// of k_grid_setup (open-direction extents by a block reduction), the bin of
// every atom, bin counts and starts (LDS), the bin lists in ascending atom
// order, and the zeroed incoming-list cursors -- the whole cell list in one
// launch.  Bin ids of frame f are frame_off[f] + local bin, as in the global
// path; ids past the frame's bins start at the next frame's first atom.
constexpr int FRAME_CELLS_MAX = 8192;
constexpr int FRAME_CELLS_THREADS = 1024;
constexpr int FC_PER = 4;  // atoms per thread whose loads are issued together
// LDS bin word of an atom: local bin id (13 bits, FRAME_CELLS_MAX bins at
// most), then its wrap + 32 in 6 bits per direction; the sign bit marks a
// wrap outside [-32, 31] (recomputed where needed).
constexpr int BIN_ID_BITS = 13, BIN_ID_MASK = (1 << BIN_ID_BITS) - 1;
static_assert(FRAME_CELLS_MAX <= (1 << BIN_ID_BITS), "bin ids must fit the bin word");
__device__ __forceinline__ int bin_word(int id, const int* w) {
  bool fit = true;
  int word = id;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    fit = fit && w[d] >= -32 && w[d] < 32;
    word |= ((w[d] + 32) & 63) << (BIN_ID_BITS + 6 * d);
  }
  return fit ? word : (id | INT32_MIN);
}

// One atom of a bin-ordered cell list: its position moved into the home cell
// (x - wrap . lattice rows), index and species, so that a neighbour candidate
// is one 32-byte load.  The per-atom copy (atom_home, in atom order, written
// with the cell lists) holds the global bin id in place of the index.
// Fractional coordinate q of x in the grid's basis, with every product and
// sum rounded on its own (no contraction), so that two evaluations in one
// kernel bin an atom identically.
__device__ __forceinline__ double frac_coord(const Grid& g, const double* x, int q) {
  return __dadd_rn(__dadd_rn(__dmul_rn(x[0] - g.org[0], g.inv[q]), __dmul_rn(x[1] - g.org[1], g.inv[3 + q])),
                   __dmul_rn(x[2] - g.org[2], g.inv[6 + q]));
}

struct __align__(32) CellAtom {
  double x, y, z;
  int j, t;
};

// n . lattice rows, component q, every product and sum rounded on its own: the
// same value in every kernel, and exactly negated for -n.
__device__ __forceinline__ double lattice_comb(const int* n, const double* a, int q) {
  return __dadd_rn(__dadd_rn(__dmul_rn((double)n[0], a[q]), __dmul_rn((double)n[1], a[3 + q])),
                   __dmul_rn((double)n[2], a[6 + q]));
}

__device__ __forceinline__ CellAtom cell_atom(const double* x, const int* w, const double* a, int j, int t) {
  CellAtom c;
  c.x = x[0] - lattice_comb(w, a, 0);
  c.y = x[1] - lattice_comb(w, a, 1);
  c.z = x[2] - lattice_comb(w, a, 2);
  c.j = j;
  c.t = t;
  return c;
}

// Separation of a candidate from the centre, (x_j - x_i) + shift over home-cell
// positions: exactly antisymmetric, so that i -> j and j -> i pass the cutoff
// test together and the directed list stays symmetric.
__device__ __forceinline__ double pair_dist2(double d0, double d1, double d2) {
  return __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)), __dmul_rn(d2, d2));
}

// Workgroup barrier over LDS only: global stores still in flight are not
// waited for (k_frame_cells never reads back what it writes to global memory).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__global__ __launch_bounds__(FRAME_CELLS_THREADS) void k_frame_cells(
    int nframes, const int64_t* __restrict__ frame_off, const double* __restrict__ pos,
    const int32_t* __restrict__ types, const double* __restrict__ cell, const int32_t* __restrict__ periodic,
    double rc, Grid* __restrict__ grid, int* __restrict__ cell_start, CellAtom* __restrict__ cell_pack,
    CellAtom* __restrict__ atom_home, int* __restrict__ in_cur, int* __restrict__ atom_frame) {
  extern __shared__ int lds_cells[];
  int* bin = lds_cells;                      // [na] local bin of each atom
  int* cnt = bin + FRAME_CELLS_MAX;          // [nbins + 1] counts, then exclusive starts
  __shared__ double red_lo[FRAME_CELLS_THREADS], red_hi[FRAME_CELLS_THREADS];
  __shared__ Grid g_sh;
  const int f = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const long b0 = frame_off[f];
  const int na = (int)(frame_off[f + 1] - b0);
  // one pass of FC_PER atoms per thread covers the frame: the positions and
  // species loaded here serve the binning and the packing
  const bool one = na <= FC_PER * nt;
  const int per = grid_mask(cell + 9 * f, periodic[f]);  // the same in every thread
  const double* fpos = pos + 3 * b0;  // the frame's atoms (32-bit offsets below)
  const int32_t* ftypes = types + b0;
  double xs[FC_PER][3];
  int ts[FC_PER], bws[FC_PER];
#pragma unroll
  for (int r = 0; r < FC_PER; ++r) {
    const int k = t + r * nt;
    if (k < na) {
#pragma unroll
      for (int q = 0; q < 3; ++q) xs[r][q] = fpos[3 * k + q];
      ts[r] = ftypes[k];
    }
  }
  // bin counts start at zero (a frame has at most max(na, 1) bins)
  for (int k = t; k <= max(na, 1); k += nt) cnt[k] = 0;
  // basis rows: lattice rows for periodic directions, unit normals for open ones
  double a[9];
  grid_basis(cell + 9 * f, per, a);
  // open directions: extent of the atoms along each normal (block min / max)
  double org[3] = {0, 0, 0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (per & (1 << q)) continue;
    double lo = 1e300, hi = -1e300;
    for (int k = t; k < na; k += nt) {
      const double* x = pos + 3 * (b0 + k);
      const double v = x[0] * a[3 * q] + x[1] * a[3 * q + 1] + x[2] * a[3 * q + 2];
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
    red_lo[t] = lo;
    red_hi[t] = hi;
    lds_barrier();
    for (int w = nt / 2; w > 0; w >>= 1) {
      if (t < w) {
        red_lo[t] = fmin(red_lo[t], red_lo[t + w]);
        red_hi[t] = fmax(red_hi[t], red_hi[t + w]);
      }
      lds_barrier();
    }
    lo = na > 0 ? red_lo[0] : 0.0;
    hi = na > 0 ? red_hi[0] : 0.0;
    lds_barrier();
    const double ext = fmax(hi - lo, 1e-6) * (1 + 1e-9) + 1e-9;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      org[k] += lo * a[3 * q + k];
      a[3 * q + k] *= ext;
    }
  }
  if (t == 0) {
    Grid g;
    g.per = per;
#pragma unroll
    for (int q = 0; q < 3; ++q) g.org[q] = org[q];
    grid_finish(g, a, rc, na);
    g_sh = g;
    grid[f] = g;
  }
  lds_barrier();
  const Grid g = g_sh;
  const int nbins = g.n[0] * g.n[1] * g.n[2];
  int* cur = cnt + FRAME_CELLS_MAX + 1;  // [nbins] placement cursors
  int* tmp = cur + FRAME_CELLS_MAX;      // [na] atoms placed per bin (unordered)
  // the thread's atoms FC_PER at a time: every position loaded before any is
  // binned (the first FC_PER are already in registers)
  for (int k0 = t; k0 < na; k0 += FC_PER * nt) {
    if (k0 != t)
#pragma unroll
      for (int r = 0; r < FC_PER; ++r) {
        const int k = k0 + r * nt;
        if (k < na)
#pragma unroll
          for (int q = 0; q < 3; ++q) xs[r][q] = fpos[3 * k + q];
      }
#pragma unroll
    for (int r = 0; r < FC_PER; ++r) {
      const int k = k0 + r * nt;
      if (k >= na) break;
      double fr[3];
      for (int q = 0; q < 3; ++q) fr[q] = frac_coord(g, xs[r], q);
      int c[3], wi[3];
      for (int q = 0; q < 3; ++q) {
        const double w = (g.per & (1 << q)) ? floor(fr[q]) : 0.0;
        wi[q] = (int)w;
        c[q] = min(g.n[q] - 1, max(0, (int)((fr[q] - w) * g.n[q])));
      }
      const int id = c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
      bws[r] = bin[k] = bin_word(id, wi);
      in_cur[b0 + k] = 0;
      atom_frame[b0 + k] = f;
      atomicAdd(cnt + id, 1);
    }
  }
  lds_barrier();
  {  // exclusive scan of the bin counts: chunk sums per thread, wave prefix
     // sums, then a scan of the wave totals
    __shared__ int wtot[FRAME_CELLS_THREADS / WAVE];
    const int chunk = (nbins + nt - 1) / nt, lane = t % WAVE, wv = t / WAVE;
    const int c0 = t * chunk, c1 = min(nbins, c0 + chunk);
    int s = 0;
    for (int k = c0; k < c1; ++k) s += cnt[k];
    int incl = s;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const int y = __shfl_up(incl, d, WAVE);
      if (lane >= d) incl += y;
    }
    if (lane == WAVE - 1) wtot[wv] = incl;
    lds_barrier();
    if (wv == 0) {
      constexpr int NW = FRAME_CELLS_THREADS / WAVE;
      int x = lane < NW ? wtot[lane] : 0;
#pragma unroll
      for (int d = 1; d < NW; d <<= 1) {
        const int y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
      }
      if (lane < NW) wtot[lane] = x;
    }
    lds_barrier();
    int run = (wv ? wtot[wv - 1] : 0) + incl - s;
    for (int k = c0; k < c1; ++k) {
      const int v = cnt[k];
      cnt[k] = run;
      cur[k] = run;
      run += v;
    }
    if (t == nt - 1) cnt[nbins] = wtot[FRAME_CELLS_THREADS / WAVE - 1];
  }
  lds_barrier();
  for (int k = t; k <= na; k += nt) cell_start[b0 + k] = (int)b0 + (k < nbins ? cnt[k] : na);
  for (int k = t; k < na; k += nt) tmp[atomicAdd(cur + (bin[k] & BIN_ID_MASK), 1)] = k;
  lds_barrier();
  // ascending atom order inside each bin: each atom's rank among its bin's
  // members places it, in the binning pass's atom order
  const double* a_rows = cell + 9 * f;
  for (int k0 = t; k0 < na; k0 += FC_PER * nt) {
    if (!one)
#pragma unroll
      for (int r = 0; r < FC_PER; ++r) {
        const int k = k0 + r * nt;
        if (k < na) {
#pragma unroll
          for (int q = 0; q < 3; ++q) xs[r][q] = fpos[3 * k + q];
          ts[r] = ftypes[k];
        }
      }
#pragma unroll
    for (int r = 0; r < FC_PER; ++r) {
      const int k = k0 + r * nt;
      if (k >= na) break;
      const int bw = one ? bws[r] : bin[k], id = bw & BIN_ID_MASK;
      const int s0 = cnt[id], s1 = cnt[id + 1];
      int r0 = 0, r1 = 0, r2 = 0, r3 = 0, q = s0;
      for (; q + 3 < s1; q += 4) {  // four independent LDS reads in flight
        r0 += tmp[q] < k;
        r1 += tmp[q + 1] < k;
        r2 += tmp[q + 2] < k;
        r3 += tmp[q + 3] < k;
      }
      for (; q < s1; ++q) r0 += tmp[q] < k;
      // the wrap of the binning pass: from the bin word, or recomputed from
      // the same expression when it did not fit there
      int w[3];
      if (bw >= 0)
        for (int d = 0; d < 3; ++d) w[d] = ((bw >> (BIN_ID_BITS + 6 * d)) & 63) - 32;
      else
        for (int d = 0; d < 3; ++d) w[d] = (g.per & (1 << d)) ? (int)floor(frac_coord(g, xs[r], d)) : 0;
      const CellAtom ca = cell_atom(xs[r], w, a_rows, (int)b0 + k, ts[r]);
      cell_pack[b0 + s0 + (r0 + r1) + (r2 + r3)] = ca;
      atom_home[b0 + k] = CellAtom{ca.x, ca.y, ca.z, (int)b0 + id, 0};
    }
  }
}

constexpr size_t frame_cells_lds() { return sizeof(int) * (3 * (size_t)FRAME_CELLS_MAX + 1 + FRAME_CELLS_MAX); }

__device__ __forceinline__ void frac_of(const Grid& g, const double* x, double* f) {
  for (int q = 0; q < 3; ++q)
    f[q] = (x[0] - g.org[0]) * g.inv[q] + (x[1] - g.org[1]) * g.inv[3 + q] + (x[2] - g.org[2]) * g.inv[6 + q];
}

__global__ void k_bin(int natoms, int nframes, const int64_t* __restrict__ frame_off, const double* __restrict__ pos,
                      const Grid* __restrict__ grid, int* __restrict__ cell_of, int* __restrict__ wrap,
                      int* __restrict__ cell_cnt, int* __restrict__ atom_frame) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const int f = frame_of(frame_off, nframes, i);
  atom_frame[i] = f;
  const Grid g = grid[f];
  double fr[3];
  frac_of(g, pos + 3 * i, fr);
  int c[3];
  for (int q = 0; q < 3; ++q) {
    const double w = (g.per & (1 << q)) ? floor(fr[q]) : 0.0;
    wrap[3 * i + q] = (int)w;
    c[q] = min(g.n[q] - 1, max(0, (int)((fr[q] - w) * g.n[q])));
  }
  const int id = (int)frame_off[f] + c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
  cell_of[i] = id;
  atomicAdd(cell_cnt + id, 1);
}

__global__ void k_fill_bins(int natoms, const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                            int* __restrict__ cell_cur, int* __restrict__ cell_atoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const int c = cell_of[i];
  cell_atoms[cell_start[c] + atomicAdd(cell_cur + c, 1)] = i;
}

__global__ void k_pack_cells(int natoms, const int* __restrict__ cell_atoms, const int* __restrict__ atom_frame,
                             const double* __restrict__ pos, const int* __restrict__ wrap,
                             const int32_t* __restrict__ types, const double* __restrict__ cell,
                             const int* __restrict__ cell_of, CellAtom* __restrict__ cell_pack,
                             CellAtom* __restrict__ atom_home) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= natoms) return;
  const int j = cell_atoms[s];
  const CellAtom ca = cell_atom(pos + 3 * j, wrap + 3 * j, cell + 9 * atom_frame[j], j, types[j]);
  cell_pack[s] = ca;
  atom_home[j] = CellAtom{ca.x, ca.y, ca.z, cell_of[j], 0};
}

// Sort the ids of one list [b, e) ascending (one workgroup of LIST_THREADS
// per list; each element's rank = number of smaller ids; ids are distinct).
// With inv, also inv[id] = its final position.
constexpr int LIST_THREADS = 128;  // per-list workgroups: a ~100-entry list in one step

__device__ void list_rank_sort(int* __restrict__ list, int b, int e, int* __restrict__ tmp, int* __restrict__ inv) {
  const int lane = threadIdx.x, n = e - b;
  for (int k = lane; k < n; k += LIST_THREADS) tmp[k] = list[b + k];
  __syncthreads();
  for (int k = lane; k < n; k += LIST_THREADS) {
    const int v = tmp[k];
    int rank = 0;
    for (int q = 0; q < n; ++q) rank += tmp[q] < v;
    list[b + rank] = v;
    if (inv) inv[v] = b + rank;
  }
  __syncthreads();
}

constexpr int SORT_MAX = 2048;

__global__ __launch_bounds__(LIST_THREADS) void k_sort_lists(int nlists, const int* __restrict__ start,
                                                     int* __restrict__ list, int* __restrict__ inv) {
  __shared__ int tmp[SORT_MAX];
  const int c = blockIdx.x;
  if (c >= nlists) return;
  const int b = start[c], e = start[c + 1];
  if (e - b <= 1) {
    if (inv && e - b == 1 && threadIdx.x == 0) inv[list[b]] = b;
    return;
  }
  if (e - b <= SORT_MAX) {
    list_rank_sort(list, b, e, tmp, inv);
    return;
  }
  if (threadIdx.x == 0) {  // oversized list: serial insertion sort
    for (int k = b + 1; k < e; ++k) {
      const int v = list[k];
      int q = k - 1;
      while (q >= b && list[q] > v) { list[q + 1] = list[q]; --q; }
      list[q + 1] = v;
    }
    if (inv)
      for (int k = b; k < e; ++k) inv[list[k]] = k;
  }
}

__global__ void k_zero_int(int n, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// One wavefront per atom.  The candidate space of the atom is flattened:
// bin offsets in (o2, o1, o0) order, then the atoms of each bin in bin order.
// Up to 64 offsets at a time are resolved by the lanes (bin, image shift, bin
// size), their sizes prefix-summed across the wave, and then every lane takes
// one candidate of the flattened range (binary search for its offset), so a
// step tests 64 candidates however small the bins are.  Hits are compacted in
// candidate order with a ballot.  Collect mode (cap > 0): hits go to the
// atom's fixed-capacity slots out[i*cap + k] for k < cap and the count to
// cnt[i].  Fill mode (cap == 0): hits go to out[off[i] + k]; only atoms whose
// count exceeds `over` run (the ones whose hits did not fit their slots).
// floor(a / n) for |a| < 2^23 and 1 <= n <= 2^16, from the float reciprocal
// rn = 1 / n: the float estimate is within one of the quotient, and one step
// each way makes it exact -- a few instructions instead of an integer
// division by a value known only at run time.
__device__ __forceinline__ int floor_div_r(int a, int n, float rn) {
  int q = (int)floorf(((float)a + 0.5f) * rn);
  q -= q * n > a;
  q += (q + 1) * n <= a;
  return q;
}

template <class T>
__global__ __launch_bounds__(WAVE) void k_neighbors(int natoms, const int* __restrict__ atom_frame,
                                                    const int64_t* __restrict__ frame_off,
                                                    const double* __restrict__ cell, const Grid* __restrict__ grid,
                                                    const CellAtom* __restrict__ atom_home,
                                                    const int* __restrict__ cell_start,
                                                    const CellAtom* __restrict__ cell_pack, double rc,
                                                    const int* __restrict__ off, int* __restrict__ cnt,
                                                    Rec<T>* __restrict__ rec, int* __restrict__ oslot,
                                                    uint16_t* __restrict__ otab, int cap, int over) {
  __shared__ int o_b[2 * WAVE], o_cum[2 * WAVE], o_z[2 * WAVE], o_hits[2 * WAVE];
  __shared__ double o_sh[3][2 * WAVE];  // image shift of each bin offset (lattice rows); o_z: shift zero
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= natoms) return;
  if (cap == 0 && cnt[i] <= over) return;
  // the atom in its home cell and its bin (written by the cell-list build, the
  // same record as the cell lists hold), then its frame's grid
  const CellAtom home = atom_home[i];
  const int f = atom_frame[i];
  const Grid g = grid[f];
  const double* a = cell + 9 * f;
  const int cbase = (int)frame_off[f];
  int ci = home.j - cbase;
  const int c0 = ci % g.n[0];
  ci /= g.n[0];
  const int c1 = ci % g.n[1], c2 = ci / g.n[1];
  const double rc2 = rc * rc;
  const double hx = home.x, hy = home.y, hz = home.z;
  // a bin box whose centre is farther than rc + (bin half-diagonal) from the
  // atom holds no neighbour (1e-6 A of slack)
  const double far2 = (rc + g.rbox + 1e-6) * (rc + g.rbox + 1e-6);
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  int count = 0;
  const long w0 = cap ? (long)i * cap : (long)off[i];
  const int n0 = 2 * g.S[0] + 1, n1 = 2 * g.S[1] + 1, n2 = 2 * g.S[2] + 1, nof = n0 * n1 * n2;
  // reciprocals for the offset decode and the bin centres (the bin test has
  // 1e-6 A of slack, far above their rounding)
  const float rq0 = 1.0f / n0, rq1 = 1.0f / n1;
  const float rb0 = 1.0f / g.n[0], rb1 = 1.0f / g.n[1], rb2 = 1.0f / g.n[2];
  const double rd0 = 1.0 / g.n[0], rd1 = 1.0 / g.n[1], rd2 = 1.0 / g.n[2];
  // offsets OPL at a time, two adjacent ones per lane (o = ob + 2 lane, +1),
  // so that a 5 x 5 x 5 search resolves all its bins in one chunk
  constexpr int OPL = 2 * WAVE;
  for (int ob = 0; ob < nof; ob += OPL) {
    int nb2[2], bs2[2], sv2[2][3];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int o = ob + 2 * lane + u;
      int nb = 0, bs = 0, s0 = 0, s1 = 0, s2 = 0;
      if (o < nof) {
        const int oq = floor_div_r(o, n0, rq0), oz = floor_div_r(oq, n1, rq1);
        int x = c0 + (o - oq * n0) - g.S[0], y = c1 + (oq - oz * n1) - g.S[1], z = c2 + oz - g.S[2];
        bool ok = true;
        if (g.per & 1) { s0 = floor_div_r(x, g.n[0], rb0); x -= s0 * g.n[0]; }
        else ok = ok && x >= 0 && x < g.n[0];
        if (g.per & 2) { s1 = floor_div_r(y, g.n[1], rb1); y -= s1 * g.n[1]; }
        else ok = ok && y >= 0 && y < g.n[1];
        if (g.per & 4) { s2 = floor_div_r(z, g.n[2], rb2); z -= s2 * g.n[2]; }
        else ok = ok && z >= 0 && z < g.n[2];
        if (ok) {
          const double u0 = (x + 0.5) * rd0 + s0, u1 = (y + 0.5) * rd1 + s1, u2 = (z + 0.5) * rd2 + s2;
          const double ex = g.org[0] + u0 * g.A[0] + u1 * g.A[3] + u2 * g.A[6] - hx;
          const double ey = g.org[1] + u0 * g.A[1] + u1 * g.A[4] + u2 * g.A[7] - hy;
          const double ez = g.org[2] + u0 * g.A[2] + u1 * g.A[5] + u2 * g.A[8] - hz;
          ok = ex * ex + ey * ey + ez * ez <= far2;
        }
        if (ok) {
          const int c = cbase + x + g.n[0] * (y + g.n[1] * z);
          bs = cell_start[c];
          nb = cell_start[c + 1] - bs;
        }
      }
      nb2[u] = nb;
      bs2[u] = bs;
      sv2[u][0] = s0; sv2[u][1] = s1; sv2[u][2] = s2;
    }
    const int pair = nb2[0] + nb2[1];
    int incl = pair;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const int t = __shfl_up(incl, d, WAVE);
      if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, WAVE - 1, WAVE);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = 2 * lane + u;
      o_b[q] = bs2[u];
      o_cum[q] = incl - pair + (u ? nb2[0] : 0);
      // open directions have zero shifts
      o_sh[0][q] = lattice_comb(sv2[u], a, 0);
      o_sh[1][q] = lattice_comb(sv2[u], a, 1);
      o_sh[2][q] = lattice_comb(sv2[u], a, 2);
      o_z[q] = (sv2[u][0] | sv2[u][1] | sv2[u][2]) == 0;
      o_hits[q] = 0;
    }
    lds_barrier();  // LDS tables only (the records in flight are not read here)
    const int nvalid = min(OPL, nof - ob), chunk_start = count;
    // candidate k of the flattened range: its bin offset (binary search), the
    // packed atom, and whether it is a neighbour (d = (x_j - x_i) + shift, both
    // in the home cell; the atom itself at zero shift is not)
    auto probe = [&](int k, bool& hit, double& d0, double& d1, double& d2, int& j, int& t, int& oi) {
      int lo = 0, hi = nvalid;  // largest q with o_cum[q] <= k
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (o_cum[mid] <= k) lo = mid;
        else hi = mid;
      }
      const CellAtom p = cell_pack[o_b[lo] + k - o_cum[lo]];
      const double sx = o_sh[0][lo], sy = o_sh[1][lo], sz = o_sh[2][lo];
      j = p.j;
      t = p.t;
      d0 = (p.x - hx) + sx;
      d1 = (p.y - hy) + sy;
      d2 = (p.z - hz) + sz;
      const bool self = j == i && o_z[lo];
      hit = !self && pair_dist2(d0, d1, d2) < rc2;
      oi = ob + lo;
    };
    // two candidates per lane per step: their loads overlap
    for (int k0 = 0; k0 < total; k0 += 2 * WAVE) {
      const int ka = k0 + lane, kb = k0 + WAVE + lane;
      bool ha = false, hb = false;
      double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
      int ja = 0, ta = 0, jb = 0, tb = 0, oa = 0, obb = 0;
      if (ka < total) probe(ka, ha, a0, a1, a2, ja, ta, oa);
      if (kb < total) probe(kb, hb, b0, b1, b2, jb, tb, obb);
      const unsigned long long ma = __ballot(ha);
      const int sa = count + __popcll(ma & below);
      if (ha && (cap == 0 || sa < cap)) {
        rec[w0 + sa] = Rec<T>{T(a0), T(a1), T(a2), rec_jbits(ja, ta, T(0))};
        if (oslot) oslot[w0 + sa] = oa;
      }
      if (ha && otab) atomicAdd(&o_hits[oa - ob], 1);
      count += __popcll(ma);
      const unsigned long long mb = __ballot(hb);
      const int sb = count + __popcll(mb & below);
      if (hb && (cap == 0 || sb < cap)) {
        rec[w0 + sb] = Rec<T>{T(b0), T(b1), T(b2), rec_jbits(jb, tb, T(0))};
        if (oslot) oslot[w0 + sb] = obb;
      }
      if (hb && otab) atomicAdd(&o_hits[obb - ob], 1);
      count += __popcll(mb);
    }
    lds_barrier();
    if (otab && nof <= NOF_TAB) {  // hits before each offset of the chunk
      const int h0 = o_hits[2 * lane], h1 = o_hits[2 * lane + 1], hp = h0 + h1;
      int inc = hp;
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const int y = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += y;
      }
      const int e0 = chunk_start + inc - hp;
      const size_t trow = (size_t)i * NOF_TAB + ob + 2 * lane;
      if (ob + 2 * lane < nof) otab[trow] = (uint16_t)e0;
      if (ob + 2 * lane + 1 < nof) otab[trow + 1] = (uint16_t)(e0 + h0);
    }
  }
  if (cap && lane == 0) cnt[i] = count;
}

// Copy every atom's collected hits from its fixed-capacity slots to the CSR
// records (atoms whose count exceeded the capacity are filled directly).
template <class T>
__global__ __launch_bounds__(WAVE) void k_compact(int natoms, int cap, const int* __restrict__ cnt,
                                                  const int* __restrict__ off, const Rec<T>* __restrict__ slots,
                                                  Rec<T>* __restrict__ rec) {
  const int i = blockIdx.x;
  if (i >= natoms) return;
  const int n = cnt[i];
  if (n > cap) return;
  const Rec<T>* src = slots + (size_t)i * cap;
  Rec<T>* dst = rec + off[i];
  for (int k = threadIdx.x; k < n; k += WAVE) dst[k] = src[k];
}

// Copy of every atom's hits from its slots to the CSR records, and the
// position of each edge's reverse, which places the edge in its target's
// incoming list without atomics: the lists are exactly symmetric, so for
// e = i -> t found at bin offset o, t found i at the mirrored offset
// nof - 1 - o (image shift negated), in i's bin.  Its position in t's list
// is the number of t's hits at earlier offsets (binary search over t's
// ascending hit offsets) plus t's hits among the atoms of i's bin before i,
// recomputed with the neighbour search's own expressions.  Incoming lists
// are thereby in the target's outgoing order.  Requires every count within
// the slots.
template <class T>
__global__ __launch_bounds__(WAVE) void k_compact_rev(int natoms, int cap, const int* __restrict__ cnt,
                                                      const int* __restrict__ off,
                                                      const Rec<T>* __restrict__ slots,
                                                      const int* __restrict__ oslot,
                                                      const uint16_t* __restrict__ otab, Rec<T>* __restrict__ rec,
                                                      int* __restrict__ in_pos, const int* __restrict__ atom_frame,
                                                      const int64_t* __restrict__ frame_off,
                                                      const Grid* __restrict__ grid, const double* __restrict__ cell,
                                                      const CellAtom* __restrict__ atom_home,
                                                      const int* __restrict__ cell_start,
                                                      const CellAtom* __restrict__ cell_pack, double rc) {
  const int i = blockIdx.x;
  if (i >= natoms) return;
  const int n = cnt[i];
  const int f = atom_frame[i];
  const Grid& g = grid[f];
  const double* a = cell + 9 * f;
  const int cbase = (int)frame_off[f];
  const int n0 = 2 * g.S[0] + 1, n1 = 2 * g.S[1] + 1, n2 = 2 * g.S[2] + 1, nof = n0 * n1 * n2;
  const float rq0 = 1.0f / n0, rq1 = 1.0f / n1;
  const float rb0 = 1.0f / g.n[0], rb1 = 1.0f / g.n[1], rb2 = 1.0f / g.n[2];
  const bool tabled = nof <= NOF_TAB;
  const int bin_i = atom_home[i].j;
  const int bin_start = cell_start[bin_i], bin_end = cell_start[bin_i + 1];
  const double rc2 = rc * rc;
  // i's bin (the same for all of i's edges) staged once when it is small
  __shared__ CellAtom sbin[WAVE];
  const bool staged = bin_end - bin_start <= WAVE;
  if (staged && bin_start + (int)threadIdx.x < bin_end) sbin[threadIdx.x] = cell_pack[bin_start + threadIdx.x];
  __syncthreads();
  const Rec<T>* src = slots + (size_t)i * cap;
  const int* osrc = oslot + (size_t)i * cap;
  const int o_i = off[i];
  for (int k = threadIdx.x; k < n; k += WAVE) {
    const Rec<T> q = src[k];
    rec[o_i + k] = q;
    const int t = rec_j(q), os = nof - 1 - osrc[k];
    const CellAtom ht = atom_home[t];  // home position and bin of t
    const int t0 = off[t], nt = off[t + 1] - t0;
    // t's hits at offsets before os
    int lo;
    if (tabled) {
      lo = otab[(size_t)t * NOF_TAB + os];
    } else {  // binary search over t's ascending hit offsets
      const int* ot = oslot + (size_t)t * cap;
      int hi = nt;
      lo = 0;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ot[mid] < os) lo = mid + 1;
        else hi = mid;
      }
    }
    // t's bin coordinates and the image shift of offset os, as t's search forms them
    int ct = ht.j - cbase;
    const int c0 = ct % g.n[0];
    ct /= g.n[0];
    const int c1 = ct % g.n[1], c2 = ct / g.n[1];
    const int oq = floor_div_r(os, n0, rq0), oz = floor_div_r(oq, n1, rq1);
    const int x = c0 + (os - oq * n0) - g.S[0], y = c1 + (oq - oz * n1) - g.S[1], z = c2 + oz - g.S[2];
    int sv[3] = {0, 0, 0};
    if (g.per & 1) sv[0] = floor_div_r(x, g.n[0], rb0);
    if (g.per & 2) sv[1] = floor_div_r(y, g.n[1], rb1);
    if (g.per & 4) sv[2] = floor_div_r(z, g.n[2], rb2);
    const double sx = lattice_comb(sv, a, 0), sy = lattice_comb(sv, a, 1), sz = lattice_comb(sv, a, 2);
    const bool zero = (sv[0] | sv[1] | sv[2]) == 0;
    int before = 0;
    for (int p = bin_start; p < bin_end; ++p) {  // i is in this bin: the loop ends at it
      const CellAtom c = staged ? sbin[p - bin_start] : cell_pack[p];
      if (c.j == i) break;
      if (c.j == t && zero) continue;
      const double d0 = (c.x - ht.x) + sx, d1 = (c.y - ht.y) + sy, d2 = (c.z - ht.z) + sz;
      before += pair_dist2(d0, d1, d2) < rc2;
    }
    // (clamped into t's list: a position outside it would mean a broken symmetry)
    in_pos[o_i + k] = t0 + max(0, min(lo + before, nt - 1));
  }
}

// Incoming edges of every atom (edges e' = j -> i grouped by i).  The full
// list is symmetric, so atom i has as many incoming as outgoing edges and the
// outgoing offsets serve both; lists are sorted afterwards for a fixed order.
template <class T>
__global__ void k_incoming(long nedges, const Rec<T>* __restrict__ rec, const int* __restrict__ off,
                           int* __restrict__ cur, int* __restrict__ in_edges) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nedges) return;
  const int dst = rec_j(rec[e]);
  in_edges[off[dst] + atomicAdd(cur + dst, 1)] = (int)e;
}

// ------------------------------------------------------------------ finalize
// Frame reduction in chunks of FIN_CHUNK atoms: block (f, s) covers atoms
// [frame_off[f] + s FIN_CHUNK, ...) of frame f, writes their FP64 forces and
// site energies, and sums energy and virial in a fixed order (per-thread
// strided sums, a fixed wave-shuffle tree, then the waves in order).  With
// one chunk per frame (nsub == 1) the block writes the frame's energy and
// virial; otherwise its partial.  Up to FIN_FOLD chunks per frame (done given),
// the last of the frame's chunks to finish (a per-frame count, left at zero
// again by that chunk) adds the partials in a fixed order -- which chunk is
// last does not change the sum; more chunks than that would queue on the one
// counter, and k_finalize_frames adds their partials in the same order.
// Chunks of 512 atoms spread a large frame over many compute units (one block
// per 4096-atom frame was a single CU's load bandwidth).
constexpr int FIN_THREADS = 256, FIN_CHUNK = 512, FIN_FOLD = 16;

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_finalize(int nsub, const int64_t* __restrict__ frame_off,
                                                          const int32_t* __restrict__ types, SiteConst sc,
                                                          const T* __restrict__ site, const T* __restrict__ fw,
                                                          double* __restrict__ energy, double* __restrict__ virial,
                                                          double* __restrict__ force, double* __restrict__ site_out,
                                                          double* __restrict__ part, int* __restrict__ done) {
  constexpr int CMAX = 64;  // per-species site constants staged in LDS
  __shared__ double red[7][FIN_THREADS / WAVE];
  __shared__ double spc[CMAX];
  const int f = blockIdx.x, s = blockIdx.y, t = threadIdx.x;
  if (t < CMAX && t < sc.C) spc[t] = sc(t);
  __syncthreads();
  const long fe = frame_off[f + 1];
  const long b = frame_off[f] + (long)s * FIN_CHUNK, e = min(fe, b + FIN_CHUNK);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  // all of the thread's loads first (FIN_CHUNK / FIN_THREADS atoms), then
  // the sums and stores in atom order
  constexpr int PER = FIN_CHUNK / FIN_THREADS;
  T wv[PER][9], sv[PER];
  int tv[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const long i = b + t + (long)p * FIN_THREADS;
    if (i < e) {
#pragma unroll
      for (int q = 0; q < 9; ++q) wv[p][q] = fw[9 * i + q];
      sv[p] = site[i];
      tv[p] = types[i];
    }
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const long i = b + t + (long)p * FIN_THREADS;
    if (i < e) {
      const int ty = tv[p];
      const double se = (double)sv[p] + (ty < CMAX ? spc[ty] : sc(ty));
      acc[0] += se;
#pragma unroll
      for (int q = 0; q < 6; ++q) acc[1 + q] += (double)wv[p][3 + q];
      force[3 * i] = (double)wv[p][0];
      force[3 * i + 1] = (double)wv[p][1];
      force[3 * i + 2] = (double)wv[p][2];
      if (site_out) site_out[i] = se;
    }
  }
  wave_sum_n(acc);
  if (t % WAVE == 0)
#pragma unroll
    for (int q = 0; q < 7; ++q) red[q][t / WAVE] = acc[q];
  __syncthreads();
  if (t < 7) {
    double v = 0;
    for (int k = 0; k < FIN_THREADS / WAVE; ++k) v += red[t][k];
    if (nsub == 1) {
      if (t == 0) energy[f] = v;
      else virial[6 * f + t - 1] = v;
    } else {
      part[((size_t)f * nsub + s) * 7 + t] = v;
    }
  }
  if (nsub == 1 || !done) return;
  // the partial (written by wave 0) is visible device-wide before it is counted
  __shared__ int last;
  if (t == 0) {
    __threadfence();
    last = atomicAdd(done + f, 1) == nsub - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // the other chunks' partials, counted before this one
  // the frame's partials: per-thread strided sums, a fixed wave-shuffle
  // tree, the waves in order
  double sum[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int c = t; c < nsub; c += FIN_THREADS)
#pragma unroll
    for (int q = 0; q < 7; ++q) sum[q] += part[((size_t)f * nsub + c) * 7 + q];
  wave_sum_n(sum);
  if (t % WAVE == 0)
#pragma unroll
    for (int q = 0; q < 7; ++q) red[q][t / WAVE] = sum[q];
  __syncthreads();
  if (t < 7) {
    double v = 0;
    for (int k = 0; k < FIN_THREADS / WAVE; ++k) v += red[t][k];
    if (t == 0) energy[f] = v;
    else virial[6 * f + t - 1] = v;
  }
  if (t == 0) done[f] = 0;  // zero again for the next call
}

// One block per frame: the frame's chunk partials summed in the order of the
// folded form above.
__global__ __launch_bounds__(FIN_THREADS) void k_finalize_frames(int nsub, const double* __restrict__ part,
                                                                 double* __restrict__ energy,
                                                                 double* __restrict__ virial) {
  __shared__ double red[7][FIN_THREADS / WAVE];
  const int f = blockIdx.x, t = threadIdx.x;
  double sum[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int c = t; c < nsub; c += FIN_THREADS)
#pragma unroll
    for (int q = 0; q < 7; ++q) sum[q] += part[((size_t)f * nsub + c) * 7 + q];
  wave_sum_n(sum);
  if (t % WAVE == 0)
#pragma unroll
    for (int q = 0; q < 7; ++q) red[q][t / WAVE] = sum[q];
  __syncthreads();
  if (t < 7) {
    double v = 0;
    for (int k = 0; k < FIN_THREADS / WAVE; ++k) v += red[t][k];
    if (t == 0) energy[f] = v;
    else virial[6 * f + t - 1] = v;
  }
}

// Forces of atoms [first, natoms) outside every frame (the ghosts of a pair-style call).
template <class T>
__global__ void k_atom_out(int first, int natoms, const T* __restrict__ fw, double* __restrict__ force) {
  const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  force[3 * i] = (double)fw[9 * (size_t)i];
  force[3 * i + 1] = (double)fw[9 * (size_t)i + 1];
  force[3 * i + 2] = (double)fw[9 * (size_t)i + 2];
}

template <class T>
__global__ void k_fill_zero(size_t n, T* __restrict__ a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = T(0);
}

template <class T>
void fill_zero(Workspace& ws, T* p, size_t n, hipStream_t st) {
  if (n == 0) return;
  SUS2_LAUNCH(ws, st, "fill_zero", k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, n, p);
}
template void fill_zero<float>(Workspace&, float*, size_t, hipStream_t);
template void fill_zero<double>(Workspace&, double*, size_t, hipStream_t);
template void fill_zero<int>(Workspace&, int*, size_t, hipStream_t);

// ------------------------------------------------------------------ driver
static void grow_slots(Workspace& ws, long maxcnt);

// Records from the slots and the incoming positions, when every count fits
// the slots.
template <class T>
static void compact_rev(Workspace& ws, const Batch& b, double rc, hipStream_t st) {
  const int N = b.natoms;
  const std::string kn = std::string("compact<") + tname<T>() + ">";
  SUS2_LAUNCH(ws, st, kn, k_compact_rev<T>, dim3(N), dim3(WAVE), 0, N,
              ws.nbr_cap, ws.nbr_cnt.get<int>(), ws.nbr_off.get<int>(), ws.nbr_slots.get<Rec<T>>(),
              ws.nbr_oslot.get<int>(), ws.nbr_otab.get<uint16_t>(), ws.rec.get<Rec<T>>(), ws.in_pos.get<int>(),
              ws.atom_frame.get<int>(), b.frame_off, ws.grid.get<Grid>(), b.cell, ws.atom_home.get<CellAtom>(),
              ws.cell_start.get<int>(), ws.cell_pack.get<CellAtom>(), rc);
}

// Largest frame of a batch (atoms), written with its tag (the sequence
// number in the upper 32 bits) into word 2 of the coherent host report.
__global__ __launch_bounds__(256) void k_frame_max(int nframes, const int64_t* __restrict__ frame_off,
                                                   int64_t* report, int64_t seq) {
  __shared__ int64_t red[256 / WAVE];
  const int t = threadIdx.x;
  int64_t m = 0;
  for (int f = t; f < nframes; f += 256) m = max(m, frame_off[f + 1] - frame_off[f]);
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) m = max(m, (int64_t)__shfl_xor(m, d, WAVE));
  if (t % WAVE == 0) red[t / WAVE] = m;
  __syncthreads();
  if (t == 0) {
    for (int k = 1; k < 256 / WAVE; ++k) m = max(m, red[k]);
    const uint64_t tag = (uint64_t)(seq & 0x7fffffff) << 32;
    __hip_atomic_store(report, (int64_t)(tag | (uint64_t)min(m, (int64_t)0x7fffffff)), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Largest frame of the batch, read by polling the host report (no copy and
// stream synchronisation on the call's critical path).
static int64_t largest_frame(Workspace& ws, const Batch& b, hipStream_t st) {
  if (b.nframes == 1) return b.natoms;  // one frame: no device round trip
  if (current_arena()) {  // allocation-free call: a device word and a plain copy
    ws.scan_tmp.ensure(8 * sizeof(int64_t));
    int64_t* word = ws.scan_tmp.get<int64_t>() + 7;
    SUS2_LAUNCH(ws, st, "frame_max", k_frame_max, dim3(1), dim3(256), 0, b.nframes, b.frame_off, word, int64_t(0));
    int64_t v = 0;
    SUS2_CHECK(hipMemcpyAsync(&v, word, sizeof(v), hipMemcpyDeviceToHost, st));
    SUS2_CHECK(hipStreamSynchronize(st));
    return v & 0x7fffffff;
  }
  const int64_t seq = ++ws.report_seq;
  SUS2_LAUNCH(ws, st, "frame_max", k_frame_max, dim3(1), dim3(256), 0, b.nframes, b.frame_off, report_word(ws, 2),
              seq);
  uint64_t w = 0;
  poll_report(ws, 2, 1, seq, st, "frame size query", &w);
  return (int64_t)(w & 0x7fffffff);
}

template <class T>
void build_neighbors(Workspace& ws, const Batch& b, double rc, hipStream_t st) {
  const int N = b.natoms;
  ws.ensure_atoms(N);
  const size_t n1 = std::max(1, N);
  ws.grid.ensure(sizeof(Grid) * std::max(1, b.nframes));
  ws.cell_of.ensure(sizeof(int) * n1);
  ws.atom_frame.ensure(sizeof(int) * n1);
  ws.wrap.ensure(sizeof(int) * 3 * n1);
  ws.cell_cnt.ensure(sizeof(int) * n1);
  ws.cell_start.ensure(sizeof(int) * (n1 + 1));
  ws.cell_cur.ensure(sizeof(int) * n1);
  ws.cell_atoms.ensure(sizeof(int) * n1);
  ws.cell_pack.ensure(sizeof(CellAtom) * n1);
  ws.atom_home.ensure(sizeof(CellAtom) * n1);
  ws.nbr_otab.ensure(sizeof(uint16_t) * n1 * NOF_TAB);
  ws.in_cur.ensure(sizeof(int) * n1);
  ws.scan_tmp.ensure(8 * sizeof(int64_t));
  const int cap = ws.nbr_cap;
  ws.nbr_slots.ensure(sizeof(Rec<T>) * n1 * cap);
  ws.nbr_oslot.ensure(sizeof(int) * n1 * cap);
  const int nb = (N + 255) / 256;
  const std::string tn = std::string("<") + tname<T>() + ">";
  // frames of at most FRAME_CELLS_MAX atoms build their cell lists in one
  // workgroup each; otherwise the multi-kernel global path
  bool per_frame = N <= FRAME_CELLS_MAX;
  ws.max_frame_atoms = N;
  if (!per_frame && b.nframes > 0) {
    const int64_t mx = ws.known_max_frame >= 0 ? ws.known_max_frame : largest_frame(ws, b, st);
    ws.max_frame_atoms = mx;
    per_frame = mx <= FRAME_CELLS_MAX;
  }
  if (per_frame) {
    if (b.nframes > 0) {
      static bool lds_set = false;
      if (!lds_set) {
        SUS2_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_frame_cells),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)frame_cells_lds()));
        lds_set = true;
      }
      SUS2_LAUNCH(ws, st, "frame_cells", k_frame_cells, dim3(b.nframes), dim3(FRAME_CELLS_THREADS),
                  frame_cells_lds(), b.nframes, b.frame_off, b.pos, b.types, b.cell, b.periodic, rc,
                  ws.grid.get<Grid>(), ws.cell_start.get<int>(), ws.cell_pack.get<CellAtom>(),
                  ws.atom_home.get<CellAtom>(), ws.in_cur.get<int>(), ws.atom_frame.get<int>());
    }
  } else {
    SUS2_LAUNCH(ws, st, "grid_setup", k_grid_setup, dim3((std::max(b.nframes, N) + 63) / 64), dim3(64), 0,
                b.nframes, N, b.frame_off, b.pos, b.cell, b.periodic, rc, ws.grid.get<Grid>(), ws.cell_cnt.get<int>(),
                ws.cell_cur.get<int>(), ws.in_cur.get<int>());
    SUS2_LAUNCH(ws, st, "bin_atoms", k_bin, dim3(nb), dim3(256), 0, N, b.nframes, b.frame_off, b.pos,
                ws.grid.get<Grid>(), ws.cell_of.get<int>(), ws.wrap.get<int>(), ws.cell_cnt.get<int>(),
                ws.atom_frame.get<int>());
    device_scan(ws.cell_cnt.get<int>(), ws.cell_start.get<int>(), N, ws.scan_tmp.get<int64_t>() + 4, st);
    SUS2_LAUNCH(ws, st, "fill_bins", k_fill_bins, dim3(nb), dim3(256), 0, N, ws.cell_of.get<int>(),
                ws.cell_start.get<int>(), ws.cell_cur.get<int>(), ws.cell_atoms.get<int>());
    SUS2_LAUNCH(ws, st, "sort_bins", k_sort_lists, dim3(N), dim3(LIST_THREADS), 0, N, ws.cell_start.get<int>(),
                ws.cell_atoms.get<int>(), (int*)nullptr);
    SUS2_LAUNCH(ws, st, "pack_cells", k_pack_cells, dim3(nb), dim3(256), 0, N, ws.cell_atoms.get<int>(),
                ws.atom_frame.get<int>(), b.pos, ws.wrap.get<int>(), b.types, b.cell, ws.cell_of.get<int>(),
                ws.cell_pack.get<CellAtom>(), ws.atom_home.get<CellAtom>());
  }
  if (N > 0) {
    SUS2_LAUNCH(ws, st, "neighbors" + tn, k_neighbors<T>, dim3(N), dim3(WAVE), 0, N, ws.atom_frame.get<int>(),
                b.frame_off, b.cell, ws.grid.get<Grid>(), ws.atom_home.get<CellAtom>(), ws.cell_start.get<int>(),
                ws.cell_pack.get<CellAtom>(), rc, (const int*)nullptr, ws.nbr_cnt.get<int>(),
                ws.nbr_slots.get<Rec<T>>(), ws.nbr_oslot.get<int>(), ws.nbr_otab.get<uint16_t>(), cap, 0);
  }
  // edge capacity left by earlier builds: the host-sized one, in records of T
  // (outside capture the edge buffers are widened to it first, as after a
  // change of precision)
  if (!ws.capture && ws.nbr_known && ws.edge_cap > 0) {
    ws.rec.ensure(sizeof(Rec<T>) * ws.edge_cap);
    ws.in_edges.ensure(sizeof(int) * ws.edge_cap);
    ws.in_pos.ensure(sizeof(int) * ws.edge_cap);
  }
  const long ecap = (long)std::min({ws.rec.bytes / sizeof(Rec<T>), ws.in_edges.bytes / sizeof(int),
                                    ws.in_pos.bytes / sizeof(int), (size_t)INT32_MAX - 1,
                                    (size_t)(ws.edge_cap > 0 ? ws.edge_cap : INT32_MAX - 1)});
  if (ws.capture && !(ws.nbr_known && ecap > 0))
    throw Error(SUS2_E_INVALID, "capture-safe neighbour build needs the edge capacity of an earlier call");
  if ((ws.nbr_optimistic || ws.capture) && ws.nbr_known && ecap > 0) {
    // no host round trip: the scan reports (and empties every list) when the
    // counts exceed the slots or the edge buffers; neighbors_verify reads the
    // report once the call's work has been queued.  Capture-safe calls set the
    // sticky device flag instead and nothing is read back.
    if (ws.capture) {
      device_scan(ws.nbr_cnt.get<int>(), ws.nbr_off.get<int>(), N, ws.scan_tmp.get<int64_t>(), st, cap, ecap, nullptr,
                  0, ws.overflow_flag);
    } else {
      device_scan(ws.nbr_cnt.get<int>(), ws.nbr_off.get<int>(), N, ws.scan_tmp.get<int64_t>(), st, cap, ecap,
                  report_word(ws, 0), ws.scan_seq = ++ws.report_seq);
      ws.nbr_pending = true;
    }
    ws.nedges = ecap;
    if (N > 0) compact_rev<T>(ws, b, rc, st);
    return;
  }
  long maxcnt = 0;
  const long total = exclusive_scan(ws, N, st, &maxcnt);
  if (total > INT32_MAX - 1) throw Error(SUS2_E_RANGE, "neighbour list exceeds 2^31 edges");
  ws.nedges = total;
  ws.nbr_known = true;
  // edge buffers with 1/8 headroom, so that later builds of similar frames fit
  const long ebuf = std::max<long>(total + total / 8, 1);
  ws.edge_cap = ebuf;
  ws.rec.ensure(sizeof(Rec<T>) * ebuf);
  ws.in_edges.ensure(sizeof(int) * ebuf);
  ws.in_pos.ensure(sizeof(int) * ebuf);
  if (N > 0 && maxcnt <= cap) {
    compact_rev<T>(ws, b, rc, st);
  } else if (N > 0) {
    // atoms with more neighbours than slots: the others are copied, these are
    // filled directly, and the incoming lists are placed (atomic cursors) and
    // sorted into a fixed order afterwards
    SUS2_LAUNCH(ws, st, "compact" + tn, k_compact<T>, dim3(N), dim3(WAVE), 0, N, cap, ws.nbr_cnt.get<int>(),
                ws.nbr_off.get<int>(), ws.nbr_slots.get<Rec<T>>(), ws.rec.get<Rec<T>>());
    SUS2_LAUNCH(ws, st, "neighbors_fill" + tn, k_neighbors<T>, dim3(N), dim3(WAVE), 0, N, ws.atom_frame.get<int>(),
                b.frame_off, b.cell, ws.grid.get<Grid>(), ws.atom_home.get<CellAtom>(), ws.cell_start.get<int>(),
                ws.cell_pack.get<CellAtom>(), rc, ws.nbr_off.get<int>(), ws.nbr_cnt.get<int>(), ws.rec.get<Rec<T>>(),
                (int*)nullptr, (uint16_t*)nullptr, 0, cap);
    SUS2_LAUNCH(ws, st, "incoming" + tn, k_incoming<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, total,
                ws.rec.get<Rec<T>>(), ws.nbr_off.get<int>(), ws.in_cur.get<int>(), ws.in_edges.get<int>());
    SUS2_LAUNCH(ws, st, "sort_incoming", k_sort_lists, dim3(N), dim3(LIST_THREADS), 0, N, ws.nbr_off.get<int>(),
                ws.in_edges.get<int>(), ws.in_pos.get<int>());
  }
  grow_slots(ws, maxcnt);
}

// slots for the next build: the largest count seen, with headroom, a multiple of 16
static void grow_slots(Workspace& ws, long maxcnt) {
  const long want = std::max<long>(16, ((maxcnt + maxcnt / 4 + 15) / 16) * 16);
  if (want > ws.nbr_cap) ws.nbr_cap = (int)std::min<long>(want, 4096);
}


bool neighbors_verify(Workspace& ws, hipStream_t st) {
  if (!ws.nbr_pending) return false;
  ws.nbr_pending = false;
  // the scan runs early in the queued work: poll its report
  uint64_t w[2];
  poll_report(ws, 0, 2, ws.scan_seq, st, "neighbour build", w);
  const uint64_t w0 = w[0], w1 = w[1];
  const long total = (long)(w1 & 0x7fffffff), maxcnt = (long)(w0 & 0xffffffff);
  const bool over = (w1 >> 31) & 1;
  grow_slots(ws, maxcnt);
  if (over) {  // the queued work ran on empty lists: drain it; the next build sizes on the host
    SUS2_CHECK(hipStreamSynchronize(st));
    ws.nbr_known = false;
    return true;
  }
  ws.nedges = total;
  return false;
}

template <class T>
void finish_efs(Workspace& ws, const Batch& b, const EfsOut& out, const SiteConst& sc, hipStream_t st, int nsite) {
  const int N = b.natoms;
  const int first = nsite < 0 ? N : nsite;  // atoms past the frames (pair-style ghosts)
  if (b.nframes > 0) {
    const long mx = std::max<long>(1, std::min<long>(ws.max_frame_atoms, N));
    const int nsub = (int)((mx + FIN_CHUNK - 1) / FIN_CHUNK);
    const bool fold = nsub > 1 && nsub <= FIN_FOLD;
    if (nsub > 1) ws.fin_part.ensure(sizeof(double) * 7 * (size_t)b.nframes * nsub);
    if (fold) {
      // per-frame chunk counts: zero when the buffer is new (each call leaves
      // them at zero)
      const void* p0 = ws.fin_done.p;
      ws.fin_done.ensure(sizeof(int) * (size_t)b.nframes);
      if (ws.fin_done.p != p0) fill_zero(ws, ws.fin_done.get<int>(), ws.fin_done.bytes / sizeof(int), st);
    }
    SUS2_LAUNCH(ws, st, "finalize", k_finalize<T>, dim3(b.nframes, nsub), dim3(FIN_THREADS), 0, nsub, b.frame_off,
                b.types, sc, ws.site.get<T>(), ws.fw.get<T>(), out.energy, out.virial, out.force, out.site,
                ws.fin_part.get<double>(), fold ? ws.fin_done.get<int>() : (int*)nullptr);
    if (nsub > FIN_FOLD)
      SUS2_LAUNCH(ws, st, "finalize_frames", k_finalize_frames, dim3(b.nframes), dim3(FIN_THREADS), 0, nsub,
                  ws.fin_part.get<double>(), out.energy, out.virial);
  }
  if (first < N)
    SUS2_LAUNCH(ws, st, "atom_out", k_atom_out<T>, dim3((N - first + 255) / 256), dim3(256), 0, first, N,
                ws.fw.get<T>(), out.force);
}

template void build_neighbors<float>(Workspace&, const Batch&, double, hipStream_t);
template void build_neighbors<double>(Workspace&, const Batch&, double, hipStream_t);
template void finish_efs<float>(Workspace&, const Batch&, const EfsOut&, const SiteConst&, hipStream_t, int);
template void finish_efs<double>(Workspace&, const Batch&, const EfsOut&, const SiteConst&, hipStream_t, int);

// ------------------------------------------------------------------ pair-style lists
// CSR of caller-supplied directed edges (i owned -> j any, d = x_j - x_i):
// edges grouped by centre in their input order, then the incoming lists of all
// atoms (CSR positions grouped by target, ascending).
__global__ void k_count_keys(long n, const int32_t* __restrict__ keys, int* __restrict__ cnt) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) atomicAdd(cnt + keys[e], 1);
}

__global__ void k_place_keys(long n, const int32_t* __restrict__ keys, const int* __restrict__ off,
                             int* __restrict__ cur, int* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[off[keys[e]] + atomicAdd(cur + keys[e], 1)] = (int)e;
}

template <class T>
__global__ void k_build_rec(long n, const int* __restrict__ order, const int32_t* __restrict__ ej,
                            const double* __restrict__ d, const int32_t* __restrict__ types,
                            Rec<T>* __restrict__ rec, int32_t* __restrict__ tgt) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int e = order[p];
  const int j = ej[e];
  rec[p] = Rec<T>{T(d[3 * (size_t)e]), T(d[3 * (size_t)e + 1]), T(d[3 * (size_t)e + 2]),
                  rec_jbits(j, types[j], T(0))};
  tgt[p] = j;
}

template <class T>
void build_pair_lists(Workspace& ws, int nowned, int nall, long nedges, const int32_t* ei, const int32_t* ej,
                      const double* d, const int32_t* types, int* in_off, hipStream_t st) {
  const size_t n1 = std::max(1, nall), e1 = std::max<long>(1, nedges);
  ws.ensure_atoms(nowned);
  ws.cell_cnt.ensure(sizeof(int) * n1);
  ws.cell_cur.ensure(sizeof(int) * n1);
  ws.cell_atoms.ensure(sizeof(int) * e1);  // edge order by centre
  ws.wrap.ensure(sizeof(int32_t) * e1);    // target of each CSR position
  ws.scan_tmp.ensure(4 * sizeof(int64_t));
  ws.rec.ensure(sizeof(Rec<T>) * e1);
  ws.in_edges.ensure(sizeof(int) * e1);
  ws.in_pos.ensure(sizeof(int) * e1);
  ws.nedges = nedges;
  const unsigned ge = (unsigned)((nedges + 255) / 256);
  int* cnt = ws.cell_cnt.get<int>();
  int* cur = ws.cell_cur.get<int>();
  int* order = ws.cell_atoms.get<int>();
  int32_t* tgt = ws.wrap.get<int32_t>();
  // out lists (owned centres)
  fill_zero(ws, cnt, (size_t)std::max(1, nowned), st);
  fill_zero(ws, cur, (size_t)std::max(1, nowned), st);
  if (nedges > 0) SUS2_LAUNCH(ws, st, "pair_count", k_count_keys, dim3(ge), dim3(256), 0, nedges, ei, cnt);
  device_scan(cnt, ws.nbr_off.get<int>(), nowned, ws.scan_tmp.get<int64_t>(), st);
  if (nedges > 0) {
    SUS2_LAUNCH(ws, st, "pair_place", k_place_keys, dim3(ge), dim3(256), 0, nedges, ei, ws.nbr_off.get<int>(), cur,
                order);
    if (nowned > 0)
      SUS2_LAUNCH(ws, st, "pair_sort", k_sort_lists, dim3(nowned), dim3(LIST_THREADS), 0, nowned, ws.nbr_off.get<int>(),
                  order, (int*)nullptr);
    SUS2_LAUNCH(ws, st, "pair_rec", k_build_rec<T>, dim3(ge), dim3(256), 0, nedges, order, ej, d, types,
                ws.rec.get<Rec<T>>(), tgt);
  }
  // in lists (all atoms) over CSR positions
  fill_zero(ws, cnt, n1, st);
  fill_zero(ws, cur, n1, st);
  if (nedges > 0) SUS2_LAUNCH(ws, st, "pair_count", k_count_keys, dim3(ge), dim3(256), 0, nedges, tgt, cnt);
  device_scan(cnt, in_off, nall, ws.scan_tmp.get<int64_t>() + 2, st);
  if (nedges > 0) {
    SUS2_LAUNCH(ws, st, "pair_place", k_place_keys, dim3(ge), dim3(256), 0, nedges, tgt, in_off, cur,
                ws.in_edges.get<int>());
    if (nall > 0)
      SUS2_LAUNCH(ws, st, "pair_sort", k_sort_lists, dim3(nall), dim3(LIST_THREADS), 0, nall, in_off,
                  ws.in_edges.get<int>(), ws.in_pos.get<int>());
  }
}
template void build_pair_lists<float>(Workspace&, int, int, long, const int32_t*, const int32_t*, const double*,
                                      const int32_t*, int*, hipStream_t);
template void build_pair_lists<double>(Workspace&, int, int, long, const int32_t*, const int32_t*, const double*,
                                       const int32_t*, int*, hipStream_t);

template <class T>
void pair_stage_t(int stage, const DevModelSet& dm, Workspace& ws, const PairView& pv, hipStream_t st) {
  shape_ops(*dm.host).pair[prec_index<T>()](stage, dm, ws, pv, st);
}

void pair_stage_dispatch(int stage, int precision, const DevModelSet& dm, Workspace& ws, const PairView& pv,
                         hipStream_t st) {
  if (precision == SUS2_FP64) pair_stage_t<double>(stage, dm, ws, pv, st);
  else if (precision == SUS2_FP32) pair_stage_t<float>(stage, dm, ws, pv, st);
  else throw Error(SUS2_E_INVALID, "precision must be SUS2_FP64 or SUS2_FP32");
  ws.prof.collect_ready();
}

template <class T>
void run_efs(const DevModelSet& dm, Workspace& ws, const Batch& b, const EfsOut& out, hipStream_t st) {
  shape_ops(*dm.host).efs[prec_index<T>()](dm, ws, b, out, st);
}

void efs_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const EfsOut& out,
                  hipStream_t st) {
  check_rec_limits(precision, b.natoms, dm.host->C);
  if (precision != SUS2_FP64 && precision != SUS2_FP32)
    throw Error(SUS2_E_INVALID, "precision must be SUS2_FP64 or SUS2_FP32");
  // first pass with the edge capacity of earlier calls (no host round trip
  // inside the neighbour build); when the device reports that the lists did
  // not fit, the outputs are recomputed with host-sized buffers
  for (int pass = 0; pass < 2; ++pass) {
    ws.nbr_optimistic = pass == 0;
    try {
      if (precision == SUS2_FP64) run_efs<double>(dm, ws, b, out, st);
      else run_efs<float>(dm, ws, b, out, st);
    } catch (...) {
      ws.nbr_optimistic = ws.nbr_pending = false;
      throw;
    }
    ws.nbr_optimistic = false;
    if (!neighbors_verify(ws, st)) break;
  }
  ws.prof.collect_ready();
}

}  // namespace sus2
