// SUS2-SH FP64 oracle -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's Static / coordination-Effect energy,
// force and virial (EFS) evaluation and of its training loss -> parameter
// gradient.  Only the test-suite, the smoke check and the CPU-baseline leg of
// the benchmark load this library, always as the checker; the product path in
// superlinear_amd/ never links or calls it.
//
// What it follows (reference paths relative to the superlinear repository):
//   model format + validation  interfaces/lammps/ML-SUS2/sus2_sl_model.cpp:174-257
//   coefficient vector layout  dev_src/mtpr.cpp:925-975,1586-1740,2108-2210,2474-2560
//   radial basis + cutoff      interfaces/lammps/ML-SUS2/sus2_sl_radial_math.cpp:256-335,468-503
//   real spherical harmonics   interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp:16-42
//   edge basis / type factors  interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp:166-216,319-338
//   rank contraction           interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp:104-105,217-221
//   product graph              interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp:308-313
//   pass order + readout       interfaces/lammps/ML-SUS2/sus2_sl_kernel.cpp:385-562
//   neighbour images           interfaces/lammps/tests/embedded_efs.cpp:44-66
//   loss and its adjoints      dev_src/non_linear_regression.cpp:892-1065
//   edge weights of the grad   src/basic_mlip.cpp:421-476
//
// Derivatives inside this file are taken by forward-mode automatic
// differentiation over a templated number type (Dual3 for the coordinate
// gradient of an edge basis, HyperDual for the parameter gradient), so they
// are independent of the hand-derived derivative formulas of the HIP kernels.
//
// Pinning: EFS against the reference's FP64 JSON fixtures
// (interfaces/lammps/tests/fixtures/two_cutoff/*), against the reference
// embedded engine compiled into oracle/_ref, against central finite
// differences, and the invariants listed in tests/test_oracle.py.  The
// parameter gradient is pinned by central finite differences of the loss.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- numbers
struct Dual3 {  // value and gradient w.r.t. three inputs
  double v, d[3];
  Dual3(double x = 0) : v(x), d{0, 0, 0} {}
};
inline Dual3 operator+(const Dual3& a, const Dual3& b) { Dual3 r(a.v + b.v); for (int q = 0; q < 3; ++q) r.d[q] = a.d[q] + b.d[q]; return r; }
inline Dual3 operator-(const Dual3& a, const Dual3& b) { Dual3 r(a.v - b.v); for (int q = 0; q < 3; ++q) r.d[q] = a.d[q] - b.d[q]; return r; }
inline Dual3 operator-(const Dual3& a) { Dual3 r(-a.v); for (int q = 0; q < 3; ++q) r.d[q] = -a.d[q]; return r; }
inline Dual3 operator*(const Dual3& a, const Dual3& b) { Dual3 r(a.v * b.v); for (int q = 0; q < 3; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q]; return r; }
inline Dual3 operator/(const Dual3& a, const Dual3& b) { Dual3 r(a.v / b.v); for (int q = 0; q < 3; ++q) r.d[q] = (a.d[q] * b.v - a.v * b.d[q]) / (b.v * b.v); return r; }
inline Dual3 sqrt(const Dual3& a) { Dual3 r(std::sqrt(a.v)); for (int q = 0; q < 3; ++q) r.d[q] = a.d[q] / (2 * r.v); return r; }
inline Dual3 tanh(const Dual3& a) { Dual3 r(std::tanh(a.v)); const double s = 1 - r.v * r.v; for (int q = 0; q < 3; ++q) r.d[q] = s * a.d[q]; return r; }
inline double val(const Dual3& a) { return a.v; }

struct HD {  // a + b e1 + c e2 + d e1 e2 with e1^2 = e2^2 = 0
  double a, b, c, d;
  HD(double x = 0) : a(x), b(0), c(0), d(0) {}
  HD(double a_, double b_, double c_, double d_) : a(a_), b(b_), c(c_), d(d_) {}
};
inline HD operator+(const HD& x, const HD& y) { return HD(x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d); }
inline HD operator-(const HD& x, const HD& y) { return HD(x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d); }
inline HD operator-(const HD& x) { return HD(-x.a, -x.b, -x.c, -x.d); }
inline HD operator*(const HD& x, const HD& y) {
  return HD(x.a * y.a, x.a * y.b + x.b * y.a, x.a * y.c + x.c * y.a,
            x.a * y.d + x.b * y.c + x.c * y.b + x.d * y.a);
}
// f(x) for a scalar function with f, f', f'' at x.a
inline HD chain(const HD& x, double f, double f1, double f2) {
  return HD(f, f1 * x.b, f1 * x.c, f1 * x.d + f2 * x.b * x.c);
}
inline HD operator/(const HD& x, const HD& y) {
  const double i = 1.0 / y.a;
  return x * chain(y, i, -i * i, 2 * i * i * i);
}
inline HD sqrt(const HD& x) { const double s = std::sqrt(x.a); return chain(x, s, 0.5 / s, -0.25 / (s * x.a)); }
inline HD tanh(const HD& x) { const double t = std::tanh(x.a), s = 1 - t * t; return chain(x, t, s, -2 * t * s); }
inline double val(const HD& a) { return a.a; }
inline double val(double a) { return a; }
using std::sqrt;
using std::tanh;

// ---------------------------------------------------------------- model
[[noreturn]] void fail(const std::string& why) { throw std::runtime_error("oracle: " + why); }

struct Product { int a, b, out; double c; };

// Static / Effect subset of the SUS2-SH text model.  Coefficients live in one
// flat vector in the reference optimiser layout (dev_src/mtpr.cpp:925-975,
// 1586-1740, 2108-2210, 2474-2560):
//   shift[C] | scal[2 C^2 Kw] | radial[F (B+C)] | gate radial[F B] |
//   gate scal[2 C^2 Kw] | alpha[C F] | gate type[C F] | gate W[F Qg] |
//   gate rank W[(L+1) R K] | outer rank W[(L+1) O K] | species[C] | readout[Q or C Q]
struct Model {
  std::map<std::string, std::string> field;
  int C = 0, L = 0, H = 0, K = 0, F = 0, B = 0, M = 0, Q = 0, Kw = 0, smode = 3;
  int R = 0, O = 0, Qg = 0, gate_begin = 0, outer_begin = 0;
  int gate_pb = 0, gate_pe = 0, outer_pb = 0;
  double rc = 0, rca = 0, scaling = 1;
  bool gated = false, species_readout = false;
  std::vector<int> bk, bl, bm;  // basic rows (k, l, m)
  std::vector<Product> prod;
  std::vector<int> scalar_map, gate_nodes;
  // coefficient layout
  std::vector<double> theta;
  int o_shift = 0, o_scal = 0, o_rad = 0, o_grad = -1, o_gscal = -1, o_alpha = -1, o_gtype = -1,
      o_W = -1, o_gw = -1, o_ow = -1, o_species = 0, o_read = 0, ncoeff = 0;
  int rad_stride_gate = 0;
  // coefficients whose gradient entries are computed (empty: all); the others
  // are left zero (lets large batches be checked on a sample of entries)
  std::vector<int> grad_subset;
};

static std::string trim(const std::string& s) {
  const auto a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  return s.substr(a, s.find_last_not_of(" \t\r\n") - a + 1);
}
static std::vector<double> parse_list(const std::string& s) {
  std::vector<double> out;
  const char* p = s.c_str();
  while (*p) {
    if (*p == '{' || *p == '}' || *p == ',' || std::isspace((unsigned char)*p)) { ++p; continue; }
    char* end = nullptr;
    const double x = std::strtod(p, &end);
    if (end == p) fail("bad number in list: " + s.substr(0, 40));
    out.push_back(x);
    p = end;
  }
  return out;
}

struct Reader {
  const std::map<std::string, std::string>& f;
  bool has(const char* k) const { return f.count(k) != 0; }
  std::string text(const char* k, const std::string& d = "") const { auto it = f.find(k); return it == f.end() ? d : it->second; }
  double real(const char* k, double d = 0) const { return has(k) ? std::strtod(text(k).c_str(), nullptr) : d; }
  int integer(const char* k, int d = 0) const { return has(k) ? (int)std::lround(real(k)) : d; }
  bool boolean(const char* k) const { const auto s = text(k, "false"); return s == "true" || s == "1"; }
  std::vector<double> list(const char* k) const { return has(k) ? parse_list(text(k)) : std::vector<double>(); }
};

void load_model(const std::string& path, Model& m) {
  std::ifstream in(path);
  if (!in) fail("cannot open " + path);
  std::string line, radial_rows;
  bool header = false, in_radial = false;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    if (!header) { if (line != "MTP") fail("missing MTP header"); header = true; continue; }
    const auto eq = line.find('=');
    if (eq == std::string::npos) {
      if (line == "radial_coeffs") { in_radial = true; continue; }
      if (!in_radial) fail("unexpected line " + line);
      if (line[0] == '{') radial_rows += line + ",";
      continue;  // the species pair label ("0-0") carries no data
    }
    in_radial = false;
    std::string key = trim(line.substr(0, eq)), value = trim(line.substr(eq + 1));
    int depth = 0;
    for (char c : value) depth += (c == '{') - (c == '}');
    while (depth > 0 && std::getline(in, line)) {
      value += " " + trim(line);
      for (char c : line) depth += (c == '{') - (c == '}');
    }
    m.field[key] = value;
  }
  if (!header) fail("empty model");
  Reader r{m.field};
  if (r.text("potential_tag") != "SUS2-SH") fail("not a SUS2-SH model");
  if (r.text("radial_basis_type") != "RBChebyshev_sss") fail("only RBChebyshev_sss is supported");
  if (r.text("radial_cutoff_type", "legacy") != "legacy") fail("only the legacy cutoff is supported");
  if (r.integer("alpha_index_times_count") != 0) fail("Cartesian graphs are unsupported");
  const std::string outer_mode = r.text("outer_radial_mode", "legacy-plus-mixer");
  if (outer_mode != "legacy-plus-mixer") fail("outer_radial_mode " + outer_mode + " is another model family");
  if (r.boolean("zbl_enabled") || r.boolean("joint_readout_enabled") || r.boolean("two_layer_gate_readout_enabled") ||
      r.boolean("two_layer_gate_edge_l1_enabled") || r.boolean("two_layer_residual_enabled") ||
      r.boolean("multi_layer_gate_enabled") || r.text("two_layer_gate_radial_mixing", "none") != "none" ||
      r.integer("tensor_message_lmax") != 0)
    fail("model uses a family outside Static/Effect");
  m.C = r.integer("species_count"); m.L = r.integer("sh_l_max"); m.K = r.integer("sh_k_max");
  m.B = r.integer("radial_basis_size"); m.F = r.integer("radial_funcs_count");
  m.M = r.integer("alpha_moments_count"); m.Q = r.integer("alpha_scalar_moments");
  m.H = (m.L + 1) * (m.L + 1);
  if (m.C <= 0 || m.K <= 0 || m.B <= 0 || m.L < 0 || m.L > 6 || m.F != m.K * (m.L + 1)) fail("inconsistent dims");
  m.rc = r.real("max_dist"); m.rca = r.real("angular_max_dist", m.rc); m.scaling = r.real("scaling", 1.0);
  if (!(m.rca > 0 && m.rca <= m.rc)) fail("bad angular cutoff");
  const std::string sm = r.text("scaling_map", "LK");
  m.smode = sm == "pair" ? 0 : sm == "L" ? 1 : sm == "K" ? 2 : sm == "LK" ? 3 : -1;
  if (m.smode < 0) fail("bad scaling_map");
  m.Kw = m.smode == 0 ? 1 : m.smode == 1 ? m.L + 1 : m.smode == 2 ? m.K : m.F;
  m.species_readout = r.text("readout_mode", "shared") == "species";
  const auto basic = r.list("alpha_index_basic");
  if (basic.size() % 3) fail("basic rows");
  for (size_t i = 0; i < basic.size(); i += 3) {
    m.bk.push_back((int)basic[i]); m.bl.push_back((int)basic[i + 1]); m.bm.push_back((int)basic[i + 2]);
  }
  const auto prods = r.list("sh_products");
  if (prods.size() % 4) fail("product rows");
  for (size_t i = 0; i < prods.size(); i += 4)
    m.prod.push_back({(int)prods[i], (int)prods[i + 1], (int)prods[i + 2], prods[i + 3]});
  for (auto& p : m.prod)
    if (p.a < 0 || p.b < 0 || p.a >= p.out || p.b >= p.out || p.out >= m.M) fail("cyclic or out-of-range product");
  for (double x : r.list("alpha_moment_mapping")) m.scalar_map.push_back((int)x);
  if ((int)m.scalar_map.size() != m.Q) fail("scalar mapping count");
  m.gated = r.boolean("two_layer_gate_enabled");
  const bool ranked = r.boolean("outer_k_rank1_enabled");
  m.O = ranked ? r.integer("outer_k_rank", 1) : 0;
  m.R = r.integer("two_layer_gate_k_rank");
  m.gate_begin = r.integer("two_layer_gate_k_rank_basic_begin");
  m.outer_begin = r.integer("outer_k_rank1_basic_begin");
  m.outer_pb = m.O ? r.integer("outer_k_rank1_product_begin") : 0;
  m.gate_pb = r.integer("two_layer_gate_k_rank_product_begin", 0);
  m.gate_pe = m.R ? r.integer("two_layer_gate_k_rank_product_end", m.outer_pb) : m.outer_pb;
  if (!m.O && !m.R) m.gate_pe = (int)m.prod.size();
  if (m.gated) {
    if (r.text("two_layer_gate_mode") != "mu-scalar-full") fail("only the mu-scalar-full gate is supported");
    if (r.text("two_layer_gate_site_mode", "neighbor") != "neighbor") fail("only neighbour gate sites");
    if (r.has("two_layer_gate_scalar_moment_indices"))
      for (double x : r.list("two_layer_gate_scalar_moment_indices")) m.gate_nodes.push_back((int)x);
    else
      for (double x : r.list("two_layer_gate_scalar_indices")) m.gate_nodes.push_back(m.scalar_map.at((int)x));
    m.Qg = (int)m.gate_nodes.size();
  }
  // ---- flat coefficient vector
  const int C = m.C, F = m.F, B = m.B, L1 = m.L + 1;
  auto& th = m.theta;
  auto put = [&](int& off, const std::vector<double>& v, size_t n, const char* what) {
    if (v.size() != n) fail(std::string("wrong coefficient count in ") + what);
    off = (int)th.size();
    th.insert(th.end(), v.begin(), v.end());
  };
  std::vector<double> v = r.list("shift_coeffs");
  if (v.empty()) v.assign(C, -1.0);
  put(m.o_shift, v, C, "shift_coeffs");
  v = r.list("scal_coeffs");
  if (v.empty()) {  // default warps, sus2_sl_model.cpp:217-226
    const int P = C * C;
    v.assign(2 * P * m.Kw, 0);
    for (int b = 0; b < m.Kw; ++b) for (int t = 0; t < C; ++t) for (int u = 0; u < C; ++u) {
      const double x = C == 1 ? .5 : (std::sqrt(t + 1.) + std::sqrt(u + 1.) - 2.) / (2. * std::sqrt(double(C)) - 2.);
      v[2 * b * P + t * C + u] = 4. + (2. - 4.) * x;
      v[(2 * b + 1) * P + t * C + u] = 1.5 + x;
    }
  }
  put(m.o_scal, v, 2 * C * C * m.Kw, "scal_coeffs");
  v = parse_list(radial_rows);
  if (v.empty()) {
    v.assign(F * (B + C), .01);
    for (int mu = 0; mu < F; ++mu) for (int t = 0; t < C; ++t) v[mu * (B + C) + B + t] = 2.;
  }
  put(m.o_rad, v, F * (B + C), "radial_coeffs");
  if (m.gated) {
    const std::string rm = r.text("two_layer_gate_radial_mode", "base-radial");
    v = r.list("two_layer_gate_radial_coeffs");
    if (rm == "shared-radial") { put(m.o_grad, v, F * B, "two_layer_gate_radial_coeffs"); m.rad_stride_gate = B; }
    else { if (!v.empty()) fail("gate radial coefficients without shared-radial"); m.rad_stride_gate = B + C; }
    if (r.text("two_layer_gate_scaling_mode", "shared-scaling") == "decoupled")
      put(m.o_gscal, r.list("two_layer_gate_scaling_coeffs"), 2 * C * C * m.Kw, "two_layer_gate_scaling_coeffs");
    v = r.list("two_layer_gate_additive_coeffs");
    if (v.empty()) v.assign(C * F, 1.0);  // sus2_sl_kernel.cpp:334 default
    put(m.o_alpha, v, C * F, "two_layer_gate_additive_coeffs");
    v = r.list("two_layer_gate_type_coeffs");
    if (v.empty()) {  // sus2_sl_kernel.cpp:56: take the outer type columns
      v.assign(C * F, 0);
      for (int t = 0; t < C; ++t) for (int mu = 0; mu < F; ++mu) v[t * F + mu] = th[m.o_rad + mu * (B + C) + B + t];
    }
    put(m.o_gtype, v, C * F, "two_layer_gate_type_coeffs");
    put(m.o_W, r.list("two_layer_gate_weights"), F * m.Qg, "two_layer_gate_weights");
  }
  if (m.R > 0) {
    v = r.list("two_layer_gate_k_rank_weights");
    if (v.empty()) v.assign(L1 * m.R * m.K, 1.0);
    put(m.o_gw, v, L1 * m.R * m.K, "two_layer_gate_k_rank_weights");
  }
  if (m.O > 0) {
    v = r.list("outer_k_rank1_weights");
    if (v.empty()) v.assign(L1 * m.O * m.K, 1.0);
    put(m.o_ow, v, L1 * m.O * m.K, "outer_k_rank1_weights");
  }
  v = r.list("species_coeffs");
  std::vector<double> w = r.list("moment_coeffs");
  if (v.empty()) { v.assign(C, .001); w.assign(m.Q * (m.species_readout ? C : 1), m.species_readout ? 1e-6 : .001); }
  put(m.o_species, v, C, "species_coeffs");
  put(m.o_read, w, m.Q * (m.species_readout ? C : 1), "moment_coeffs");
  m.ncoeff = (int)th.size();
}

// ---------------------------------------------------------------- math
inline int hidx(int l, int m) { return l * l + l + m; }

// Real spherical harmonics Y_lm(d/|d|), l <= L, Condon-Shortley phase
// cancelled (Y_1{-1,0,1} ~ (y, z, x)), flat index l^2+l+m
// (sus2_sl_kernel.cpp:16-42, dev_src/mtpr_sh.cpp:923-1002).  Evaluated from
// the unit vector u: P_l^m(u_z)/sin^m via the three-term recurrence in l and
// (u_x + i u_y)^m for the azimuthal factor.
template <class T>
void real_sh(const T& x, const T& y, const T& z, int L, T* Y) {
  const T r = sqrt(x * x + y * y + z * z);
  const T ux = x / r, uy = y / r, uz = z / r;
  T cm[16], sm[16];  // Re/Im of (ux + i uy)^m (L <= 15)
  cm[0] = T(1.0); sm[0] = T(0.0);
  for (int m = 1; m <= L; ++m) {
    cm[m] = ux * cm[m - 1] - uy * sm[m - 1];
    sm[m] = ux * sm[m - 1] + uy * cm[m - 1];
  }
  const double pi = std::acos(-1.0);
  for (int m = 0; m <= L; ++m) {
    T p2(0.0), p1(0.0), p(0.0);
    double dfact = 1;  // (2m-1)!!
    for (int a = 1; a <= 2 * m - 1; a += 2) dfact *= a;
    for (int l = m; l <= L; ++l) {
      if (l == m) p = T(dfact);
      else if (l == m + 1) p = T(double(2 * m + 1)) * uz * p1;
      else p = (T(double(2 * l - 1)) * uz * p1 - T(double(l + m - 1)) * p2) / T(double(l - m));
      double ratio = 1;
      for (int a = l - m + 1; a <= l + m; ++a) ratio /= a;
      const double n = std::sqrt((2 * l + 1) * ratio / (4 * pi));
      if (m == 0) Y[hidx(l, 0)] = T(n) * p;
      else {
        const double s = std::sqrt(2.0) * n;
        Y[hidx(l, m)] = T(s) * p * cm[m];
        Y[hidx(l, -m)] = T(s) * p * sm[m];
      }
      p2 = p1; p1 = p;
    }
  }
}

// phi_b(r) = scaling * T_b(tanh(slope (r - shift) / 2)) * (r - rc)^2, zero for
// r >= rc (RBChebyshev_sss with the legacy cutoff; sus2_sl_radial_math.cpp:468-503,
// 256-265; the two-cutoff envelope of :296-335).
template <class T>
void cheb_basis(const T& r, const T& slope, const T& shift, double rc, double scaling, int B, T* phi) {
  if (val(r) >= rc) { for (int b = 0; b < B; ++b) phi[b] = T(0.0); return; }
  const T xi = tanh(slope * (r - shift) * T(0.5));
  const T f = (r - T(rc)) * (r - T(rc));
  phi[0] = T(scaling) * f;
  if (B > 1) phi[1] = T(scaling) * (xi * f);
  for (int b = 2; b < B; ++b) phi[b] = T(2.0) * xi * phi[b - 1] - phi[b - 2];
}

// ---------------------------------------------------------------- neighbours
struct Edge { int i, j; double d[3]; };

// Full directed neighbour list with every periodic image inside rc
// (interfaces/lammps/tests/embedded_efs.cpp:44-66 semantics; strict r < rc).
// A lattice row of zeros marks a non-periodic direction, as in the
// reference's Configuration (src/neighborhoods.cpp:216-220): images are only
// taken along the non-zero rows.
std::vector<Edge> neighbours(int n, const double* x, const double* cell, bool periodic, double rc) {
  std::vector<Edge> out;
  double inv[3][3] = {{0}};
  double reach[3] = {0, 0, 0};
  bool act[3] = {false, false, false};
  if (periodic)
    for (int q = 0; q < 3; ++q) act[q] = cell[3 * q] != 0.0 || cell[3 * q + 1] != 0.0 || cell[3 * q + 2] != 0.0;
  const bool any = act[0] || act[1] || act[2];
  if (any) {
    // basis: active lattice rows, completed by unit normals for the open axes
    double a[9];
    for (int k = 0; k < 9; ++k) a[k] = cell[k];
    auto cross = [](const double* u, const double* v, double* w) {
      w[0] = u[1] * v[2] - u[2] * v[1]; w[1] = u[2] * v[0] - u[0] * v[2]; w[2] = u[0] * v[1] - u[1] * v[0];
    };
    auto unit = [](double* w) {
      const double s = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      for (int k = 0; k < 3; ++k) w[k] /= s;
    };
    const int nact = act[0] + act[1] + act[2];
    if (nact == 1) {
      const int p = act[0] ? 0 : act[1] ? 1 : 2;
      const double* v = a + 3 * p;
      double e[3] = {0, 0, 0};
      const int m = std::fabs(v[0]) <= std::fabs(v[1]) && std::fabs(v[0]) <= std::fabs(v[2]) ? 0
                    : std::fabs(v[1]) <= std::fabs(v[2]) ? 1 : 2;
      e[m] = 1;
      const int q1 = (p + 1) % 3, q2 = (p + 2) % 3;
      cross(v, e, a + 3 * q1); unit(a + 3 * q1);
      cross(v, a + 3 * q1, a + 3 * q2); unit(a + 3 * q2);
    } else if (nact == 2) {
      const int q = !act[0] ? 0 : !act[1] ? 1 : 2;
      cross(a + 3 * ((q + 1) % 3), a + 3 * ((q + 2) % 3), a + 3 * q); unit(a + 3 * q);
    }
    double b[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
      b[j][i] = a[((i + 1) % 3) * 3 + (j + 1) % 3] * a[((i + 2) % 3) * 3 + (j + 2) % 3] -
                a[((i + 1) % 3) * 3 + (j + 2) % 3] * a[((i + 2) % 3) * 3 + (j + 1) % 3];
    double det = 0;
    for (int i = 0; i < 3; ++i) det += a[i] * b[i][0];
    if (!(std::fabs(det) > 0)) fail("singular cell");
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) inv[i][j] = b[i][j] / det;
    for (int q = 0; q < 3; ++q) {
      double s = 0;
      for (int a2 = 0; a2 < 3; ++a2) s += inv[a2][q] * inv[a2][q];
      reach[q] = rc * std::sqrt(s);
    }
  }
  const double rc2 = rc * rc;
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {
    double d[3], f[3] = {0, 0, 0};
    for (int q = 0; q < 3; ++q) d[q] = x[3 * j + q] - x[3 * i + q];
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (any) {
      for (int q = 0; q < 3; ++q) for (int a2 = 0; a2 < 3; ++a2) f[q] += d[a2] * inv[a2][q];
      for (int q = 0; q < 3; ++q)
        if (act[q]) { lo[q] = (int)std::ceil(-f[q] - reach[q]); hi[q] = (int)std::floor(-f[q] + reach[q]); }
    }
    for (int a2 = lo[0]; a2 <= hi[0]; ++a2) for (int b2 = lo[1]; b2 <= hi[1]; ++b2) for (int c2 = lo[2]; c2 <= hi[2]; ++c2) {
      if (i == j && a2 == 0 && b2 == 0 && c2 == 0) continue;
      Edge e{i, j, {0, 0, 0}};
      double r2 = 0;
      for (int q = 0; q < 3; ++q) {
        e.d[q] = d[q] + a2 * cell[q] + b2 * cell[3 + q] + c2 * cell[6 + q];
        r2 += e.d[q] * e.d[q];
      }
      if (r2 < rc2) out.push_back(e);
    }
  }
  return out;
}

// ---------------------------------------------------------------- passes
template <class T>
struct Coeff {  // typed view of the flat coefficient vector
  const Model& m;
  const T* th;
  T slope(bool gate, int blk, int ti, int tj) const {
    const int o = (gate && m.o_gscal >= 0) ? m.o_gscal : m.o_scal;
    return th[o + 2 * blk * m.C * m.C + ti * m.C + tj];
  }
  T shift(bool gate, int blk, int ti, int tj) const {
    const int o = (gate && m.o_gscal >= 0) ? m.o_gscal : m.o_scal;
    return th[o + (2 * blk + 1) * m.C * m.C + ti * m.C + tj];
  }
  T rad(bool gate, int mu, int b) const {
    if (gate && m.o_grad >= 0) return th[m.o_grad + mu * m.B + b];
    return th[m.o_rad + mu * (m.B + m.C) + b];
  }
  T tau(int t, int mu) const { return th[m.o_rad + mu * (m.B + m.C) + m.B + t]; }
  T alpha(int t, int mu) const { return th[m.o_alpha + t * m.F + mu]; }
  T gtype(int t, int mu) const { return th[m.o_gtype + t * m.F + mu]; }
  T W(int mu, int q) const { return th[m.o_W + mu * m.Qg + q]; }
  T rankw(bool gate, int l, int a, int k) const {
    const int rank = gate ? m.R : m.O, o = gate ? m.o_gw : m.o_ow;
    return th[o + (l * rank + a) * m.K + k];
  }
  T readout(int t, int q) const {  // includes the shared-readout species factor
    if (m.species_readout) return th[m.o_read + t * m.Q + q];
    return th[m.o_species + t] * th[m.o_read + q];
  }
  T site0(int t) const { return th[m.o_species + t] + th[m.o_shift + t]; }
  int block(int k, int l) const {
    return m.smode == 2 ? k : m.smode == 1 ? l : m.smode == 0 ? 0 : k * (m.L + 1) + l;
  }
};

// Radial functions R_mu(r) for one directed edge of pass `gate`; active l
// range and per-l cutoff as in sus2_sl_kernel.cpp:173-207.
template <class T>
int edge_radial(const Coeff<T>& c, bool gate, int ti, int tj, const T& r, T* R) {
  const Model& m = c.m;
  const int L1 = m.L + 1;
  const int activeL = val(r) >= m.rca ? 1 : L1;
  std::vector<T> phi(m.B);
  for (int mu = 0; mu < m.F; ++mu) R[mu] = T(0.0);
  for (int k = 0; k < m.K; ++k) for (int l = 0; l < activeL; ++l) {
    const int blk = c.block(k, l);
    cheb_basis(r, c.slope(gate, blk, ti, tj), c.shift(gate, blk, ti, tj), l == 0 ? m.rc : m.rca, m.scaling, m.B, phi.data());
    T a(0.0);
    for (int b = 0; b < m.B; ++b) a = a + c.rad(gate, k * L1 + l, b) * phi[b];
    R[k * L1 + l] = a;
  }
  return activeL;
}

template <class T>
void contract(const Model& m, const Coeff<T>& c, bool gate, T* v) {
  const int rank = gate ? m.R : m.O, begin = gate ? m.gate_begin : m.outer_begin;
  for (int a = 0; a < rank; ++a) for (size_t b = 0; b < m.bk.size(); ++b) {
    const int k = m.bk[b], l = m.bl[b], mm = m.bm[b];
    const int dst = begin + rank * (m.H - (l + 1) * (l + 1)) + (rank - 1 - a) * (2 * l + 1) + mm + l;
    v[dst] = v[dst] + c.rankw(gate, l, a, k) * v[b];
  }
}
template <class T>
void graph(const Model& m, bool gate, T* v) {
  const int lo = gate ? m.gate_pb : m.outer_pb, hi = gate ? m.gate_pe : (int)m.prod.size();
  for (int q = lo; q < hi; ++q) { const auto& p = m.prod[q]; v[p.out] = v[p.out] + T(p.c) * v[p.a] * v[p.b]; }
}

// Basic moments of one pass for every atom.  Edge contribution to basic row b
// = x0 * x1 * R_mu(r) * Y_lm(d) with x0, x1 the centre/neighbour type
// factors (sus2_sl_kernel.cpp:319-353).
template <class T, class E>
void moments(const Model& m, const Coeff<T>& c, bool gate, int n, const int* type, const std::vector<E>& edges,
             const T* h, std::vector<T>& mom) {
  mom.assign((size_t)n * m.M, T(0.0));
  std::vector<T> R(m.F), Y(m.H);
  for (const auto& e : edges) {
    const T dx = e.d[0], dy = e.d[1], dz = e.d[2];
    const T r = sqrt(dx * dx + dy * dy + dz * dz);
    if (val(r) >= m.rc) continue;
    const int ti = type[e.i], tj = type[e.j];
    const int activeL = edge_radial(c, gate, ti, tj, r, R.data());
    real_sh(dx, dy, dz, activeL - 1, Y.data());
    T* row = mom.data() + (size_t)e.i * m.M;
    for (size_t b = 0; b < m.bk.size(); ++b) {
      const int l = m.bl[b];
      if (l >= activeL) continue;
      const int mu = m.bk[b] * (m.L + 1) + l;
      T x0, x1;
      if (gate) { x0 = c.gtype(ti, mu); x1 = c.gtype(tj, mu); }
      else {
        x0 = c.tau(ti, mu);
        x1 = c.tau(tj, mu);
        if (m.gated) x1 = x1 + c.alpha(tj, mu) * h[(size_t)e.j * m.F + mu];
      }
      row[b] = row[b] + x0 * x1 * (R[mu] * Y[hidx(l, m.bm[b])]);
    }
  }
}

// Forward energy of one frame in number type T (gate pass -> h -> outer pass
// -> site energies), edges given with typed vectors so that coordinate
// tangents can ride along.
template <class T>
struct TEdge { int i, j; T d[3]; };

template <class T>
void forward(const Model& m, const T* theta, int n, const int* type, const std::vector<TEdge<T>>& edges,
             std::vector<T>& site, std::vector<T>* h_out = nullptr, std::vector<T>* gmom_out = nullptr,
             std::vector<T>* omom_out = nullptr) {
  Coeff<T> c{m, theta};
  std::vector<T> h, gm, om;
  if (m.gated) {
    moments(m, c, true, n, type, edges, (const T*)nullptr, gm);
    h.assign((size_t)n * m.F, T(0.0));
    for (int i = 0; i < n; ++i) {
      T* a = gm.data() + (size_t)i * m.M;
      contract(m, c, true, a);
      graph(m, true, a);
      for (int mu = 0; mu < m.F; ++mu) {
        T s(0.0);
        for (int q = 0; q < m.Qg; ++q) s = s + c.W(mu, q) * a[m.gate_nodes[q]];
        h[(size_t)i * m.F + mu] = s;
      }
    }
  }
  moments(m, c, false, n, type, edges, h.data(), om);
  site.assign(n, T(0.0));
  for (int i = 0; i < n; ++i) {
    T* a = om.data() + (size_t)i * m.M;
    contract(m, c, false, a);
    graph(m, false, a);
    T E = c.site0(type[i]);
    for (int q = 0; q < m.Q; ++q) E = E + c.readout(type[i], q) * a[m.scalar_map[q]];
    site[i] = E;
  }
  if (h_out) *h_out = h;
  if (gmom_out) *gmom_out = gm;
  if (omom_out) *omom_out = om;
}

// ---------------------------------------------------------------- EFS (FP64)
struct EFS { double energy; std::vector<double> site, force; double virial[6]; };

// Reverse of one pass's node adjoints back to the basic rows (graph reverse in
// reverse row order, then the transposed rank contraction;
// sus2_sl_kernel.cpp:308-313, 217-221).
void reverse_nodes(const Model& m, const Coeff<double>& c, bool gate, const double* v, double* a) {
  const int lo = gate ? m.gate_pb : m.outer_pb, hi = gate ? m.gate_pe : (int)m.prod.size();
  for (int q = hi - 1; q >= lo; --q) {
    const auto& p = m.prod[q];
    const double z = p.c * a[p.out];
    a[p.a] += z * v[p.b];
    a[p.b] += z * v[p.a];
  }
  const int rank = gate ? m.R : m.O, begin = gate ? m.gate_begin : m.outer_begin;
  for (int aa = 0; aa < rank; ++aa) for (size_t b = 0; b < m.bk.size(); ++b) {
    const int k = m.bk[b], l = m.bl[b], mm = m.bm[b];
    const int dst = begin + rank * (m.H - (l + 1) * (l + 1)) + (rank - 1 - aa) * (2 * l + 1) + mm + l;
    a[b] += c.rankw(gate, l, aa, k) * a[dst];
  }
}

// Coordinate reverse of one pass: g_e = sum_b adj_b d(x0 x1 R Y)/dd over the
// pass's edges; optional gate adjoint h_bar_j += adj_b x0 alpha_j R Y
// (sus2_sl_kernel.cpp:356-380, 545-546).
void coordinate_reverse(const Model& m, const Coeff<double>& c, bool gate, const int* type, const std::vector<Edge>& edges,
                        const double* h, const std::vector<double>& adj, double* force, double* virial, double* hbar) {
  const int L1 = m.L + 1;
  std::vector<Dual3> R(m.F), Y(m.H);
  Coeff<Dual3> cd{m, nullptr};
  (void)cd;
  for (const auto& e : edges) {
    Dual3 dx(e.d[0]), dy(e.d[1]), dz(e.d[2]);
    dx.d[0] = 1; dy.d[1] = 1; dz.d[2] = 1;
    const Dual3 r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r.v >= m.rc) continue;
    const int ti = type[e.i], tj = type[e.j];
    // radial functions in Dual3 (coefficients are constants here)
    const int activeL = r.v >= m.rca ? 1 : L1;
    std::vector<Dual3> phi(m.B);
    for (int mu = 0; mu < m.F; ++mu) R[mu] = Dual3(0.0);
    for (int k = 0; k < m.K; ++k) for (int l = 0; l < activeL; ++l) {
      const int blk = c.block(k, l);
      cheb_basis(r, Dual3(c.slope(gate, blk, ti, tj)), Dual3(c.shift(gate, blk, ti, tj)), l == 0 ? m.rc : m.rca, m.scaling, m.B, phi.data());
      Dual3 a(0.0);
      for (int b = 0; b < m.B; ++b) a = a + Dual3(c.rad(gate, k * L1 + l, b)) * phi[b];
      R[k * L1 + l] = a;
    }
    real_sh(dx, dy, dz, activeL - 1, Y.data());
    const double* arow = adj.data() + (size_t)e.i * m.M;
    double g[3] = {0, 0, 0};
    for (size_t b = 0; b < m.bk.size(); ++b) {
      const int l = m.bl[b];
      if (l >= activeL) continue;
      const int mu = m.bk[b] * L1 + l;
      const Dual3 basis = R[mu] * Y[hidx(l, m.bm[b])];
      double x0, x1;
      if (gate) { x0 = c.gtype(ti, mu); x1 = c.gtype(tj, mu); }
      else {
        x0 = c.tau(ti, mu);
        x1 = c.tau(tj, mu);
        if (m.gated) {
          x1 += c.alpha(tj, mu) * h[(size_t)e.j * m.F + mu];
          hbar[(size_t)e.j * m.F + mu] += arow[b] * x0 * c.alpha(tj, mu) * basis.v;
        }
      }
      const double w = arow[b] * x0 * x1;
      for (int q = 0; q < 3; ++q) g[q] += w * basis.d[q];
    }
    for (int q = 0; q < 3; ++q) { force[3 * e.i + q] += g[q]; force[3 * e.j + q] -= g[q]; }
    virial[0] -= e.d[0] * g[0]; virial[1] -= e.d[1] * g[1]; virial[2] -= e.d[2] * g[2];
    virial[3] -= e.d[0] * g[1]; virial[4] -= e.d[0] * g[2]; virial[5] -= e.d[1] * g[2];
  }
}

EFS efs(const Model& m, const double* theta, int n, const int* type, const std::vector<Edge>& edges) {
  for (int i = 0; i < n; ++i) if (type[i] < 0 || type[i] >= m.C) fail("species index out of range");
  std::vector<TEdge<double>> te(edges.size());
  for (size_t k = 0; k < edges.size(); ++k) {
    te[k].i = edges[k].i; te[k].j = edges[k].j;
    for (int q = 0; q < 3; ++q) te[k].d[q] = edges[k].d[q];
  }
  EFS out;
  std::vector<double> h, gm, om;
  forward<double>(m, theta, n, type, te, out.site, &h, &gm, &om);
  out.energy = 0;
  for (int i = 0; i < n; ++i) out.energy += out.site[i];
  out.force.assign(3 * (size_t)n, 0.0);
  for (double& w : out.virial) w = 0;
  Coeff<double> c{m, theta};
  // outer adjoints
  std::vector<double> adj((size_t)n * m.M, 0.0);
  for (int i = 0; i < n; ++i) {
    double* a = adj.data() + (size_t)i * m.M;
    for (int q = 0; q < m.Q; ++q) a[m.scalar_map[q]] += c.readout(type[i], q);
    reverse_nodes(m, c, false, om.data() + (size_t)i * m.M, a);
  }
  std::vector<double> hbar(m.gated ? (size_t)n * m.F : 0, 0.0);
  coordinate_reverse(m, c, false, type, edges, h.data(), adj, out.force.data(), out.virial, hbar.data());
  if (m.gated) {
    std::vector<double> gadj((size_t)n * m.M, 0.0);
    for (int i = 0; i < n; ++i) {
      double* a = gadj.data() + (size_t)i * m.M;
      for (int mu = 0; mu < m.F; ++mu)
        for (int q = 0; q < m.Qg; ++q) a[m.gate_nodes[q]] += hbar[(size_t)i * m.F + mu] * c.W(mu, q);
      reverse_nodes(m, c, true, gm.data() + (size_t)i * m.M, a);
    }
    coordinate_reverse(m, c, true, type, edges, nullptr, gadj, out.force.data(), out.virial, nullptr);
  }
  return out;
}

// ---------------------------------------------------------------- loss
// Objective options (NonLinearRegression members, dev_src/non_linear_regression.h:110-140):
// opts[7] = rel_forces, norm_by_forces, force_loss (0 l2, 1 log-cosh),
// log_cosh_scale, std_weight, stdd_weight, constr_weight.  Null = the
// objective of `mlp-sus2 train` for these models (dev_src/mlp/mtpr_train.cpp:
// 1827-1849, 2077-2090, 2832: constr 1e-9, log-cosh scale 2, the rest 0).
struct Objective {
  double rel = 0; int nbf = 0; int kind = 0; double lcs = 2.0; double stdw = 0, stddw = 0, constr = 1e-9;
};
Objective objective_of(const double* o) {
  Objective ob;
  if (!o) return ob;
  ob.rel = o[0]; ob.nbf = (int)o[1]; ob.kind = (int)o[2]; ob.lcs = o[3]; ob.stdw = o[4]; ob.stddw = o[5];
  ob.constr = o[6];
  if (ob.kind != 0 && ob.kind != 1) fail("force loss kind must be 0 (l2) or 1 (log-cosh)");
  if (!(ob.lcs > 0)) fail("log-cosh scale must be positive");
  return ob;
}

// Force residual loss and derivative (dev_src/force_loss.h:47-70): L2 and the
// log-cosh residual s^2 log cosh(d / s), here by its series-free identity
// log cosh x = log((1 + e^{-2|x|}) / 2) + |x|.
double rho(double d, const Objective& o) {
  if (o.kind == 0) return d * d;
  const double x = std::fabs(d) / o.lcs;
  return o.lcs * o.lcs * (x + std::log((1.0 + std::exp(-2.0 * x)) / 2.0));
}
double drho(double d, const Objective& o) { return o.kind == 0 ? 2.0 * d : o.lcs * std::tanh(d / o.lcs); }

// Per-frame loss of dev_src/non_linear_regression.cpp:936-1052; `mode` 0 =
// vibrations, 1 = structures, 2 = molecules.  Returns the frame loss and
// writes the adjoints handed to the model (dL/dE_i, dL/dF_i, str_weights =
// dLdS of the reference, i.e. -2 w_S (S - S0) / N).  The std terms need the
// site energies `se` and species `type` of the frame's atoms.
double frame_loss(int n, int mode, double wE, double wF, double wS, double E, const double* F, const double* S,
                  double E0, const double* F0, const double* S0, bool hasE, bool hasF, bool hasS,
                  double* dEi, double* dF, double* dS, const Objective& ob = Objective{},
                  const double* se = nullptr, const int* type = nullptr) {
  double loss = 0;
  for (int i = 0; i < 3 * n; ++i) dF[i] = 0;
  for (int q = 0; q < 9; ++q) dS[q] = 0;
  auto f0sq = [&](int i) { return F0[3 * i] * F0[3 * i] + F0[3 * i + 1] * F0[3 * i + 1] + F0[3 * i + 2] * F0[3 * i + 2]; };
  double avef = 0;
  if (hasF)
    for (int i = 0; i < n; ++i) avef += f0sq(i) / n;
  const double nb = 0.1 / (0.1 + ob.nbf * avef);
  if (wF != 0 && hasF) {
    for (int i = 0; i < n; ++i) {
      double w = ob.rel <= 0 ? wF : wF * ob.rel / (f0sq(i) + ob.rel);
      if (mode == 1) w /= n;
      w *= nb;
      for (int a = 0; a < 3; ++a) {
        const double d = F[3 * i + a] - F0[3 * i + a];
        loss += w * rho(d, ob);
        dF[3 * i + a] = w * drho(d, ob);
      }
    }
  }
  double we = wE, ws = wS;
  if (mode == 1) { we /= n; ws /= n; }
  else if (mode == 2) { we *= n; ws *= n; }
  if (ws != 0 && hasS) {
    double fro = 0;
    for (int q = 0; q < 9; ++q) { const double d = S[q] - S0[q]; fro += d * d; dS[q] = -2.0 * ws * d / n; }
    loss += ws * fro / n;
  }
  double dE = 0;
  if (we != 0 && hasE) { loss += we * (E - E0) * (E - E0) * nb / n; dE = 2.0 * we * (E - E0) * nb / n; }
  for (int i = 0; i < n; ++i) dEi[i] = dE;
  if ((ob.stdw != 0 || ob.stddw != 0) && n > 0) {
    if (!se || !type) fail("the std terms need the frame's site energies and species");
    // per-species count, mean site energy, and the joint term of the adjoint
    std::map<int, double> cnt, mean, joint;
    for (int i = 0; i < n; ++i) cnt[type[i]] += 1;
    for (int i = 0; i < n; ++i) mean[type[i]] += se[i] / cnt[type[i]];
    const double mean0 = E0 / n;
    double sd = 0, sdd = 0;
    for (int i = 0; i < n; ++i) {
      const double g = 200.0 / (200.0 + f0sq(i)), d = se[i] - mean[type[i]];
      joint[type[i]] += d * g / cnt[type[i]] / n;
      sd += d * d * g / n;
    }
    for (const auto& kv : mean) sdd += (kv.second - mean0) * (kv.second - mean0);
    loss += ob.stdw * sd + ob.stddw * sdd;
    for (int i = 0; i < n; ++i) {
      const int t = type[i];
      dEi[i] += ob.stdw * 2.0 * ((se[i] - mean[t]) * 200.0 / (200.0 + f0sq(i)) / n - joint[t]) +
                ob.stddw * 2.0 * (mean[t] - mean0) / cnt[t];
    }
  }
  return loss;
}

// Legacy radial penalty (MLMTPR::AddPenaltyGrad, dev_src/mtpr.cpp:14050):
// rows c_mu = theta[o_rad + mu (B + C) + b], b < B, mu < F;
// constr (sum_mu (|c_mu|^2 - 1)^2 + sum_{mu<nu} (c_mu . c_nu)^2), gradient
// added to grad (may be null).
double radial_penalty(const Model& m, const double* theta, double constr, double* grad) {
  if (constr == 0) return 0.0;
  const int F = m.F, B = m.B, st = m.B + m.C, o = m.o_rad;
  auto c = [&](int k, int l) { return theta[o + k * st + l]; };
  double v = 0;
  for (int k = 0; k < F; ++k) {
    double nn = 0;
    for (int l = 0; l < B; ++l) nn += c(k, l) * c(k, l);
    v += constr * (nn - 1) * (nn - 1);
    if (grad) for (int l = 0; l < B; ++l) grad[o + k * st + l] += 4 * constr * (nn - 1) * c(k, l);
  }
  for (int k1 = 0; k1 < F; ++k1)
    for (int k2 = k1 + 1; k2 < F; ++k2) {
      double sc = 0;
      for (int l = 0; l < B; ++l) sc += c(k1, l) * c(k2, l);
      v += constr * sc * sc;
      if (grad)
        for (int l = 0; l < B; ++l) {
          grad[o + k1 * st + l] += 2 * constr * sc * c(k2, l);
          grad[o + k2 * st + l] += 2 * constr * sc * c(k1, l);
        }
    }
  return v;
}

// Parameter gradient of one frame:  sum_i dEi_i dE_i/dtheta + d/dtheta of
// the directional derivative of E along w_e = dF_i - dF_j + dS d_e for every
// directed edge e = (i -> j)  (src/basic_mlip.cpp:421-446).  Hyper-dual
// numbers carry the coefficient direction in e1 and w in e2, one coefficient
// at a time.  out[ncoeff] is overwritten.
void frame_grad(const Model& m, int n, const int* type, const std::vector<Edge>& e, const double* dEi,
                const double* dF, const double* dS, double* out, int nthreads) {
  const int P = m.ncoeff;
  std::vector<TEdge<HD>> te(e.size());
  for (size_t k = 0; k < e.size(); ++k) {
    const int i = e[k].i, j = e[k].j;
    te[k].i = i; te[k].j = j;
    for (int q = 0; q < 3; ++q) {
      double w = dF[3 * i + q] - dF[3 * j + q];
      for (int b = 0; b < 3; ++b) w += dS[3 * q + b] * e[k].d[b];
      te[k].d[q] = HD(e[k].d[q], 0, w, 0);
    }
  }
  // coefficients are independent forward passes: spread them over threads
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<HD> th(P);
    for (int p = 0; p < P; ++p) th[p] = HD(m.theta[p]);
    std::vector<HD> site;
    const bool sub = !m.grad_subset.empty();
    const int np = sub ? (int)m.grad_subset.size() : P;
#pragma omp for schedule(dynamic, 4)
    for (int p = 0; p < P; ++p) out[p] = 0.0;
#pragma omp for schedule(dynamic, 4)
    for (int k = 0; k < np; ++k) {
      const int p = sub ? m.grad_subset[k] : k;
      th[p].b = 1.0;
      forward<HD>(m, th.data(), n, type, te, site);
      double g = 0;
      for (int i = 0; i < n; ++i) g += dEi[i] * site[i].b + site[i].d;
      out[p] = g;
      th[p].b = 0.0;
    }
  }
}

}  // namespace orc

// ================================================================= C ABI
using namespace orc;
namespace {
thread_local std::string g_err;
int guard(const std::function<void()>& f) {
  try { f(); return 0; } catch (const std::exception& e) { g_err = e.what(); return 1; }
}
void voigt_to_full(const double* w, double* S) {  // plus_stress matrix from xx,yy,zz,xy,xz,yz
  S[0] = w[0]; S[1] = w[3]; S[2] = w[4]; S[3] = w[3]; S[4] = w[1]; S[5] = w[5]; S[6] = w[4]; S[7] = w[5]; S[8] = w[2];
}
}

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_model_load(const char* path) {
  Model* m = new Model();
  if (guard([&] { load_model(path, *m); })) { delete m; return nullptr; }
  return m;
}
void orc_model_free(void* h) { delete static_cast<Model*>(h); }
int orc_coeff_count(void* h) { return static_cast<Model*>(h)->ncoeff; }
void orc_coeff_get(void* h, double* out) { auto* m = static_cast<Model*>(h); std::copy(m->theta.begin(), m->theta.end(), out); }
void orc_coeff_set(void* h, const double* in) { auto* m = static_cast<Model*>(h); std::copy(in, in + m->ncoeff, m->theta.begin()); }
double orc_cutoff(void* h) { return static_cast<Model*>(h)->rc; }
// dims: C L K B F M Q R O Qg Kw nbasic nprod
void orc_dims(void* h, int* d) {
  auto* m = static_cast<Model*>(h);
  const int v[] = {m->C, m->L, m->K, m->B, m->F, m->M, m->Q, m->R, m->O, m->Qg, m->Kw, (int)m->bk.size(), (int)m->prod.size()};
  std::copy(v, v + 13, d);
}
// offsets: shift scal rad grad gscal alpha gtype W gw ow species read ncoeff
void orc_offsets(void* h, int* d) {
  auto* m = static_cast<Model*>(h);
  const int v[] = {m->o_shift, m->o_scal, m->o_rad, m->o_grad, m->o_gscal, m->o_alpha, m->o_gtype, m->o_W, m->o_gw, m->o_ow, m->o_species, m->o_read, m->ncoeff};
  std::copy(v, v + 13, d);
}

// Real Y_lm and its Cartesian gradient for one vector (test hook).
void orc_sh(int L, const double* d, double* Y, double* dY) {
  if (L < 0 || L > 15) return;
  Dual3 x(d[0]), y(d[1]), z(d[2]);
  x.d[0] = 1; y.d[1] = 1; z.d[2] = 1;
  std::vector<Dual3> v((L + 1) * (L + 1));
  real_sh(x, y, z, L, v.data());
  for (int h = 0; h < (L + 1) * (L + 1); ++h) { Y[h] = v[h].v; for (int q = 0; q < 3; ++q) dY[3 * h + q] = v[h].d[q]; }
}

// Neighbour count / fill (for tests): returns edge count, writes up to cap edges.
long orc_neighbors(int n, const double* x, const double* cell, int periodic, double rc, long cap, int* ij, double* d) {
  std::vector<Edge> e;
  if (guard([&] { e = neighbours(n, x, cell, periodic != 0, rc); })) return -1;
  for (long k = 0; k < (long)e.size() && k < cap; ++k) {
    ij[2 * k] = e[k].i; ij[2 * k + 1] = e[k].j;
    for (int q = 0; q < 3; ++q) d[3 * k + q] = e[k].d[q];
  }
  return (long)e.size();
}

// EFS of a batch of frames.  Frames are independent; with nthreads > 1 they
// run on OpenMP threads (the DP analogue of the reference's per-rank frame
// shards).  virial is 6 per frame (xx yy zz xy xz yz, reference PlusStress).
int orc_efs_batch(void* h, int nframes, const long* off, const int* type, const double* x, const double* cell,
                  const int* periodic, double* energy, double* force, double* virial, double* site, int nthreads) {
  auto* m = static_cast<Model*>(h);
  int bad = 0;
  std::string err;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int f = 0; f < nframes; ++f) {
    try {
      const long a = off[f];
      const int n = (int)(off[f + 1] - a);
      auto e = neighbours(n, x + 3 * a, cell + 9 * f, periodic[f] != 0, m->rc);
      EFS r = efs(*m, m->theta.data(), n, type + a, e);
      energy[f] = r.energy;
      for (int q = 0; q < 6; ++q) virial[6 * f + q] = r.virial[q];
      for (long i = 0; i < 3L * n; ++i) force[3 * a + i] = r.force[i];
      if (site) for (int i = 0; i < n; ++i) site[a + i] = r.site[i];
    } catch (const std::exception& ex) {
#pragma omp critical
      { bad = 1; err = ex.what(); }
    }
  }
  if (bad) { g_err = err; return 1; }
  return 0;
}

// EFS on caller-supplied directed edges of one domain that owns every atom
// (the pair_style compute() edge layout with no ghosts: i, j in [0, n),
// d = x_j - x_i).
int orc_efs_edges(void* h, int n, int nall, const int* type, long nedge, const int* ij, const double* d,
                  double* energy, double* force, double* virial, double* site) {
  auto* m = static_cast<Model*>(h);
  return guard([&] {
    if (nall != n) fail("the oracle edge entry evaluates single-domain edge lists only");
    std::vector<Edge> e(nedge);
    for (long k = 0; k < nedge; ++k) {
      e[k].i = ij[2 * k]; e[k].j = ij[2 * k + 1];
      if (e[k].i < 0 || e[k].i >= n || e[k].j < 0 || e[k].j >= nall) fail("edge index out of range");
      for (int q = 0; q < 3; ++q) e[k].d[q] = d[3 * k + q];
    }
    EFS r = efs(*m, m->theta.data(), nall, type, e);
    double E = 0;
    for (int i = 0; i < n; ++i) { E += r.site[i]; if (site) site[i] = r.site[i]; }
    *energy = E;
    for (long i = 0; i < 3L * nall; ++i) force[i] = r.force[i];
    for (int q = 0; q < 6; ++q) virial[q] = r.virial[q];
  });
}

// Training objective and its parameter gradient over a batch of labelled
// frames: objective = (1/nframes) sum_f frame_loss, gradient by hyper-dual
// forward differentiation of  sum_i Ebar_i E_i + sum_e w_e . dE/dd_e  with
// w_e = dLdF_i - dLdF_j + str_weights d_e (src/basic_mlip.cpp:421-446), one
// coefficient at a time.  Slow by design: it is the checker.
int orc_loss_grad(void* h, int nframes, const long* off, const int* type, const double* x, const double* cell,
                  const int* periodic, const double* E0, const double* F0, const double* S0, int mode, double wE,
                  double wF, double wS, const double* opts, double* loss_out, double* grad, double* ebar_out,
                  double* fbar_out, double* sbar_out, int nthreads) {
  auto* m = static_cast<Model*>(h);
  return guard([&] {
    const Objective ob = objective_of(opts);
    const int P = m->ncoeff;
    std::vector<double> G((size_t)nframes * P, 0.0);
    std::vector<double> Lf(nframes, 0.0);
    int bad = 0;
    std::string err;
    // frames in order; frame_grad spreads the coefficients over the threads
    for (int f = 0; f < nframes; ++f) {
      try {
        const long a = off[f];
        const int n = (int)(off[f + 1] - a);
        auto e = neighbours(n, x + 3 * a, cell + 9 * f, periodic[f] != 0, m->rc);
        EFS r = efs(*m, m->theta.data(), n, type + a, e);
        double S[9], S0f[9];
        voigt_to_full(r.virial, S);
        voigt_to_full(S0 + 6 * f, S0f);
        std::vector<double> dEi(n), dF(3 * n);
        double dS[9];
        Lf[f] = frame_loss(n, mode, wE, wF, wS, r.energy, r.force.data(), S, E0[f], F0 + 3 * a, S0f, true, true, true,
                           dEi.data(), dF.data(), dS, ob, r.site.data(), type + a);
        if (ebar_out) for (int i = 0; i < n; ++i) ebar_out[a + i] = dEi[i];
        if (fbar_out) for (int i = 0; i < 3 * n; ++i) fbar_out[3 * a + i] = dF[i];
        if (sbar_out) for (int q = 0; q < 9; ++q) sbar_out[9 * f + q] = dS[q];
        frame_grad(*m, n, type + a, e, dEi.data(), dF.data(), dS, G.data() + (size_t)f * P, nthreads);
      } catch (const std::exception& ex) {
#pragma omp critical
        { bad = 1; err = ex.what(); }
      }
    }
    if (bad) fail(err);
    double L = 0;
    for (int f = 0; f < nframes; ++f) L += Lf[f];
    *loss_out = L / nframes;
    for (int p = 0; p < P; ++p) {
      double s = 0;
      for (int f = 0; f < nframes; ++f) s += G[(size_t)f * P + p];
      grad[p] = s / nframes;
    }
    // the penalty is added once per frame and divided by the frame count
    // with the rest (dev_src/non_linear_regression.cpp:1052, mtpr_trainer.cpp:4737)
    if (nframes > 0) *loss_out += radial_penalty(*m, m->theta.data(), ob.constr, grad);
  });
}

// Parameter gradient with caller-given weights (AccumulateEFSCombinationGrad
// over a batch, dev_src/mtpr.cpp:9188): grad[p] = sum over frames of
// frame_grad with ene_w [natoms], frc_w [natoms*3], str_w [nframes*9].
int orc_efs_grad(void* h, int nframes, const long* off, const int* type, const double* x, const double* cell,
                 const int* periodic, const double* ene_w, const double* frc_w, const double* str_w, double* grad,
                 int nthreads) {
  auto* m = static_cast<Model*>(h);
  return guard([&] {
    const int P = m->ncoeff;
    std::vector<double> G((size_t)nframes * P, 0.0);
    int bad = 0;
    std::string err;
    // frames in order; frame_grad spreads the coefficients over the threads
    for (int f = 0; f < nframes; ++f) {
      try {
        const long a = off[f];
        const int n = (int)(off[f + 1] - a);
        auto e = neighbours(n, x + 3 * a, cell + 9 * f, periodic[f] != 0, m->rc);
        frame_grad(*m, n, type + a, e, ene_w + a, frc_w + 3 * a, str_w + 9 * f, G.data() + (size_t)f * P,
                   nthreads);
      } catch (const std::exception& ex) {
#pragma omp critical
        { bad = 1; err = ex.what(); }
      }
    }
    if (bad) fail(err);
    for (int p = 0; p < P; ++p) {
      double s = 0;
      for (int f = 0; f < nframes; ++f) s += G[(size_t)f * P + p];
      grad[p] = s;
    }
  });
}

// Plain loss value (for finite differences of the gradient).
int orc_loss(void* h, int nframes, const long* off, const int* type, const double* x, const double* cell,
             const int* periodic, const double* E0, const double* F0, const double* S0, int mode, double wE,
             double wF, double wS, const double* opts, double* loss_out) {
  auto* m = static_cast<Model*>(h);
  return guard([&] {
    const Objective ob = objective_of(opts);
    double L = 0;
    for (int f = 0; f < nframes; ++f) {
      const long a = off[f];
      const int n = (int)(off[f + 1] - a);
      auto e = neighbours(n, x + 3 * a, cell + 9 * f, periodic[f] != 0, m->rc);
      EFS r = efs(*m, m->theta.data(), n, type + a, e);
      double S[9], S0f[9], dS[9];
      voigt_to_full(r.virial, S);
      voigt_to_full(S0 + 6 * f, S0f);
      std::vector<double> dEi(n), dF(3 * n);
      L += frame_loss(n, mode, wE, wF, wS, r.energy, r.force.data(), S, E0[f], F0 + 3 * a, S0f, true, true, true,
                      dEi.data(), dF.data(), dS, ob, r.site.data(), type + a);
    }
    *loss_out = L / nframes + (nframes > 0 ? radial_penalty(*m, m->theta.data(), ob.constr, nullptr) : 0.0);
  });
}

// Restrict the gradient entries of orc_loss_grad / orc_efs_grad to the n
// listed coefficients (n = 0: all); the other entries are returned as zero.
int orc_set_grad_subset(void* h, int n, const int* idx) {
  auto* m = static_cast<Model*>(h);
  return guard([&] {
    std::vector<int> v(idx, idx + n);
    for (int p : v)
      if (p < 0 || p >= m->ncoeff) fail("gradient subset index out of range");
    m->grad_subset = v;
  });
}

// One frame's loss and adjoints from given predictions and labels (no model):
// E, E0 scalars, F, F0 [n*3], S, S0 Voigt xx yy zz xy xz yz, objective options
// opts[7] (null = defaults), site energies se [n] and species type [n] (used
// by the std terms only; may be null otherwise).  Outputs dEi [n], dF [n*3],
// dS [9] row-major; returns the frame loss (NaN on an invalid option)
// (dev_src/non_linear_regression.cpp:936-1052).
double orc_frame_loss(int n, int mode, double wE, double wF, double wS, double E, const double* F, const double* S,
                      double E0, const double* F0, const double* S0, const double* opts, const double* se,
                      const int* type, double* dEi, double* dF, double* dS) {
  double Sf[9], S0f[9];
  voigt_to_full(S, Sf);
  voigt_to_full(S0, S0f);
  double L = 0;
  if (guard([&] {
        L = frame_loss(n, mode, wE, wF, wS, E, F, Sf, E0, F0, S0f, true, true, true, dEi, dF, dS, objective_of(opts),
                       se, type);
      }) != 0)
    return std::nan("");
  return L;
}

// The legacy radial penalty of the model's coefficients theta (ncoeff) and
// its gradient (grad [ncoeff] overwritten; may be null).
double orc_radial_penalty(void* h, const double* theta, double constr, double* grad) {
  auto* m = static_cast<Model*>(h);
  if (grad) std::fill(grad, grad + m->ncoeff, 0.0);
  return radial_penalty(*m, theta, constr, grad);
}

}  // extern "C"
