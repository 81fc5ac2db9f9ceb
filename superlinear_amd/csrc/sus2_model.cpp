// SUS2-SH text model reader/writer and execution-plan builder.
//
// Format semantics follow interfaces/lammps/ML-SUS2/sus2_sl_model.cpp:174-257
// (reader, defaults) and :258-453 (validation) for the Static and
// coordination-Effect families; the coefficient vector follows the reference
// optimiser layout of dev_src/mtpr.cpp:925-975, 1586-1740, 2108-2210,
// 2474-2560.  Other model families (adaptive, mix-basis, joint readout, ZBL,
// residual/multilayer gates, body-linear-combo gates) are rejected with
// SUS2_E_UNSUPPORTED.
#include "sus2_model.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <set>
#include <sstream>

#include "../../include/sus2amd.h"

namespace sus2 {
namespace {

std::string trim(const std::string& s) {
  const auto a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  return s.substr(a, s.find_last_not_of(" \t\r\n") - a + 1);
}
[[noreturn]] void bad(const std::string& w) { throw Error(SUS2_E_INVALID, "SUS2 model: " + w); }
[[noreturn]] void unsupported(const std::string& w) { throw Error(SUS2_E_UNSUPPORTED, "SUS2 model: " + w); }

std::vector<double> numbers(const std::string& s, const std::string& key) {
  std::vector<double> out;
  const char* p = s.c_str();
  int depth = 0;
  while (*p) {
    const char c = *p;
    if (c == '{') { ++depth; ++p; continue; }
    if (c == '}') { --depth; ++p; continue; }
    if (c == ',' || std::isspace((unsigned char)c)) { ++p; continue; }
    char* end = nullptr;
    const double x = std::strtod(p, &end);
    if (end == p || !std::isfinite(x)) bad("invalid number in " + key);
    out.push_back(x);
    p = end;
  }
  if (depth != 0) bad("unbalanced braces in " + key);
  return out;
}

struct Fields {
  const std::map<std::string, std::string>& f;
  bool has(const std::string& k) const { return f.count(k) != 0; }
  std::string text(const std::string& k, const std::string& d = "") const {
    auto it = f.find(k);
    return it == f.end() ? d : it->second;
  }
  double real(const std::string& k, double d = 0) const {
    if (!has(k)) return d;
    char* end = nullptr;
    const std::string s = text(k);
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || !trim(end).empty() || !std::isfinite(v)) bad("invalid number " + k);
    return v;
  }
  int integer(const std::string& k, int d = 0) const {
    if (!has(k)) return d;
    const double v = real(k);
    if (std::floor(v) != v || std::fabs(v) > 2e9) bad("non-integer " + k);
    return (int)v;
  }
  bool boolean(const std::string& k) const {
    if (!has(k)) return false;
    const auto s = text(k);
    if (s != "true" && s != "false" && s != "0" && s != "1") bad("invalid boolean " + k);
    return s == "true" || s == "1";
  }
  std::vector<double> list(const std::string& k) const { return has(k) ? numbers(text(k), k) : std::vector<double>(); }
};

// Layered schedule of one pass: forward layers by output dependency depth,
// reverse layers by consumer depth; both in fixed (serialised) row order so
// every accumulation is deterministic.  A depth level with more than
// kLayerRows rows is cut into consecutive sub-layers at node boundaries (the
// nodes of one level are independent, so every node's sum is unchanged):
// the kernels stage one (sub-)layer's row products in LDS, and the cap keeps
// that scratch -- and so the LDS footprint per atom -- bounded at large L.
constexpr int kLayerRows = 1024;

// Rows per sub-layer when a level of `total` rows is cut: equal shares of the
// fewest sub-layers under the cap.
int split_target(int total) {
  const int parts = std::max(1, (total + kLayerRows - 1) / kLayerRows);
  return (total + parts - 1) / parts;
}

PassPlan make_plan(const Model& m, bool gate) {
  PassPlan p;
  int lo = m.M, hi = -1;
  auto touch = [&](int n) { lo = std::min(lo, n); hi = std::max(hi, n); };
  p.prod_lo = gate ? m.gate_pb : m.outer_pb;
  p.prod_hi = gate ? m.gate_pe : (int)m.prod.size();
  p.rank = gate ? m.R : m.O;
  p.begin = gate ? m.gate_begin : m.outer_begin;
  std::vector<int> layer(m.M, 0);
  std::vector<char> is_out(m.M, 0);
  for (int q = p.prod_lo; q < p.prod_hi; ++q) {
    const auto& r = m.prod[q];
    layer[r.out] = std::max(layer[r.out], 1 + std::max(layer[r.a], layer[r.b]));
    is_out[r.out] = 1;
    touch(r.a); touch(r.b); touch(r.out);
  }
  if (p.rank > 0) { touch(p.begin); touch(p.begin + p.rank * m.H - 1); }
  for (int n : gate ? m.gate_nodes : m.scalar_map) touch(n);
  if (hi < lo) { lo = 0; hi = 0; }
  p.node_lo = lo;
  p.node_cnt = hi - lo + 1;
  int nl = 0;
  for (int n = 0; n < m.M; ++n) if (is_out[n]) nl = std::max(nl, layer[n]);
  std::vector<std::vector<int>> rows_of(m.M);
  for (int q = p.prod_lo; q < p.prod_hi; ++q) rows_of[m.prod[q].out].push_back(q);
  p.fw_layer_off.push_back(0);
  p.fw_row_off.push_back(0);
  for (int l = 1; l <= nl; ++l) {
    int total = 0, rows = 0;
    for (int n = 0; n < m.M; ++n) if (is_out[n] && layer[n] == l) total += (int)rows_of[n].size();
    const int target = split_target(total);
    for (int n = 0; n < m.M; ++n) if (is_out[n] && layer[n] == l) {
      const int nr = (int)rows_of[n].size();
      if (rows > 0 && (rows + nr > kLayerRows || rows >= target)) {
        p.fw_layer_off.push_back((int)p.fw_out.size());
        rows = 0;
      }
      rows += nr;
      p.fw_out.push_back(n);
      for (int q : rows_of[n]) { p.fw_a.push_back(m.prod[q].a); p.fw_b.push_back(m.prod[q].b); p.fw_c.push_back(m.prod[q].c); }
      p.fw_row_off.push_back((int)p.fw_a.size());
    }
    p.fw_layer_off.push_back((int)p.fw_out.size());
  }
  // reverse: consumers of each node
  std::vector<std::vector<std::array<int, 2>>> cons(m.M);  // (row, role)
  for (int q = p.prod_lo; q < p.prod_hi; ++q) { cons[m.prod[q].a].push_back({q, 0}); cons[m.prod[q].b].push_back({q, 1}); }
  std::vector<int> rl(m.M, 0);
  // nodes in decreasing id order: a consumer's out id is > the input id
  for (int n = m.M - 1; n >= 0; --n)
    for (const auto& c : cons[n]) rl[n] = std::max(rl[n], 1 + rl[m.prod[c[0]].out]);
  int nr = 0;
  for (int n = 0; n < m.M; ++n) if (!cons[n].empty()) nr = std::max(nr, rl[n]);
  p.rv_layer_off.push_back(0);
  p.rv_row_off.push_back(0);
  for (int l = 1; l <= nr; ++l) {
    int total = 0, rows = 0;
    for (int n = 0; n < m.M; ++n) if (!cons[n].empty() && rl[n] == l) total += (int)cons[n].size();
    const int target = split_target(total);
    for (int n = 0; n < m.M; ++n) if (!cons[n].empty() && rl[n] == l) {
      const int nc = (int)cons[n].size();
      if (rows > 0 && (rows + nc > kLayerRows || rows >= target)) {
        p.rv_layer_off.push_back((int)p.rv_node.size());
        rows = 0;
      }
      rows += nc;
      p.rv_node.push_back(n);
      for (const auto& c : cons[n]) {
        const auto& r = m.prod[c[0]];
        p.rv_out.push_back(r.out);
        p.rv_partner.push_back(c[1] == 0 ? r.b : r.a);
        p.rv_c.push_back(r.c);
      }
      p.rv_row_off.push_back((int)p.rv_out.size());
    }
    p.rv_layer_off.push_back((int)p.rv_node.size());
  }
  for (size_t l = 0; l + 1 < p.fw_layer_off.size(); ++l)
    p.max_fw_rows = std::max(p.max_fw_rows, p.fw_row_off[p.fw_layer_off[l + 1]] - p.fw_row_off[p.fw_layer_off[l]]);
  for (size_t l = 0; l + 1 < p.rv_layer_off.size(); ++l)
    p.max_rv_rows = std::max(p.max_rv_rows, p.rv_row_off[p.rv_layer_off[l + 1]] - p.rv_row_off[p.rv_layer_off[l]]);
  return p;
}

std::string fmt_list(const std::vector<double>& v) {
  std::ostringstream o;
  o << std::scientific << std::setprecision(16) << "{";  // 17 significant digits: exact round trip
  for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
  o << "}";
  return o.str();
}

}  // namespace

void Model::load(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error(SUS2_E_IO, "SUS2 model: cannot open " + path);
  std::string line;
  bool header = false, in_radial = false;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    if (!header) { if (line != "MTP") bad("missing MTP header"); header = true; continue; }
    const auto eq = line.find('=');
    if (eq == std::string::npos) {
      if (line == "radial_coeffs") {
        if (in_radial || !radial_rows.empty()) bad("duplicate radial_coeffs");
        in_radial = true;
        key_order.push_back("radial_coeffs");
        continue;
      }
      if (!in_radial) bad("unrecognized line: " + line);
      if (line[0] == '{') radial_rows.push_back(numbers(line, "radial_coeffs"));
      else if (line.find('-') == std::string::npos) bad("invalid radial coefficient block");
      continue;
    }
    in_radial = false;
    const std::string key = trim(line.substr(0, eq));
    std::string value = trim(line.substr(eq + 1));
    if (key.empty() || field.count(key)) bad("duplicate or empty field " + key);
    int depth = 0;
    for (char c : value) depth += (c == '{') - (c == '}');
    while (depth > 0 && std::getline(in, line)) {
      value += " " + trim(line);
      for (char c : line) depth += (c == '{') - (c == '}');
    }
    if (depth != 0) bad("unbalanced braces in " + key);
    field[key] = value;
    key_order.push_back(key);
  }
  if (!header) bad("empty model");
  Fields f{field};
  if (f.text("potential_tag") != "SUS2-SH") bad("requires a SUS2-SH spherical harmonic model");
  if (f.text("radial_basis_type") != "RBChebyshev_sss") unsupported("radial basis " + f.text("radial_basis_type"));
  if (f.text("radial_cutoff_type", "legacy") != "legacy") unsupported("radial cutoff type");
  if (f.integer("alpha_index_times_count") != 0) unsupported("Cartesian product graph");
  if (f.text("outer_radial_mode", "legacy-plus-mixer") != "legacy-plus-mixer") unsupported("outer radial mode " + f.text("outer_radial_mode"));
  for (const char* k : {"zbl_enabled", "joint_readout_enabled", "two_layer_gate_readout_enabled", "two_layer_gate_edge_l1_enabled",
                        "two_layer_residual_enabled", "multi_layer_gate_enabled"})
    if (f.boolean(k)) unsupported(std::string("feature ") + k);
  if (f.text("two_layer_gate_radial_mixing", "none") != "none") unsupported("radial mixing");
  if (f.integer("tensor_message_lmax") != 0) unsupported("tensor messages");
  C = f.integer("species_count"); L = f.integer("sh_l_max"); K = f.integer("sh_k_max");
  B = f.integer("radial_basis_size"); F = f.integer("radial_funcs_count");
  M = f.integer("alpha_moments_count"); Q = f.integer("alpha_scalar_moments");
  if (C <= 0 || C > 64 || K <= 0 || K > 16 || B <= 0 || B > 32 || L < 0 || L > 6 || F != K * (L + 1) || M <= 0 || Q < 0)
    bad("inconsistent L/K/B/species dimensions");
  if (f.integer("L", L) != L) bad("L and sh_l_max differ");
  H = (L + 1) * (L + 1);
  rc = f.real("max_dist"); rca = f.real("angular_max_dist", rc); scaling = f.real("scaling", 1.0);
  if (!(rc > 0) || !(rca > 0) || rca > rc || !(scaling > 0)) bad("invalid distance or scaling");
  const std::string sm = f.text("scaling_map", "LK");
  smode = sm == "pair" ? 0 : sm == "L" ? 1 : sm == "K" ? 2 : sm == "LK" ? 3 : -1;
  if (smode < 0) bad("unknown scaling_map");
  Kw = smode == 0 ? 1 : smode == 1 ? L + 1 : smode == 2 ? K : F;
  const std::string ro = f.text("readout_mode", "shared");
  if (ro != "shared" && ro != "species") bad("invalid readout_mode");
  species_readout = ro == "species";
  const auto basic = f.list("alpha_index_basic");
  if (basic.size() % 3 || (int)basic.size() / 3 != f.integer("alpha_index_basic_count")) bad("invalid basic graph dimensions");
  basic_node.assign(K * H, -1);
  for (size_t i = 0; i < basic.size(); i += 3) {
    const int k = (int)basic[i], l = (int)basic[i + 1], m = (int)basic[i + 2];
    if (k < 0 || k >= K || l < 0 || l > L || m < -l || m > l) bad("invalid basic (k,l,m)");
    bk.push_back(k); bl.push_back(l); bm.push_back(m);
    const int c = (l * l + l + m) * K + k;  // canonical moment order: h-major, k-minor
    if (basic_node[c] >= 0) bad("duplicate basic row");
    basic_node[c] = (int)i / 3;
  }
  if ((int)bk.size() > M) bad("more basic rows than moments");
  const auto pr = f.list("sh_products");
  if (pr.size() % 4 || (int)pr.size() / 4 != f.integer("sh_product_count")) bad("wrong sh_product_count");
  for (size_t i = 0; i < pr.size(); i += 4) {
    Product p{(int)pr[i], (int)pr[i + 1], (int)pr[i + 2], pr[i + 3]};
    if (p.a < 0 || p.b < 0 || p.a >= p.out || p.b >= p.out || p.out >= M) bad("invalid or cyclic product graph");
    prod.push_back(p);
  }
  for (double x : f.list("alpha_moment_mapping")) {
    const int n = (int)x;
    if (n < 0 || n >= M) bad("scalar mapping out of range");
    scalar_map.push_back(n);
  }
  if ((int)scalar_map.size() != Q) bad("wrong scalar mapping count");
  gated = f.boolean("two_layer_gate_enabled");
  const bool ranked = f.boolean("outer_k_rank1_enabled");
  O = ranked ? f.integer("outer_k_rank", 1) : 0;
  R = f.integer("two_layer_gate_k_rank");
  gate_begin = f.integer("two_layer_gate_k_rank_basic_begin");
  outer_begin = f.integer("outer_k_rank1_basic_begin");
  outer_pb = O ? f.integer("outer_k_rank1_product_begin") : 0;
  gate_pb = f.integer("two_layer_gate_k_rank_product_begin", 0);
  gate_pe = R ? f.integer("two_layer_gate_k_rank_product_end", outer_pb) : outer_pb;
  if (!O && !R) gate_pe = (int)prod.size();
  if (O < 0 || R < 0 || (O && (outer_begin < (int)bk.size() || outer_begin + O * H > M)) ||
      (R && (gate_begin < (int)bk.size() || gate_begin + R * H > M)) || outer_pb < 0 || outer_pb > (int)prod.size() ||
      gate_pb < 0 || gate_pe < gate_pb || gate_pe > (int)prod.size())
    bad("invalid rank graph segment");
  if (gated) {
    if (f.text("two_layer_gate_mode") != "mu-scalar-full") unsupported("gate mode " + f.text("two_layer_gate_mode"));
    if (f.text("two_layer_gate_site_mode", "neighbor") != "neighbor") unsupported("gate site mode");
    if (f.has("two_layer_gate_scalar_moment_indices")) {
      for (double x : f.list("two_layer_gate_scalar_moment_indices")) gate_nodes.push_back((int)x);
    } else {
      for (double x : f.list("two_layer_gate_scalar_indices")) {
        const int q = (int)x;
        if (q < 0 || q >= Q) bad("gate scalar index out of range");
        gate_nodes.push_back(scalar_map[q]);
      }
    }
    for (int n : gate_nodes) if (n < 0 || n >= M) bad("gate moment out of range");
    Qg = (int)gate_nodes.size();
    if (Qg <= 0) bad("empty gate scalar basis");
  }
  // ---- coefficient vector
  auto put = [&](int& off, std::vector<double> v, size_t n, const char* what) {
    if (v.size() != n) bad(std::string("wrong coefficient count in ") + what);
    off = (int)theta.size();
    theta.insert(theta.end(), v.begin(), v.end());
  };
  std::vector<double> v = f.list("shift_coeffs");
  if (v.empty()) v.assign(C, -1.0);
  put(o_shift, v, C, "shift_coeffs");
  v = f.list("scal_coeffs");
  if (v.empty()) {
    const int P = C * C;
    v.assign(2 * P * Kw, 0);
    for (int b = 0; b < Kw; ++b) for (int t = 0; t < C; ++t) for (int u = 0; u < C; ++u) {
      const double x = C == 1 ? .5 : (std::sqrt(t + 1.) + std::sqrt(u + 1.) - 2.) / (2. * std::sqrt(double(C)) - 2.);
      v[2 * b * P + t * C + u] = 4. - 2. * x;
      v[(2 * b + 1) * P + t * C + u] = 1.5 + x;
    }
  }
  put(o_scal, v, 2 * C * C * Kw, "scal_coeffs");
  v.clear();
  for (const auto& r : radial_rows) v.insert(v.end(), r.begin(), r.end());
  if (v.empty()) {
    v.assign(F * (B + C), .01);
    for (int mu = 0; mu < F; ++mu) for (int t = 0; t < C; ++t) v[mu * (B + C) + B + t] = 2.;
  }
  put(o_rad, v, F * (B + C), "radial_coeffs");
  if (gated) {
    const std::string rm = f.text("two_layer_gate_radial_mode", "base-radial");
    v = f.list("two_layer_gate_radial_coeffs");
    if (rm == "shared-radial") put(o_grad, v, F * B, "two_layer_gate_radial_coeffs");
    else if (!v.empty()) bad("gate radial coefficients require shared-radial");
    const std::string gs = f.text("two_layer_gate_scaling_mode", "shared-scaling");
    if (gs == "decoupled") put(o_gscal, f.list("two_layer_gate_scaling_coeffs"), 2 * C * C * Kw, "two_layer_gate_scaling_coeffs");
    v = f.list("two_layer_gate_additive_coeffs");
    if (v.empty()) v.assign(C * F, 1.0);
    put(o_alpha, v, C * F, "two_layer_gate_additive_coeffs");
    v = f.list("two_layer_gate_type_coeffs");
    if (v.empty()) {
      v.assign(C * F, 0);
      for (int t = 0; t < C; ++t) for (int mu = 0; mu < F; ++mu) v[t * F + mu] = theta[o_rad + mu * (B + C) + B + t];
    }
    put(o_gtype, v, C * F, "two_layer_gate_type_coeffs");
    put(o_W, f.list("two_layer_gate_weights"), F * Qg, "two_layer_gate_weights");
  }
  if (R > 0) {
    v = f.list("two_layer_gate_k_rank_weights");
    if (v.empty()) v.assign((L + 1) * R * K, 1.0);
    put(o_gw, v, (L + 1) * R * K, "two_layer_gate_k_rank_weights");
  }
  if (O > 0) {
    v = f.list("outer_k_rank1_weights");
    if (v.empty()) v.assign((L + 1) * O * K, 1.0);
    put(o_ow, v, (L + 1) * O * K, "outer_k_rank1_weights");
  }
  v = f.list("species_coeffs");
  std::vector<double> w = f.list("moment_coeffs");
  if (v.empty()) { v.assign(C, .001); w.assign(Q * (species_readout ? C : 1), species_readout ? 1e-6 : .001); }
  put(o_species, v, C, "species_coeffs");
  put(o_read, w, Q * (species_readout ? C : 1), "moment_coeffs");
  ncoeff = (int)theta.size();
  if (gated) gate_plan = make_plan(*this, true);
  outer_plan = make_plan(*this, false);
}

// Writes the model back with the current coefficients in every coefficient
// field (the reference's Save is dev_src/mtpr.cpp:6893); structural fields
// are written as read.
void Model::save(const std::string& path) const {
  std::ofstream out(path);
  if (!out) throw Error(SUS2_E_IO, "SUS2 model: cannot write " + path);
  auto slice = [&](int off, int n) { return std::vector<double>(theta.begin() + off, theta.begin() + off + n); };
  std::map<std::string, std::string> fld = field;
  fld["shift_coeffs"] = fmt_list(slice(o_shift, C));
  fld["scal_coeffs"] = fmt_list(slice(o_scal, 2 * C * C * Kw));
  if (o_grad >= 0) fld["two_layer_gate_radial_coeffs"] = fmt_list(slice(o_grad, F * B));
  if (o_gscal >= 0) fld["two_layer_gate_scaling_coeffs"] = fmt_list(slice(o_gscal, 2 * C * C * Kw));
  if (o_alpha >= 0) { fld["two_layer_gate_additive_coeffs"] = fmt_list(slice(o_alpha, C * F)); fld["two_layer_gate_additive_coeff_count"] = std::to_string(C * F); }
  if (o_gtype >= 0) { fld["two_layer_gate_type_coeffs"] = fmt_list(slice(o_gtype, C * F)); fld["two_layer_gate_type_coeff_count"] = std::to_string(C * F); }
  if (o_W >= 0) fld["two_layer_gate_weights"] = fmt_list(slice(o_W, F * Qg));
  if (o_gw >= 0) { fld["two_layer_gate_k_rank_weights"] = fmt_list(slice(o_gw, (L + 1) * R * K)); fld["two_layer_gate_k_rank_weight_count"] = std::to_string((L + 1) * R * K); }
  if (o_ow >= 0) { fld["outer_k_rank1_weights"] = fmt_list(slice(o_ow, (L + 1) * O * K)); fld["outer_k_rank1_weight_count"] = std::to_string((L + 1) * O * K); }
  fld["species_coeffs"] = fmt_list(slice(o_species, C));
  fld["moment_coeffs"] = fmt_list(slice(o_read, Q * (species_readout ? C : 1)));
  std::vector<std::string> order = key_order;
  for (const auto& kv : fld) if (std::find(order.begin(), order.end(), kv.first) == order.end()) order.push_back(kv.first);
  static const std::set<std::string> indented = {"min_dist", "max_dist", "angular_max_dist", "radial_basis_size", "radial_funcs_count"};
  out << "MTP\n";
  for (const auto& k : order) {
    if (k == "radial_coeffs") {
      out << "\tradial_coeffs\n\t\t0-0\n";
      for (int mu = 0; mu < F; ++mu) out << "\t\t\t" << fmt_list(slice(o_rad + mu * (B + C), B + C)) << "\n";
      continue;
    }
    out << (indented.count(k) ? "\t" : "") << k << " = " << fld.at(k) << "\n";
  }
  if (std::find(order.begin(), order.end(), "radial_coeffs") == order.end()) {
    out << "\tradial_coeffs\n\t\t0-0\n";
    for (int mu = 0; mu < F; ++mu) out << "\t\t\t" << fmt_list(slice(o_rad + mu * (B + C), B + C)) << "\n";
  }
}

}  // namespace sus2
