// Host-side SUS2-SH model: text-format reader/writer, the flat coefficient
// vector in the reference optimiser layout, and the per-pass execution plan
// (layered product-graph schedules) that the HIP kernels walk.
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace sus2 {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Product { int a, b, out; double c; };

// One pass (gate or outer) of the product graph as a layered schedule.
//  forward:  out nodes grouped by layer; each out sums rows (a, b, c)
//  reverse:  input nodes grouped by reverse layer; each gathers rows
//            (out, partner, c) from its consumers
struct PassPlan {
  int prod_lo = 0, prod_hi = 0;
  std::vector<int> fw_layer_off, fw_out, fw_row_off, fw_a, fw_b;
  std::vector<double> fw_c;
  std::vector<int> rv_layer_off, rv_node, rv_row_off, rv_out, rv_partner;
  std::vector<double> rv_c;
  // contraction: rank, first rank node
  int rank = 0, begin = 0;
  // contiguous node range [node_lo, node_lo + node_cnt) touched by the pass
  // (rank nodes, product inputs/outputs and readout seeds)
  int node_lo = 0, node_cnt = 0;
  // largest number of product rows in one forward layer / consumer entries in
  // one reverse layer (LDS scratch of the row-parallel walks)
  int max_fw_rows = 0, max_rv_rows = 0;
};

struct Model {
  std::map<std::string, std::string> field;
  std::vector<std::string> key_order;  // for a faithful re-save
  std::vector<std::vector<double>> radial_rows;
  // dims
  int C = 0, L = 0, H = 0, K = 0, F = 0, B = 0, M = 0, Q = 0, Kw = 0, smode = 3;
  int R = 0, O = 0, Qg = 0;
  double rc = 0, rca = 0, scaling = 1;
  bool gated = false, species_readout = false;
  std::vector<int> bk, bl, bm;
  std::vector<Product> prod;
  std::vector<int> scalar_map, gate_nodes;
  int gate_begin = 0, outer_begin = 0, gate_pb = 0, gate_pe = 0, outer_pb = 0;
  // canonical moment c = h(l,m)*K + k -> basic node (or -1)
  std::vector<int> basic_node;
  // coefficient vector and offsets (reference layout)
  std::vector<double> theta;
  int o_shift = 0, o_scal = 0, o_rad = 0, o_grad = -1, o_gscal = -1, o_alpha = -1, o_gtype = -1, o_W = -1,
      o_gw = -1, o_ow = -1, o_species = 0, o_read = 0, ncoeff = 0;
  PassPlan gate_plan, outer_plan;

  void load(const std::string& path);
  void save(const std::string& path) const;
};

}  // namespace sus2
