// Timing harness of the reference embedded EFS engine -- TEST INFRASTRUCTURE
// ONLY (the CPU-baseline calibration).  Linked against the reference's own
// sources (interfaces/lammps/ML-SUS2/sus2_sl_*.cpp) by oracle/Makefile into
// oracle/_ref/; nothing in the product path uses it.
//
// Usage: ref-timing MODEL FRAMES REPS
// FRAMES is a text file: the frame count, then per frame the atom count, nine
// lattice numbers (rows), a periodic flag, and one "type x y z" line per atom.
// Every frame's full directed edge list (strict r < cutoff, every periodic
// image, as interfaces/lammps/tests/embedded_efs.cpp:44-66) is built once;
// the timed region is REPS x (Kernel::begin, outer, finish) over all frames
// with one domain per frame (the engine's serial EFS, direct radial mode).
// Prints REF_TIMING_JSON {"seconds_per_rep": s, "atoms": n, "edges": e, "energy_sum": E}.
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sus2_sl_kernel.h"

using namespace LAMMPS_NS::SUS2SL;

namespace {
struct Frame {
  std::array<double, 9> cell{};
  bool periodic = true;
  std::vector<int> type;
  std::vector<std::array<double, 3>> x;
};

std::vector<Frame> read_frames(const char* path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open frames file");
  int nf = 0;
  in >> nf;
  std::vector<Frame> out(nf);
  for (auto& f : out) {
    int n = 0, per = 0;
    in >> n;
    for (double& c : f.cell) in >> c;
    in >> per;
    f.periodic = per != 0;
    f.type.resize(n);
    f.x.resize(n);
    for (int i = 0; i < n; ++i) in >> f.type[i] >> f.x[i][0] >> f.x[i][1] >> f.x[i][2];
    if (!in) throw std::runtime_error("truncated frames file");
  }
  return out;
}

// Brute-force images: the range of lattice shifts whose image can lie within
// the cutoff along each direction (periodic frames); a molecule has none.
std::vector<Edge> frame_edges(const Frame& f, double rc) {
  const int n = (int)f.x.size();
  std::array<double, 9> inv{};
  std::array<double, 3> reach{0, 0, 0};
  if (f.periodic) {
    const auto& a = f.cell;
    double b[9];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q)
        b[3 * q + p] = a[3 * ((p + 1) % 3) + (q + 1) % 3] * a[3 * ((p + 2) % 3) + (q + 2) % 3] -
                       a[3 * ((p + 1) % 3) + (q + 2) % 3] * a[3 * ((p + 2) % 3) + (q + 1) % 3];
    const double det = a[0] * b[0] + a[1] * b[3] + a[2] * b[6];
    for (int k = 0; k < 9; ++k) inv[k] = b[k] / det;
    for (int q = 0; q < 3; ++q)
      reach[q] = rc * std::sqrt(inv[q] * inv[q] + inv[3 + q] * inv[3 + q] + inv[6 + q] * inv[6 + q]);
  }
  std::vector<Edge> out;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      std::array<double, 3> d{f.x[j][0] - f.x[i][0], f.x[j][1] - f.x[i][1], f.x[j][2] - f.x[i][2]};
      int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      if (f.periodic)
        for (int q = 0; q < 3; ++q) {
          const double fq = d[0] * inv[q] + d[1] * inv[3 + q] + d[2] * inv[6 + q];
          lo[q] = (int)std::ceil(-fq - reach[q]);
          hi[q] = (int)std::floor(-fq + reach[q]);
        }
      for (int a = lo[0]; a <= hi[0]; ++a)
        for (int b = lo[1]; b <= hi[1]; ++b)
          for (int c = lo[2]; c <= hi[2]; ++c) {
            if (i == j && a == 0 && b == 0 && c == 0) continue;
            std::array<double, 3> v = d;
            double r2 = 0;
            for (int q = 0; q < 3; ++q) {
              v[q] += a * f.cell[q] + b * f.cell[3 + q] + c * f.cell[6 + q];
              r2 += v[q] * v[q];
            }
            if (r2 < rc * rc) out.push_back({i, j, v});
          }
    }
  return out;
}
}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc != 4) throw std::runtime_error("usage: ref-timing MODEL FRAMES REPS");
    Model model;
    model.load(argv[1]);
    const auto frames = read_frames(argv[2]);
    const int reps = std::stoi(argv[3]);
    std::vector<std::vector<Edge>> edges;
    long natoms = 0, nedges = 0;
    for (const auto& f : frames) {
      edges.push_back(frame_edges(f, model.cutoff()));
      natoms += (long)f.x.size();
      nedges += (long)edges.back().size();
    }
    Kernel k(model);
    k.configure_radial(RadialTableMode::Direct);
    double esum = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r)
      for (size_t f = 0; f < frames.size(); ++f) {
        const int n = (int)frames[f].x.size();
        k.begin(n, n, frames[f].type, edges[f]);
        k.outer();
        esum += k.finish().energy;
      }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("REF_TIMING_JSON {\"seconds_per_rep\": %.9g, \"atoms\": %ld, \"edges\": %ld, \"energy_sum\": %.17g}\n",
                dt / reps, natoms, nedges, esum / reps);
  } catch (const std::exception& e) {
    std::cerr << e.what() << '\n';
    return 1;
  }
  return 0;
}
