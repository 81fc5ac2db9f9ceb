// mlp-sus2: C++ drop-in for the reference's EFS verbs over the C ABI
// (include/sus2amd.h); the arithmetic runs in libsus2amd.so on the GPU.
//
//   mlp-sus2 calc-efs pot.sus2 in.cfg out.cfg [--no-loss] [--precision fp64|fp32] [--batch-atoms N]
//   mlp-sus2 calc-errors pot.sus2 db.cfg [--precision fp64|fp32] [--batch-atoms N]
//
// calc-efs (dev_src/mlp/dev_mlp_commands.cpp:3037-3068) writes every
// configuration of in.cfg with its energy, forces, PlusStress and site
// energies (site_en = the readout part, without the per-species constant);
// --no-loss writes full double precision.  calc-errors
// (src/mlp/mlp_commands.cpp:793-808; metrics of
// dev_src/non_linear_regression.cpp:779-804, summary block of
// dev_src/mlp/mtpr_train.cpp:3893-3903) prints energy / force / stress MAE and
// RMSE against the labels of db.cfg.  CFG text as src/configuration.cpp:429
// (Load) and :1093 (Save).  Frames are evaluated in batches of up to
// --batch-atoms atoms per GPU call.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/sus2amd.h"

namespace {

struct Frame {
  int n = 0;
  bool periodic = false;
  std::array<double, 9> cell{};
  std::vector<int> type;
  std::vector<double> x, f, site;
  bool has_f = false, has_e = false, has_s = false;
  double energy = 0;
  std::array<double, 6> virial{};  // xx yy zz xy xz yz
};

std::vector<std::string> words(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> w;
  for (std::string t; is >> t;) w.push_back(t);
  return w;
}

// One BEGIN_CFG .. END_CFG block (the reader of the Python binding, cfg.py).
bool read_frame(std::istream& in, Frame& fr) {
  std::string line;
  bool inside = false;
  std::vector<std::string> block;
  while (std::getline(in, line)) {
    const auto w = words(line);
    if (!w.empty() && w[0] == "BEGIN_CFG") { inside = true; block.clear(); continue; }
    if (!w.empty() && w[0] == "END_CFG" && inside) break;
    if (inside) block.push_back(line);
  }
  if (!inside) return false;
  fr = Frame();
  for (size_t i = 0; i < block.size();) {
    const auto w = words(block[i]);
    if (w.empty()) { ++i; continue; }
    if (w[0] == "Size") {
      fr.n = std::stoi(words(block.at(i + 1))[0]);
      i += 2;
    } else if (w[0] == "Supercell") {
      size_t j = i + 1;
      int rows = 0;
      while (j < block.size() && rows < 3) {
        const auto p = words(block[j]);
        try {
          if (p.size() < 3) break;
          for (int q = 0; q < 3; ++q) fr.cell[3 * rows + q] = std::stod(p[q]);
        } catch (const std::exception&) { break; }
        ++rows;
        ++j;
      }
      if (!rows) throw std::runtime_error("Supercell without lattice vectors");
      for (double c : fr.cell) fr.periodic = fr.periodic || c != 0.0;
      i = j;
    } else if (w[0] == "AtomData:") {
      std::vector<std::string> cols(w.begin() + 1, w.end());
      auto col = [&](const char* nm) {
        for (size_t k = 0; k < cols.size(); ++k) if (cols[k] == nm) return (int)k;
        return -1;
      };
      const int ct = col("type"), cx = col("cartes_x"), dx = col("direct_x"), cf = col("fx");
      if (ct < 0 || (cx < 0 && dx < 0)) throw std::runtime_error("AtomData without type or coordinates");
      fr.type.resize(fr.n);
      fr.x.resize(3 * fr.n);
      fr.has_f = cf >= 0;
      if (fr.has_f) fr.f.resize(3 * fr.n);
      for (int a = 0; a < fr.n; ++a) {
        const auto p = words(block.at(i + 1 + a));
        fr.type[a] = (int)std::lround(std::stod(p.at(ct)));
        if (cx >= 0) {
          for (int q = 0; q < 3; ++q) fr.x[3 * a + q] = std::stod(p.at(cx + q));
        } else {
          double s[3];
          for (int q = 0; q < 3; ++q) s[q] = std::stod(p.at(dx + q));
          for (int q = 0; q < 3; ++q) fr.x[3 * a + q] = s[0] * fr.cell[q] + s[1] * fr.cell[3 + q] + s[2] * fr.cell[6 + q];
        }
        if (fr.has_f)
          for (int q = 0; q < 3; ++q) fr.f[3 * a + q] = std::stod(p.at(cf + q));
      }
      i += 1 + fr.n;
    } else if (w[0] == "Energy") {
      fr.energy = std::stod(words(block.at(i + 1))[0]);
      fr.has_e = true;
      i += 2;
    } else if (w[0].rfind("PlusStress", 0) == 0) {
      const auto names = words(block[i].substr(block[i].find(':') + 1));
      const auto vals = words(block.at(i + 1));
      for (size_t k = 0; k < names.size() && k < vals.size(); ++k) {
        const double v = std::stod(vals[k]);
        const std::string& nm = names[k];
        const int q = nm == "xx" ? 0 : nm == "yy" ? 1 : nm == "zz" ? 2 : nm == "xy" ? 3 : nm == "xz" ? 4 : nm == "yz" ? 5 : -1;
        if (q >= 0) fr.virial[q] = v;
      }
      fr.has_s = true;
      i += 2;
    } else {
      ++i;
    }
  }
  return true;
}

std::string fmt(const char* f, double v) {
  char b[64];
  std::snprintf(b, sizeof(b), f, v);
  return b;
}

// Configuration::Save layout (cfg.py write_cfg), with computed EFS.
void write_frame(std::ostream& out, const Frame& fr, bool lossless) {
  const char* lat = lossless ? "%.16e" : "%13.6f";
  const char* frc = lossless ? "%.16e" : "%11.6f";
  const char* sen = lossless ? "%.16e" : "%17.12f";
  const char* ene = lossless ? "%.16e" : "%20.12f";
  const char* st = lossless ? "%.16e" : "%11.5f";
  out << "BEGIN_CFG\n Size\n    " << fr.n << "\n";
  if (fr.periodic) {
    int count = 0;
    for (int q = 0; q < 3; ++q)
      if (fr.cell[3 * q] != 0 || fr.cell[3 * q + 1] != 0 || fr.cell[3 * q + 2] != 0) count = q + 1;
    if (count) {
      out << " Supercell\n";
      for (int q = 0; q < count; ++q)
        out << "    " << fmt(lat, fr.cell[3 * q]) << " " << fmt(lat, fr.cell[3 * q + 1]) << " "
            << fmt(lat, fr.cell[3 * q + 2]) << "\n";
    }
  }
  out << " AtomData:  id type       cartes_x      cartes_y      cartes_z    site_en           fx          fy          fz\n";
  for (int a = 0; a < fr.n; ++a) {
    char head[64];
    if (lossless) std::snprintf(head, sizeof(head), "    %d %d  ", a + 1, fr.type[a]);
    else std::snprintf(head, sizeof(head), "    %10d %4d  ", a + 1, fr.type[a]);
    out << head << fmt(lat, fr.x[3 * a]) << " " << fmt(lat, fr.x[3 * a + 1]) << " " << fmt(lat, fr.x[3 * a + 2])
        << "  " << fmt(sen, fr.site[a]) << "  " << fmt(frc, fr.f[3 * a]) << " " << fmt(frc, fr.f[3 * a + 1]) << " "
        << fmt(frc, fr.f[3 * a + 2]) << "\n";
  }
  const auto& w = fr.virial;
  out << " Energy\n    " << fmt(ene, fr.energy) << "\n";
  out << " PlusStress:  xx          yy          zz          yz          xz          xy\n     " << fmt(st, w[0]) << " "
      << fmt(st, w[1]) << " " << fmt(st, w[2]) << " " << fmt(st, w[5]) << " " << fmt(st, w[4]) << " " << fmt(st, w[3])
      << "\nEND_CFG\n\n";
}

void check(int rc) {
  if (rc != SUS2_OK) throw std::runtime_error(sus2_last_error());
}

// EFS of a batch of frames (host buffers); results written into the frames.
void evaluate(sus2_model* m, int precision, const std::vector<Frame>& batch, std::vector<Frame>* computed) {
  const int nf = (int)batch.size();
  std::vector<int64_t> off(nf + 1, 0);
  for (int k = 0; k < nf; ++k) off[k + 1] = off[k] + batch[k].n;
  const int n = (int)off[nf];
  std::vector<int32_t> types(n), per(nf);
  std::vector<double> pos(3 * (size_t)n), cell(9 * (size_t)nf), E(nf), F(3 * (size_t)n), V(6 * (size_t)nf), S(n);
  for (int k = 0; k < nf; ++k) {
    const Frame& fr = batch[k];
    for (int a = 0; a < fr.n; ++a) {
      types[off[k] + a] = fr.type[a];
      for (int q = 0; q < 3; ++q) pos[3 * (off[k] + a) + q] = fr.x[3 * a + q];
    }
    for (int q = 0; q < 9; ++q) cell[9 * k + q] = fr.cell[q];
    per[k] = fr.periodic ? 1 : 0;
  }
  check(sus2_efs_host(m, precision, nf, n, off.data(), types.data(), pos.data(), cell.data(), per.data(), E.data(),
                      F.data(), V.data(), S.data()));
  // per-species constants (shift + species) removed from the site energies
  int o[13];
  check(sus2_coeff_offsets(m, o));
  std::vector<double> theta(sus2_coeff_count(m));
  check(sus2_coeff_get(m, theta.data()));
  for (int k = 0; k < nf; ++k) {
    Frame g = batch[k];
    g.energy = E[k];
    g.f.assign(F.begin() + 3 * off[k], F.begin() + 3 * off[k + 1]);
    for (int q = 0; q < 6; ++q) g.virial[q] = V[6 * k + q];
    g.site.resize(g.n);
    for (int a = 0; a < g.n; ++a) g.site[a] = S[off[k] + a] - (theta[o[10] + g.type[a]] + theta[o[0] + g.type[a]]);
    computed->push_back(std::move(g));
  }
}

int usage() {
  std::cerr << "usage: mlp-sus2 calc-efs pot.sus2 in.cfg out.cfg [--no-loss] [--precision fp64|fp32] [--batch-atoms N]\n"
               "       mlp-sus2 calc-errors pot.sus2 db.cfg [--precision fp64|fp32] [--batch-atoms N]\n";
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc < 4) return usage();
    const std::string cmd = argv[1];
    std::vector<std::string> pos;
    bool lossless = false;
    int precision = SUS2_FP64;
    long max_atoms = 1L << 20;
    for (int k = 2; k < argc; ++k) {
      const std::string a = argv[k];
      if (a == "--no-loss") lossless = true;
      else if (a == "--precision" && k + 1 < argc) {
        const std::string p = argv[++k];
        if (p == "fp64") precision = SUS2_FP64;
        else if (p == "fp32") precision = SUS2_FP32;
        else return usage();
      } else if (a == "--batch-atoms" && k + 1 < argc) max_atoms = std::stol(argv[++k]);
      else pos.push_back(a);
    }
    if ((cmd == "calc-efs" && pos.size() != 3) || (cmd == "calc-errors" && pos.size() != 2) ||
        (cmd != "calc-efs" && cmd != "calc-errors"))
      return usage();
    sus2_model* m = nullptr;
    check(sus2_model_load(pos[0].c_str(), &m));
    std::ifstream in(pos[1]);
    if (!in) throw std::runtime_error("cannot open " + pos[1]);
    std::ofstream out;
    if (cmd == "calc-efs") {
      out.open(pos[2]);
      if (!out) throw std::runtime_error("cannot open " + pos[2]);
    }
    long cfgs = 0, atoms = 0, f_cnt = 0, s_cnt = 0;
    double e_abs = 0, e_sq = 0, f_abs = 0, f_sq = 0, s_abs = 0, s_sq = 0;
    std::vector<Frame> batch;
    long batch_atoms = 0;
    auto flush = [&]() {
      if (batch.empty()) return;
      std::vector<Frame> done;
      evaluate(m, precision, batch, &done);
      for (size_t k = 0; k < done.size(); ++k) {
        const Frame& lab = batch[k];
        const Frame& g = done[k];
        if (cmd == "calc-efs") {
          write_frame(out, g, lossless);
          continue;
        }
        ++cfgs;
        if (lab.has_e) {
          const double d = lab.energy - g.energy;
          e_abs += std::fabs(d);
          e_sq += d * d / lab.n;
          atoms += lab.n;
        }
        if (lab.has_f) {
          for (int q = 0; q < 3 * lab.n; ++q) {
            const double d = g.f[q] - lab.f[q];
            f_abs += std::fabs(d);
            f_sq += d * d;
          }
          f_cnt += 3 * lab.n;
        }
        if (lab.has_s) {
          for (int q = 0; q < 6; ++q) {
            const double d = g.virial[q] - lab.virial[q];
            s_abs += std::fabs(d);
            s_sq += d * d;
          }
          s_cnt += 6;
        }
      }
      batch.clear();
      batch_atoms = 0;
    };
    for (Frame fr; read_frame(in, fr);) {
      if (!batch.empty() && batch_atoms + fr.n > max_atoms) flush();
      batch_atoms += fr.n;
      batch.push_back(fr);
    }
    flush();
    if (cmd == "calc-errors") {
      std::printf("\n=== Train Summary ===\n"
                  "Structures           : %ld\n"
                  "Atoms                : %ld\n"
                  "Energy MAE (meV/atom): %.3f\n"
                  "Energy RMSE(meV/atom): %.3f\n"
                  "Force MAE  (meV/A)   : %.3f\n"
                  "Force RMSE (meV/A)   : %.3f\n"
                  "Stress MAE (eV)      : %.3f\n"
                  "Stress RMSE(eV)      : %.3f\n"
                  "=====================\n",
                  cfgs, atoms, atoms ? 1000.0 * e_abs / atoms : 0.0, atoms ? 1000.0 * std::sqrt(e_sq / atoms) : 0.0,
                  f_cnt ? 1000.0 * f_abs / f_cnt : 0.0, f_cnt ? 1000.0 * std::sqrt(f_sq / f_cnt) : 0.0,
                  s_cnt ? s_abs / s_cnt : 0.0, s_cnt ? std::sqrt(s_sq / s_cnt) : 0.0);
    }
    sus2_model_free(m);
  } catch (const std::exception& e) {
    std::cerr << "mlp-sus2: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
