// Test driver of the host BFGS and L-BFGS (superlinear_amd/csrc/sus2_optim.hpp),
// built with g++ by tests/test_host.py.  Usage:
//   optim_driver quad FILE   minimise 1/2 x^T A x - b^T x (FILE: n, A row-major, b)
//   optim_driver rosen       minimise the Rosenbrock function from (-1.2, 1)
//   optim_driver lbfgs-quad FILE HISTORY | lbfgs-1d | lbfgs-dw | lbfgs-rosen   (L-BFGS)
// Prints the iterations, evaluations, the final point and, for the quadratic,
// the largest secant residual |H y - s| over all inverse-Hessian updates.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../superlinear_amd/csrc/sus2_optim.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  if (mode == "quad" && argc == 3) {
    std::ifstream in(argv[2]);
    int n = 0;
    in >> n;
    std::vector<double> A(n * n), b(n);
    for (double& v : A) in >> v;
    for (double& v : b) in >> v;
    sus2::Objective fn = [&](const std::vector<double>& x, std::vector<double>& g) {
      double f = 0;
      for (int i = 0; i < n; ++i) {
        double t = 0;
        for (int j = 0; j < n; ++j) t += A[i * n + j] * x[j];
        g[i] = t - b[i];
        f += 0.5 * x[i] * t - b[i] * x[i];
      }
      return f;
    };
    double secant = 0;
    std::vector<double> x(n, 0.0);
    auto r = sus2::bfgs_minimize(fn, x, 200, 1e-12, [&](const std::vector<double>& H, const std::vector<double>& s,
                                                        const std::vector<double>& y) {
      for (int i = 0; i < n; ++i) {
        double t = 0;
        for (int j = 0; j < n; ++j) t += H[i * n + j] * y[j];
        secant = std::fmax(secant, std::fabs(t - s[i]));
      }
    });
    std::printf("iters %d evals %d converged %d secant %.3e\n", r.iters, r.evals, (int)r.converged, secant);
    for (double v : x) std::printf("%.17g\n", v);
    return 0;
  }
  if (mode == "lbfgs-quad" && argc == 4) {  // FILE HISTORY
    std::ifstream in(argv[2]);
    int n = 0;
    in >> n;
    std::vector<double> A(n * n), b(n);
    for (double& v : A) in >> v;
    for (double& v : b) in >> v;
    sus2::Objective fn = [&](const std::vector<double>& x, std::vector<double>& g) {
      double f = 0;
      for (int i = 0; i < n; ++i) {
        double t = 0;
        for (int j = 0; j < n; ++j) t += A[i * n + j] * x[j];
        g[i] = t - b[i];
        f += 0.5 * x[i] * t - b[i] * x[i];
      }
      return f;
    };
    std::vector<double> x(n, 0.0);
    auto r = sus2::lbfgs_minimize(fn, x, 2000, std::stoi(argv[3]), 1e-9);
    std::printf("iters %d evals %d converged %d damped %d\n", r.iters, r.evals, (int)r.converged, r.damped);
    for (double v : x) std::printf("%.17g\n", v);
    for (size_t k = 1; k < r.trace.size(); ++k)
      if (r.trace[k] > r.trace[k - 1]) return 3;
    return 0;
  }
  if (mode == "lbfgs-1d") {  // f = c x^2 / 2 with c = 0.5 from x = 1: gamma after the first pair is 1 / c
    sus2::Objective fn = [](const std::vector<double>& x, std::vector<double>& g) {
      g[0] = 0.5 * x[0];
      return 0.25 * x[0] * x[0];
    };
    std::vector<double> x = {1.0};
    auto r = sus2::lbfgs_minimize(fn, x, 20, 5, 0.0);
    std::printf("iters %d evals %d converged %d gamma %.17g\n%.17g\n", r.iters, r.evals, (int)r.converged,
                r.gammas.empty() ? 0.0 : r.gammas[0], x[0]);
    return 0;
  }
  if (mode == "lbfgs-dw") {  // double well x^4/4 - x^2/2 + y^2/2 from the concave region
    sus2::Objective fn = [](const std::vector<double>& x, std::vector<double>& g) {
      g[0] = x[0] * x[0] * x[0] - x[0];
      g[1] = x[1];
      return 0.25 * x[0] * x[0] * x[0] * x[0] - 0.5 * x[0] * x[0] + 0.5 * x[1] * x[1];
    };
    std::vector<double> x = {0.1, 0.3};
    auto r = sus2::lbfgs_minimize(fn, x, 200, 5, 1e-10);
    std::printf("iters %d evals %d converged %d damped %d\n%.17g\n%.17g\n", r.iters, r.evals, (int)r.converged,
                r.damped, x[0], x[1]);
    return 0;
  }
  if (mode == "lbfgs-rosen") {
    sus2::Objective fn = [](const std::vector<double>& x, std::vector<double>& g) {
      const double a = 1 - x[0], c = x[1] - x[0] * x[0];
      g[0] = -2 * a - 400 * x[0] * c;
      g[1] = 200 * c;
      return a * a + 100 * c * c;
    };
    std::vector<double> x = {-1.2, 1.0};
    auto r = sus2::lbfgs_minimize(fn, x, 1000, 8, 1e-10);
    std::printf("iters %d evals %d converged %d damped %d\n%.17g\n%.17g\n", r.iters, r.evals, (int)r.converged,
                r.damped, x[0], x[1]);
    return 0;
  }
  if (mode == "rosen") {
    sus2::Objective fn = [](const std::vector<double>& x, std::vector<double>& g) {
      const double a = 1 - x[0], c = x[1] - x[0] * x[0];
      g[0] = -2 * a - 400 * x[0] * c;
      g[1] = 200 * c;
      return a * a + 100 * c * c;
    };
    std::vector<double> x = {-1.2, 1.0};
    auto r = sus2::bfgs_minimize(fn, x, 200, 1e-10);
    std::printf("iters %d evals %d converged %d secant 0\n%.17g\n%.17g\n", r.iters, r.evals, (int)r.converged, x[0], x[1]);
    for (size_t k = 1; k < r.trace.size(); ++k)
      if (r.trace[k] > r.trace[k - 1]) return 3;  // accepted steps never raise f
    return 0;
  }
  return 2;
}
