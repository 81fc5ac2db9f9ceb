// Host FP64 optimisers of the training objective (no HIP: the objective and
// its gradient are a callback, evaluated on the GPU by the C ABI layer).
//
// BFGS follows the protocol of the reference's BFGS::Iterate
// (src/common/bfgs.cpp:251-330, src/common/bfgs.h:384-385): a dense inverse
// Hessian H starting at the identity, direction p = -H g, a line search that
// accepts a step alpha when the strong Wolfe conditions hold with c1 = 0.1,
// c2 = 0.5 (f <= f0 + c1 alpha p.g0 and |p.g| <= c2 |p.g0|), then the update
//   H += (s.y + y.Hy) / (s.y)^2 s s^T - (Hy s^T + s (Hy)^T) / (s.y),
// s = alpha p, y = g - g0, skipped when s.y <= 64 eps (bfgs.cpp:420-470).
// The line search itself is written here from the textbook bracketing/zoom
// scheme with safeguarded cubic interpolation (Nocedal & Wright, Numerical
// Optimization, 2nd ed., Algorithms 3.5 and 3.6).
#pragma once

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <stdexcept>
#include <vector>

namespace sus2 {

// f(x) with its gradient written to g.
using Objective = std::function<double(const std::vector<double>& x, std::vector<double>& g)>;

struct BfgsResult {
  int iters = 0;        // accepted steps
  int evals = 0;        // objective evaluations
  bool converged = false;
  std::vector<double> trace;  // objective at the start and after each accepted step
};

namespace optim_detail {
inline double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t k = 0; k < a.size(); ++k) s += a[k] * b[k];
  return s;
}

// Minimiser of the cubic interpolating (a, fa, ga), (b, fb, gb), clamped into
// the interval between a and b with a 10 % margin; bisection when the cubic
// has no real minimiser.
inline double cubic_step(double a, double fa, double ga, double b, double fb, double gb) {
  const double lo = std::min(a, b), hi = std::max(a, b), w = hi - lo;
  const double d1 = ga + gb - 3.0 * (fa - fb) / (a - b);
  const double disc = d1 * d1 - ga * gb;
  double t = 0.5 * (a + b);
  if (disc >= 0 && std::isfinite(disc)) {
    const double d2 = (b > a ? 1.0 : -1.0) * std::sqrt(disc);
    const double den = gb - ga + 2.0 * d2;
    if (den != 0 && std::isfinite(den)) t = b - (b - a) * (gb + d2 - d1) / den;
  }
  if (!std::isfinite(t) || t < lo + 0.1 * w || t > hi - 0.1 * w) t = 0.5 * (a + b);
  return t;
}
}  // namespace optim_detail

// Strong-Wolfe line search along p from x0 (f0, g0, slope d0 = p.g0 < 0).
// On success x, f, g hold the accepted point and the step is returned; on
// failure (no acceptable step within max_evals) returns 0 and leaves x, f, g
// at the best point seen if it lowered f, else at x0.
inline double wolfe_search(const Objective& fn, const std::vector<double>& x0, double f0,
                           const std::vector<double>& g0, const std::vector<double>& p, double d0, double step0,
                           std::vector<double>& x, double& f, std::vector<double>& g, int& evals,
                           double c1 = 0.1, double c2 = 0.5, int max_evals = 40, double step_max = 1e10) {
  using optim_detail::dot;
  const size_t n = x0.size();
  auto eval = [&](double a, std::vector<double>& xa, std::vector<double>& ga) {
    for (size_t k = 0; k < n; ++k) xa[k] = x0[k] + a * p[k];
    ++evals;
    return fn(xa, ga);
  };
  std::vector<double> xa(n), ga(n), xb(n), gb(n);
  double a_prev = 0, f_prev = f0, d_prev = d0;
  double a = step0;
  double best_f = f0, best_a = 0;
  std::vector<double> best_x = x0, best_g = g0;
  auto keep_best = [&](double aa, double fa, const std::vector<double>& xx, const std::vector<double>& gg) {
    if (std::isfinite(fa) && fa < best_f) {
      best_f = fa;
      best_a = aa;
      best_x = xx;
      best_g = gg;
    }
  };
  // zoom between lo (Armijo-satisfying, lower f) and hi
  auto zoom = [&](double lo, double flo, double dlo, double hi, double fhi, double dhi) -> double {
    for (int it = 0; evals < max_evals && it < max_evals; ++it) {
      const double t = optim_detail::cubic_step(lo, flo, dlo, hi, fhi, dhi);
      const double ft = eval(t, xb, gb);
      const double dt = dot(gb, p);
      keep_best(t, ft, xb, gb);
      if (!std::isfinite(ft) || ft > f0 + c1 * t * d0 || ft >= flo) {
        hi = t; fhi = ft; dhi = std::isfinite(dt) ? dt : dhi;
        if (!std::isfinite(ft)) fhi = std::numeric_limits<double>::max();
      } else {
        if (std::fabs(dt) <= -c2 * d0) {
          x = xb; f = ft; g = gb;
          return t;
        }
        if (dt * (hi - lo) >= 0) { hi = lo; fhi = flo; dhi = dlo; }
        lo = t; flo = ft; dlo = dt;
      }
      if (std::fabs(hi - lo) <= 1e-14 * std::max(1.0, std::fabs(lo))) break;
    }
    return 0.0;
  };
  for (int it = 0; evals < max_evals; ++it) {
    const double fa = eval(a, xa, ga);
    const double da = dot(ga, p);
    keep_best(a, fa, xa, ga);
    if (!std::isfinite(fa) || fa > f0 + c1 * a * d0 || (it > 0 && fa >= f_prev)) {
      const double r = zoom(a_prev, f_prev, d_prev, a, std::isfinite(fa) ? fa : std::numeric_limits<double>::max(),
                            std::isfinite(da) ? da : -d_prev);
      if (r > 0) return r;
      break;
    }
    if (std::fabs(da) <= -c2 * d0) {
      x = xa; f = fa; g = ga;
      return a;
    }
    if (da >= 0) {
      const double r = zoom(a, fa, da, a_prev, f_prev, d_prev);
      if (r > 0) return r;
      break;
    }
    a_prev = a; f_prev = fa; d_prev = da;
    a = std::min(2.0 * a, step_max);
  }
  x = best_x; f = best_f; g = best_g;
  (void)best_a;
  return 0.0;
}

// BFGS minimisation from x (overwritten with the last accepted point).
// Stops when max |g| <= gtol, after max_iter accepted steps, or when a line
// search finds no acceptable step (the best point it saw is kept when lower).
// `on_update` (may be empty) sees H after every update (for tests).
inline BfgsResult bfgs_minimize(const Objective& fn, std::vector<double>& x, int max_iter, double gtol,
                                const std::function<void(const std::vector<double>& H, const std::vector<double>& s,
                                                         const std::vector<double>& y)>& on_update = {}) {
  using optim_detail::dot;
  const size_t n = x.size();
  BfgsResult res;
  std::vector<double> g(n), H(n * n, 0.0), p(n), xn(n), gn(n), s(n), y(n), Hy(n);
  for (size_t k = 0; k < n; ++k) H[k * n + k] = 1.0;
  double f = fn(x, g);
  ++res.evals;
  if (!std::isfinite(f)) throw std::runtime_error("BFGS: non-finite objective at the starting point");
  res.trace.push_back(f);
  for (int it = 0; it < max_iter; ++it) {
    double gmax = 0;
    for (double v : g) gmax = std::max(gmax, std::fabs(v));
    if (gmax <= gtol) {
      res.converged = true;
      break;
    }
    for (size_t i = 0; i < n; ++i) {
      double t = 0;
      for (size_t j = 0; j < n; ++j) t -= H[i * n + j] * g[j];
      p[i] = t;
    }
    double d0 = dot(p, g);
    if (!(d0 < 0)) {  // not a descent direction: restart from the identity
      std::fill(H.begin(), H.end(), 0.0);
      for (size_t k = 0; k < n; ++k) H[k * n + k] = 1.0;
      for (size_t k = 0; k < n; ++k) p[k] = -g[k];
      d0 = dot(p, g);
    }
    double fn_new = f;
    int ev = 0;
    const double step = wolfe_search(fn, x, f, g, p, d0, 1.0, xn, fn_new, gn, ev);
    res.evals += ev;
    if (step <= 0) {  // no acceptable step: keep a lower point if one was seen
      if (fn_new < f) {
        x = xn;
        g = gn;
        f = fn_new;
        res.trace.push_back(f);
      }
      break;
    }
    for (size_t k = 0; k < n; ++k) {
      s[k] = xn[k] - x[k];
      y[k] = gn[k] - g[k];
    }
    const double sy = dot(s, y);
    if (sy > 64.0 * std::numeric_limits<double>::epsilon()) {
      for (size_t i = 0; i < n; ++i) {
        double t = 0;
        for (size_t j = 0; j < n; ++j) t += H[i * n + j] * y[j];
        Hy[i] = t;
      }
      const double yHy = dot(y, Hy);
      const double a = (sy + yHy) / (sy * sy), b = 1.0 / sy;
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) H[i * n + j] += a * s[i] * s[j] - b * (Hy[i] * s[j] + s[i] * Hy[j]);
      if (on_update) on_update(H, s, y);
    }
    x.swap(xn);
    g.swap(gn);
    f = fn_new;
    res.iters = it + 1;
    res.trace.push_back(f);
  }
  return res;
}

// L-BFGS with the update rules of the reference's LBFGS
// (src/common/lbfgs.cpp:129-327, constants src/common/lbfgs.h:58,93-98):
// direction from the two-loop recursion with H0 = gamma I (gamma = s.y / y.y
// of the last accepted pair, clamped to [1e-12, 1e12]); a non-descent
// direction clears the history and falls back to -g.  Armijo backtracking by
// halving (c1 = 1e-4) from the next initial step: 1 after an accepted history
// update, else the last step (times 1.5 when it needed no backtracking).  The
// pair (s, y) is Powell-damped when s.y < 0.2 s.B0.s (B0 = I / gamma):
//   y <- theta y + (1 - theta) B0 s,  theta = 0.8 s.B0.s / (s.B0.s - s.y),
// and skipped unless s.y > 64 eps sqrt(s.B0.s y.H0.y) and its relative
// curvature s.y / sqrt(s.B0.s y.H0.y) >= 1e-10.  The line search stops
// (stagnates) when halving no longer changes the step or it falls below 1e-20.
struct LbfgsResult {
  int iters = 0, evals = 0, damped = 0, skipped = 0, resets = 0;
  bool converged = false;
  std::vector<double> trace;  // objective at the start and after each accepted step
  std::vector<double> gammas;  // H0 scale after each accepted step
};

inline LbfgsResult lbfgs_minimize(const Objective& fn, std::vector<double>& x, int max_iter, int history,
                                  double gtol) {
  using optim_detail::dot;
  constexpr double c1 = 1e-4, powell = 0.2, rel_tol = 1e-10, min_step = 1e-20, max_step = 1e20, growth = 1.5;
  const size_t n = x.size();
  LbfgsResult res;
  std::vector<std::vector<double>> S, Y;
  std::vector<double> rho, alpha;
  double gamma = 1.0, next_step = 1.0;
  std::vector<double> g(n), p(n), q(n), xn(n), gn(n), s(n), y(n);
  double f = fn(x, g);
  ++res.evals;
  if (!std::isfinite(f)) throw std::runtime_error("L-BFGS: non-finite objective at the starting point");
  res.trace.push_back(f);
  for (int it = 0; it < max_iter; ++it) {
    double gmax = 0;
    for (double v : g) gmax = std::max(gmax, std::fabs(v));
    if (gmax <= gtol) {
      res.converged = true;
      break;
    }
    // two-loop recursion: p = -H g
    q = g;
    alpha.assign(S.size(), 0.0);
    for (int k = (int)S.size() - 1; k >= 0; --k) {
      alpha[k] = rho[k] * dot(S[k], q);
      for (size_t t = 0; t < n; ++t) q[t] -= alpha[k] * Y[k][t];
    }
    for (size_t t = 0; t < n; ++t) q[t] *= gamma;
    for (size_t k = 0; k < S.size(); ++k) {
      const double beta = rho[k] * dot(Y[k], q);
      for (size_t t = 0; t < n; ++t) q[t] += (alpha[k] - beta) * S[k][t];
    }
    for (size_t t = 0; t < n; ++t) p[t] = -q[t];
    double pg = dot(p, g);
    if (!(pg < 0)) {  // not a descent direction: clear the history, steepest descent
      S.clear(); Y.clear(); rho.clear();
      gamma = 1.0;
      ++res.resets;
      for (size_t t = 0; t < n; ++t) p[t] = -g[t];
      pg = dot(p, g);
      if (!(pg < 0)) break;
    }
    // Armijo backtracking by halving
    double a = std::min(std::max(next_step, min_step), max_step), fn_new = 0;
    int backtracks = 0;
    bool ok = false;
    for (;;) {
      for (size_t t = 0; t < n; ++t) xn[t] = x[t] + a * p[t];
      fn_new = fn(xn, gn);
      ++res.evals;
      if (std::isfinite(fn_new) && fn_new <= f + c1 * a * pg) {
        ok = true;
        break;
      }
      const double a2 = 0.5 * a;
      if (std::fabs(a2 - a) <= 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, a) || a2 < min_step)
        break;  // stagnated
      a = a2;
      ++backtracks;
    }
    if (!ok) break;  // x, g, f stay at the last accepted point
    next_step = std::min(std::max(backtracks == 0 ? a * growth : a, min_step), max_step);
    for (size_t t = 0; t < n; ++t) {
      s[t] = xn[t] - x[t];
      y[t] = gn[t] - g[t];
    }
    // curvature pair, Powell-damped against B0 = I / gamma
    const double sy_raw = dot(s, y), sBs = dot(s, s) / gamma, yHy_raw = gamma * dot(y, y);
    bool accepted = false;
    if (sBs > 0 && yHy_raw > 0 && std::isfinite(sBs * yHy_raw)) {
      std::vector<double> yd = y;
      bool damped = false;
      if (sy_raw < powell * sBs) {
        const double den = sBs - sy_raw;
        if (den > 0 && std::isfinite(den)) {
          const double theta = std::min(1.0, std::max(0.0, (1.0 - powell) * sBs / den));
          for (size_t t = 0; t < n; ++t) yd[t] = theta * y[t] + (1.0 - theta) * s[t] / gamma;
          damped = true;
        } else {
          yd.clear();
        }
      }
      if (!yd.empty()) {
        const double sy = dot(s, yd), yHy = gamma * dot(yd, yd);
        const double scale = std::sqrt(sBs * yHy);
        if (yHy > 0 && std::isfinite(yHy) && sy > 64.0 * std::numeric_limits<double>::epsilon() * scale &&
            sy / scale >= rel_tol && std::isfinite(1.0 / sy)) {
          if ((int)S.size() == history) {
            S.erase(S.begin()); Y.erase(Y.begin()); rho.erase(rho.begin());
          }
          S.push_back(s);
          Y.push_back(yd);
          rho.push_back(1.0 / sy);
          const double gc = sy / dot(yd, yd);
          if (gc > 0 && std::isfinite(gc)) gamma = std::min(1e12, std::max(1e-12, gc));
          accepted = true;
          res.damped += damped ? 1 : 0;
        }
      }
    }
    if (accepted) next_step = 1.0;
    else ++res.skipped;
    x.swap(xn);
    g.swap(gn);
    f = fn_new;
    res.iters = it + 1;
    res.trace.push_back(f);
    res.gammas.push_back(gamma);
  }
  return res;
}

}  // namespace sus2
