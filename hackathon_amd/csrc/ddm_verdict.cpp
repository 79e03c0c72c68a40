// Motility verdict layer on the host: the reference's analyzeMotility
// (motility-ddm src/motility_fit.js:808-990) over one Radial record, in FP64.
//
// Pass 1 fits the Brownian null model (log-D grid + golden section,
// :500-536) and the Schulz swimmer model (64-start scan + Nelder-Mead from the
// best 3, :577-611), both as variable projection with per-bin NNLS amplitudes
// (evalModel, :445-472); estimates the uniform drift from the phase sums
// (estimateDrift, :671-725).  A significant drift is folded into the null
// model (fitBrownianDrift, :544-567) and the swimmer refit with its speed
// (pass 2).  The models are compared by BIC on the AR(1)-effective sample size
// of the swimmer residuals (:622-656); TSI, the motile share, the q-scaling
// of tau_1/2 and gates G1-G5 give the verdict (:730-990).
//
// The steps, grids, start sets, tolerances and tie rules follow the reference
// (they decide which local optimum is reported); the arithmetic is C++ double
// with libm, so values agree with the reference's to rounding, not bit for bit.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ddm.h"
#include "ddm_isf2d.h"

namespace {

constexpr double kTsiMin = 0.15, kFlowSigma = 3.0, kFlowMin = 2.0, kDbicMin = 10.0, kAlphaMin = 0.05;
constexpr double kDriftIsfMax = 0.2, kSwimMin = 2.0;

thread_local std::string g_err;

// ---------------------------------------------------------------- kernels

double bessel_j0(double x) {   // rational approximations, |error| < 1e-8 (src/motility_fit.js:52-67)
  const double ax = std::fabs(x);
  if (ax < 8) {
    const double y = x * x;
    const double num = 57568490574.0 +
                       y * (-13362590354.0 + y * (651619640.7 + y * (-11214424.18 + y * (77392.33017 + y * -184.9052456))));
    const double den = 57568490411.0 + y * (1029532985.0 + y * (9494680.718 + y * (59272.64853 + y * (267.8532712 + y))));
    return num / den;
  }
  const double z = 8 / ax, y = z * z, xx = ax - 0.785398164;
  const double p = 1 + y * (-0.1098628627e-2 + y * (0.2734510407e-4 + y * (-0.2073370639e-5 + y * 0.2093887211e-6)));
  const double q = -0.1562499995e-1 + y * (0.1430488765e-3 + y * (-0.6911147651e-5 + y * (0.7621095161e-6 - y * 0.934935152e-7)));
  return std::sqrt(0.636619772 / ax) * (std::cos(xx) * p - z * std::sin(xx) * q);
}

// 3D swimmer ISF: sin(Z atan u) / (Z u (1+u^2)^(Z/2)), u = x/(Z+1) (:135-144)
double isf3d(double x, double Z) {
  const double u = x / (Z + 1);
  if (u < 1e-4) return 1 - x * x * (Z + 2) / (6 * (Z + 1));
  return std::sin(Z * std::atan(u)) / (Z * u * std::exp(0.5 * Z * std::log1p(u * u)));
}

// The fitter's tabulated in-plane ISF (isf2dFast, :168-219) on the shared
// table (ddm_isf2d.h): Catmull-Rom along ln Lambda, quadratic Lagrange across
// rows.
double row_eval(const ddm::Isf2dTable& T, int i, double x) {
  const double Zi = T.Z[i];
  const double Lm = x / (Zi + 1);
  if (Lm < ddm::Isf2dTable::lser) return 1 - x * x * (Zi + 2) / (4 * (Zi + 1));
  const double f = (std::log(Lm) - T.u0[i]) / T.du[i];
  if (f >= T.n[i] - 1) return 0;
  const int k = (int)std::floor(f);
  const double t = f - k;
  const double* v = T.v.data() + T.off[i];
  const double p1 = v[k], p2 = v[k + 1], p3 = v[k + 2];
  const double p0 = k > 0 ? v[k - 1] : 2 * p1 - p2;
  return p1 + 0.5 * t * (p2 - p0 + t * (2 * p0 - 5 * p1 + 4 * p2 - p3 + t * (3 * (p1 - p2) + p3 - p0)));
}
double isf2d_fast(double x, double Z) {
  const ddm::Isf2dTable& T = ddm::isf2d_table();
  const double zc = std::min(100.0, std::max(0.5, Z));
  const double zf = (std::log(zc) - T.lz0) / T.dlz;
  const int i0 = std::min(std::max((int)std::floor(zf + 0.5) - 1, 0), ddm::Isf2dTable::nz - 3);
  const double t = zf - i0;
  const double g0 = row_eval(T, i0, x), g1 = row_eval(T, i0 + 1, x), g2 = row_eval(T, i0 + 2, x);
  return 0.5 * (t - 1) * (t - 2) * g0 - t * (t - 2) * g1 + 0.5 * t * (t - 1) * g2;
}

using Kernel = double (*)(double, double);
Kernel swimmer_kernel(int geometry) { return geometry == DDM_GEOMETRY_2D ? isf2d_fast : isf3d; }

// ---------------------------------------------------------------- data

struct Bin {
  int b;
  double q, w;
  std::vector<double> qn, d;
};
struct Prepared {
  std::vector<Bin> bins;
  std::vector<int> lags;
  std::vector<double> tau;
  double dt, tauMax;
  int L;
};
struct RadialView {
  int N, NB, L, K;
  const double *kCenter, *kNodes, *D, *drift;
  const uint32_t* nModes;
};

std::vector<int> lag_grid(int L) {   // lags 1..10, then ~8 % steps (:358-369)
  std::vector<int> lags;
  for (int t = 1; t <= std::min(10, L - 1); t++) lags.push_back(t);
  double x = 10;
  for (;;) {
    x *= 1.08;
    const int t = (int)std::floor(x + 0.5);
    if (t > L - 1) break;
    if (lags.empty() || t > lags.back()) lags.push_back(t);
  }
  return lags;
}

Prepared prepare(const RadialView& r, double dx_um, double fps) {   // (:383-401)
  Prepared pr;
  pr.dt = 1 / fps;
  pr.L = r.L;
  pr.tauMax = (r.L - 1) * pr.dt;
  pr.lags = lag_grid(r.L);
  for (int l : pr.lags) pr.tau.push_back(l * pr.dt);
  const double qs = 2 * M_PI / (r.N * dx_um);
  for (int b = 0; b < r.NB; b++) {
    if (r.nModes[b] < 3) continue;
    Bin bin;
    bin.b = b;
    double scale = -INFINITY;
    bool nan = false;
    for (int l : pr.lags) {
      const double v = r.D[(long)b * r.L + l];
      bin.d.push_back(v);
      if (std::isnan(v)) nan = true;
      scale = std::max(scale, v);
    }
    if (nan || !(scale > 0) || !std::isfinite(scale)) continue;
    bin.q = qs * r.kCenter[b];
    if (r.K > 0)
      for (int k = 0; k < r.K; k++) bin.qn.push_back(r.kNodes[(long)b * r.K + k] * qs);
    else
      bin.qn.push_back(bin.q);
    bin.w = 1 / scale;
    pr.bins.push_back(std::move(bin));
  }
  return pr;
}

struct Params {
  double vbar, Z, Dd, Vmag, alpha;
};
struct Amp {
  double A, B;
};
struct Fit {
  Params p;
  double rss;
  int k;
  std::vector<Amp> amps;
};

// Model column c(tau) = 1 - <E (1 - alpha + alpha g)>_nodes (:417-432)
void model_column(const Prepared& pr, const Bin& bin, const Params& p, Kernel g, bool swim, double* out) {
  const double inv = 1.0 / bin.qn.size();
  const double alpha = swim ? p.alpha : 0;
  for (size_t j = 0; j < pr.tau.size(); j++) {
    const double t = pr.tau[j];
    double f = 0;
    for (double q : bin.qn) {
      double E = std::exp(-p.Dd * q * q * t);
      if (p.Vmag > 0) E *= bessel_j0(q * p.Vmag * t);
      f += alpha > 0 ? E * (1 - alpha + alpha * g(q * p.vbar * t, p.Z)) : E;
    }
    out[j] = 1 - f * inv;
  }
}

// Two-column NNLS on the Gram matrix [[g11, g12], [g12, n]], h = (h1, h2):
// the supports {A}, {B}, {A, B} in that order, each solved by elimination with
// partial pivoting and rejected when singular (pivot < 1e-11 of the largest
// diagonal) or infeasible; the smallest residual wins, the empty support
// (rss = dd) being the start (nnls3 with an empty third column, :238-290).
void nnls2(double g11, double g12, double n, double h1, double h2, double dd, double& A, double& B, double& rss) {
  A = B = 0;
  rss = dd;
  auto take = [&](double a, double b, double r) {
    if (r < rss) {
      A = a;
      B = b;
      rss = std::max(r, 0.0);
    }
  };
  // {A}
  if (!(std::fabs(g11) < 1e-11 * std::max(std::fabs(g11), 1e-300))) {
    const double x = h1 / g11;
    if (!(x < 0)) take(x, 0, dd - x * h1);
  }
  // {B}
  if (!(std::fabs(n) < 1e-11 * std::max(std::fabs(n), 1e-300))) {
    const double x = h2 / n;
    if (!(x < 0)) take(0, x, dd - x * h2);
  }
  // {A, B}
  {
    const double scale = std::max(std::max(std::fabs(g11), std::fabs(n)), 1e-300);
    double m[2][3] = {{g11, g12, h1}, {g12, n, h2}};
    if (std::fabs(m[1][0]) > std::fabs(m[0][0]))
      for (int j = 0; j < 3; j++) std::swap(m[0][j], m[1][j]);
    if (!(std::fabs(m[0][0]) < 1e-11 * scale)) {
      const double f = m[1][0] / m[0][0];
      for (int j = 0; j < 3; j++) m[1][j] -= f * m[0][j];
      if (!(std::fabs(m[1][1]) < 1e-11 * scale)) {
        const double x1 = m[1][2] / m[1][1];
        const double x0 = (m[0][2] - m[0][1] * x1) / m[0][0];
        if (!(x0 < 0) && !(x1 < 0)) {
          double r = dd;
          r -= x0 * h1;
          r -= x1 * h2;
          take(x0, x1, r);
        }
      }
    }
  }
}

// VarPro objective: sum over bins of w^2 times the NNLS residual (:445-472)
double eval_model(const Prepared& pr, const Params& p, Kernel g, bool swim, std::vector<Amp>* amps) {
  const int nt = (int)pr.tau.size();
  std::vector<double> c(nt);
  double total = 0;
  if (amps) amps->clear();
  for (const Bin& bin : pr.bins) {
    model_column(pr, bin, p, g, swim, c.data());
    double g11 = 0, g12 = 0, h1 = 0, h2 = 0, dd = 0;
    for (int j = 0; j < nt; j++) {
      g11 += c[j] * c[j];
      g12 += c[j];
      h1 += c[j] * bin.d[j];
      h2 += bin.d[j];
      dd += bin.d[j] * bin.d[j];
    }
    double A, B, rss;
    nnls2(g11, g12, nt, h1, h2, dd, A, B, rss);
    total += rss * bin.w * bin.w;
    if (amps) amps->push_back({A, B});
  }
  return total;
}

struct Bounds {
  double vmin, vmax, Dmin, Dmax, Zmin, Zmax;
};
Bounds bounds(const Prepared& pr) {   // (:480-491)
  double qmin = INFINITY, qmax = -INFINITY;
  for (const Bin& b : pr.bins) {
    qmin = std::min(qmin, b.q);
    qmax = std::max(qmax, b.q);
  }
  return {std::max(kSwimMin, 0.3 / (qmax * pr.tauMax)), 3 / (qmin * pr.dt), 0.01 / (qmax * qmax * pr.tauMax),
          10 / (qmin * qmin * pr.dt), 0.5, 100};
}
double clampd(double x, double lo, double hi) { return std::min(hi, std::max(lo, x)); }

// Nelder-Mead with the reference's moves and stopping rule (:302-350): the
// simplex is re-sorted stably each iteration; stop when the spread of f is
// below tol (|f_best| + 1e-30).
template <typename F>
double nelder_mead(F&& f, std::vector<double> x0, const std::vector<double>& step, int max_iter, double tol,
                   std::vector<double>& xbest) {
  const int n = (int)x0.size();
  std::vector<std::vector<double>> S(n + 1, x0);
  for (int i = 0; i < n; i++) S[i + 1][i] += step[i];
  std::vector<double> fs(n + 1);
  for (int i = 0; i <= n; i++) fs[i] = f(S[i]);
  std::vector<int> ord(n + 1);
  std::vector<double> c(n), xr(n), xe(n), xc(n);
  for (int it = 0; it < max_iter; it++) {
    for (int i = 0; i <= n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return fs[a] < fs[b]; });
    std::vector<std::vector<double>> S2(n + 1);
    std::vector<double> f2(n + 1);
    for (int i = 0; i <= n; i++) {
      S2[i] = S[ord[i]];
      f2[i] = fs[ord[i]];
    }
    S.swap(S2);
    fs.swap(f2);
    if (std::fabs(fs[n] - fs[0]) <= tol * (std::fabs(fs[0]) + 1e-30)) break;
    std::fill(c.begin(), c.end(), 0.0);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) c[j] += S[i][j] / n;
    auto point = [&](double t, std::vector<double>& out) {
      for (int j = 0; j < n; j++) out[j] = c[j] + t * (S[n][j] - c[j]);
    };
    point(-1, xr);
    const double fr = f(xr);
    if (fr < fs[0]) {
      point(-2, xe);
      const double fe = f(xe);
      if (fe < fr) {
        S[n] = xe;
        fs[n] = fe;
      } else {
        S[n] = xr;
        fs[n] = fr;
      }
    } else if (fr < fs[n - 1]) {
      S[n] = xr;
      fs[n] = fr;
    } else {
      const bool outside = fr < fs[n];
      point(outside ? -0.5 : 0.5, xc);
      const double fc = f(xc);
      if (fc < (outside ? fr : fs[n])) {
        S[n] = xc;
        fs[n] = fc;
      } else {
        for (int i = 1; i <= n; i++) {
          for (int j = 0; j < n; j++) S[i][j] = S[0][j] + 0.5 * (S[i][j] - S[0][j]);
          fs[i] = f(S[i]);
        }
      }
    }
  }
  int ib = 0;
  for (int i = 1; i <= n; i++)
    if (fs[i] < fs[ib]) ib = i;
  xbest = S[ib];
  return fs[ib];
}

Fit finish_fit(const Prepared& pr, const Params& p, Kernel g, bool swim, int k) {
  Fit ft;
  ft.p = p;
  ft.rss = eval_model(pr, p, g, swim, &ft.amps);
  ft.k = k;
  return ft;
}

Fit fit_brownian(const Prepared& pr, double Vmag) {   // (:500-536)
  const Bounds B = bounds(pr);
  auto f = [&](double lnD) { return eval_model(pr, {0, 1, std::exp(lnD), Vmag, 0}, isf3d, false, nullptr); };
  const double lo = std::log(B.Dmin), hi = std::log(B.Dmax);
  const int n = 48;
  std::vector<double> xs(n + 1);
  int bi = 0;
  double bf = INFINITY;
  for (int i = 0; i <= n; i++) {
    xs[i] = lo + (hi - lo) * i / n;
    const double v = f(xs[i]);
    if (v < bf) {
      bf = v;
      bi = i;
    }
  }
  double a = xs[std::max(0, bi - 1)], b = xs[std::min(n, bi + 1)];
  const double gr = (std::sqrt(5.0) - 1) / 2;
  double c = b - gr * (b - a), d = a + gr * (b - a);
  double fc = f(c), fd = f(d);
  for (int it = 0; it < 60 && b - a > 1e-6; it++) {
    if (fc < fd) {
      b = d;
      d = c;
      fd = fc;
      c = b - gr * (b - a);
      fc = f(c);
    } else {
      a = c;
      c = d;
      fc = fd;
      d = a + gr * (b - a);
      fd = f(d);
    }
  }
  const double lnD = fc < fd ? c : d;
  return finish_fit(pr, {0, 1, std::exp(lnD), Vmag, 0}, isf3d, false, 1 + 2 * (int)pr.bins.size());
}

Fit fit_brownian_drift(const Prepared& pr, double vlo, double vhi) {   // (:544-567)
  const Bounds B = bounds(pr);
  const double lDmin = std::log(B.Dmin), lDmax = std::log(B.Dmax);
  auto unpack = [&](const std::vector<double>& u) {
    return Params{0, 1, std::exp(clampd(u[0], lDmin, lDmax)), clampd(u[1], vlo, vhi), 0};
  };
  auto f = [&](const std::vector<double>& u) { return eval_model(pr, unpack(u), isf3d, false, nullptr); };
  std::vector<double> best;
  double bf = 0;
  for (int i = 0; i <= 24; i++) {
    const double lD = lDmin + (lDmax - lDmin) * i / 24;
    for (int j = 0; j <= 6; j++) {
      std::vector<double> u = {lD, vlo + (vhi - vlo) * j / 6};
      const double v = f(u);
      if (best.empty() || v < bf) {
        best = u;
        bf = v;
      }
    }
  }
  std::vector<double> x;
  nelder_mead(f, best, {0.2, std::max((vhi - vlo) / 6, 1e-3)}, 200, 1e-10, x);
  return finish_fit(pr, unpack(x), isf3d, false, 2 + 2 * (int)pr.bins.size());
}

Fit fit_swimmer(const Prepared& pr, int geometry, double Vmag, double Dstart) {   // (:577-611)
  const Bounds B = bounds(pr);
  const Kernel g = swimmer_kernel(geometry);
  const double lv0 = std::log(B.vmin), lv1 = std::log(B.vmax), lz0 = std::log(B.Zmin), lz1 = std::log(B.Zmax);
  const double ld0 = std::log(B.Dmin), ld1 = std::log(B.Dmax);
  auto unpack = [&](const std::vector<double>& u) {
    return Params{std::exp(clampd(u[0], lv0, lv1)), std::exp(clampd(u[1], lz0, lz1)), std::exp(clampd(u[2], ld0, ld1)),
                  Vmag, clampd(u[3], 0, 1)};
  };
  auto f = [&](const std::vector<double>& u) { return eval_model(pr, unpack(u), g, true, nullptr); };
  const double D0 = Dstart > 0 ? clampd(Dstart, B.Dmin, B.Dmax) : std::sqrt(B.Dmin * B.Dmax);
  struct Start {
    std::vector<double> u;
    double f;
  };
  std::vector<Start> starts;
  const int nv = 16;
  for (int i = 0; i < nv; i++) {
    const double lv = lv0 + (lv1 - lv0) * (i + 0.5) / nv;
    for (double lZ : {std::log(1.5), std::log(8.0)})
      for (double a0 : {0.3, 0.8}) {
        std::vector<double> u = {lv, lZ, std::log(D0), a0};
        starts.push_back({u, f(u)});
      }
  }
  std::stable_sort(starts.begin(), starts.end(), [](const Start& a, const Start& b) { return a.f < b.f; });
  std::vector<double> bx;
  double bfv = 0;
  bool have = false;
  for (int s = 0; s < 3 && s < (int)starts.size(); s++) {
    std::vector<double> x;
    const double v = nelder_mead(f, starts[s].u, {0.3, 0.7, 0.7, 0.2}, 400, 1e-10, x);
    if (!have || v < bfv) {
      bx = x;
      bfv = v;
      have = true;
    }
  }
  return finish_fit(pr, unpack(bx), g, true, 4 + 2 * (int)pr.bins.size());
}

// estimateDrift (:671-725) over the record's 8 phase sums per (bin, lag)
ddm_drift_estimate estimate_drift(const RadialView& r, const Prepared& pr, double dx_um, int geometry,
                                  const Params* swimmer) {
  const Kernel g = swimmer_kernel(geometry);
  double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long modes_used = 0;
  int pairs = 0;
  for (const Bin& bin : pr.bins) {
    bool used = false;
    for (int tau = 1; tau < r.L - 1; tau++) {
      if (swimmer && !(std::fabs(g(bin.q * swimmer->vbar * tau * pr.dt, swimmer->Z)) < kDriftIsfMax)) continue;
      const double* o = r.drift + ((long)bin.b * r.L + tau) * 8;
      if (!(o[0] > 0)) continue;
      for (int c = 0; c < 8; c++) S[c] += o[c];
      used = true;
      pairs++;
    }
    if (used) modes_used += r.nModes[bin.b];
  }
  const double Sw = S[0], Sxx = S[1], Sxy = S[2], Syy = S[3], Rx = S[4], Ry = S[5], Rpp = S[6], n = S[7];
  const double det = Sxx * Syy - Sxy * Sxy;
  ddm_drift_estimate d;
  std::memset(&d, 0, sizeof d);
  d.sigma = INFINITY;
  d.pairs = pairs;
  d.samples = n;
  if (n < 10 || !(det > 1e-12 * Sxx * Syy) || !(Sw > 0)) return d;
  const double bx = (Syy * Rx - Sxy * Ry) / det, by = (Sxx * Ry - Sxy * Rx) / det;
  const double rss = std::max(Rpp - bx * Rx - by * Ry, 0.0);
  const double inflate = std::max(1.0, n / std::max<double>(modes_used, 1));
  const double s2 = rss / (n - 2) * inflate;
  const double cxx = s2 * Syy / det, cyy = s2 * Sxx / det, cxy = -s2 * Sxy / det;
  const double conv = -(r.N * dx_um) / (2 * M_PI * pr.dt);
  d.Vx = conv * bx;
  d.Vy = conv * by;
  d.mag = std::hypot(d.Vx, d.Vy);
  double var_along;
  if (d.mag > 0) {
    const double ux = d.Vx / d.mag, uy = d.Vy / d.mag;
    var_along = conv * conv * (ux * ux * cxx + 2 * ux * uy * cxy + uy * uy * cyy);
  } else {
    var_along = conv * conv * (cxx + cyy) / 2;
  }
  d.sigma = std::sqrt(std::max(var_along, 0.0));
  d.angle_deg = std::atan2(d.Vy, d.Vx) * 180 / M_PI;
  d.significant = d.mag > kFlowSigma * d.sigma ? 1 : 0;
  return d;
}

double tsi(const RadialView& r) {   // temporalStructureIndex (:735-749)
  std::vector<double> v;
  for (int b = 0; b < r.NB; b++) {
    if (r.nModes[b] < 3) continue;
    double late = 0;
    int n = 0;
    for (int t = r.L >> 1; t < r.L; t++) {
      late += r.D[(long)b * r.L + t];
      n++;
    }
    late /= n;
    if (!(late > 0)) continue;
    v.push_back((late - r.D[(long)b * r.L + 1]) / late);
  }
  std::stable_sort(v.begin(), v.end(), [](double a, double b) { return a > b; });
  const size_t k = std::min<size_t>(4, v.size());
  if (!k) return 0;
  double s = 0;
  for (size_t i = 0; i < k; i++) s += v[i];
  return std::max(0.0, s / k);
}

double residual_autocorr(const Prepared& pr, const Fit& ft, Kernel g) {   // (:633-651)
  std::vector<Amp> amps;
  eval_model(pr, ft.p, g, true, &amps);
  const int nt = (int)pr.tau.size();
  std::vector<double> c(nt);
  double num = 0, den = 0;
  for (size_t i = 0; i < pr.bins.size(); i++) {
    const Bin& bin = pr.bins[i];
    model_column(pr, bin, ft.p, g, true, c.data());
    double prev = 0;
    for (int j = 0; j < nt; j++) {
      const double res = (bin.d[j] - (amps[i].A * c[j] + amps[i].B)) * bin.w;
      den += res * res;
      if (j > 0) num += prev * res;
      prev = res;
    }
  }
  return den > 0 ? std::min(0.99, std::max(0.0, num / den)) : 0;
}

double bic(double rss, double n, double k, double n_raw) {
  return n * std::log(std::max(rss, 1e-300) / n_raw) + k * std::log(n);
}

}  // namespace

extern "C" {

const char* ddm_verdict_last_error(void) { return g_err.c_str(); }

int ddm_analyze_motility(const ddm_radial_record* rad, const ddm_motility_opts* opt, ddm_motility_result* res,
                         ddm_motility_detail* det) {
  if (!rad || !opt || !res) {
    g_err = "analyzeMotility: null argument";
    return DDM_EINVAL;
  }
  if (rad->N < 2 || rad->NB < 1 || rad->L < 3 || !rad->kCenter || !rad->nModes || !rad->D || !rad->drift ||
      rad->knodes < 0 || (rad->knodes > 0 && !rad->kNodes) || !(opt->dx_um > 0) || !(opt->fps > 0) ||
      (opt->geometry != DDM_GEOMETRY_3D && opt->geometry != DDM_GEOMETRY_2D)) {
    g_err = "analyzeMotility: invalid radial record or options (N >= 2, NB >= 1, L >= 3, drift sums, dxUm, fps > 0)";
    return DDM_EINVAL;
  }
  const RadialView r{rad->N, rad->NB, rad->L, rad->knodes, rad->kCenter, rad->kNodes, rad->D, rad->drift, rad->nModes};
  std::memset(res, 0, sizeof *res);
  const Prepared pr = prepare(r, opt->dx_um, opt->fps);
  const double TSI = tsi(r);
  res->TSI = TSI;
  res->nbins = (int)pr.bins.size();
  res->mu = NAN;
  auto gate = [&](int id, double value, int pass) {
    ddm_gate& g = res->gates[res->ngates++];
    g.id = id;
    g.value = value;
    g.pass = pass;
  };
  const bool g1 = TSI >= kTsiMin;
  gate(1, TSI, g1);
  int verdict;
  if (pr.bins.size() < 4) {
    res->insufficient = 1;
    verdict = DDM_VERDICT_NO_DYNAMICS;
    gate(0, (double)pr.bins.size(), 0);
  } else {
    const Kernel g = swimmer_kernel(opt->geometry);
    const double nData = (double)pr.bins.size() * pr.tau.size();
    // pass 1: no drift in the models
    Fit brown = fit_brownian(pr, 0);
    Fit swim = fit_swimmer(pr, opt->geometry, 0, brown.p.Dd);
    bool has_swim = swim.p.alpha > 0.01;
    ddm_drift_estimate drift = estimate_drift(r, pr, opt->dx_um, opt->geometry, has_swim ? &swim.p : nullptr);
    // pass 2: a significant drift joins the null model, within its measured band
    int drift_params = 0;
    if (drift.significant && drift.mag > 0) {
      const double half = kFlowSigma * drift.sigma + 0.25 * drift.mag;
      brown = fit_brownian_drift(pr, std::max(0.0, drift.mag - half), drift.mag + half);
      swim = fit_swimmer(pr, opt->geometry, brown.p.Vmag, brown.p.Dd);
      drift_params = 1;
      has_swim = swim.p.alpha > 0.01;
      const ddm_drift_estimate d2 = estimate_drift(r, pr, opt->dx_um, opt->geometry, has_swim ? &swim.p : nullptr);
      if (d2.samples >= 10) drift = d2;
    }
    const double rho = residual_autocorr(pr, swim, g);
    const double nEff = std::max(2.0, nData * (1 - rho) / (1 + rho));
    const double dBIC = bic(brown.rss, nEff, brown.k, nData) - bic(swim.rss, nEff, swim.k + drift_params, nData);
    // motile share: alpha is identifiable when a bin resolves the swimmer decay
    int resolved = 0;
    for (const Bin& bin : pr.bins)
      if (bin.q * swim.p.vbar * pr.dt <= 0.5 && bin.q * swim.p.vbar * pr.tauMax >= 3) resolved++;
    // q-scaling of tau_1/2 from the data normalised by the better model's amplitudes
    const Fit& best = dBIC > 0 ? swim : brown;
    std::vector<double> sq, st;
    for (size_t i = 0; i < pr.bins.size(); i++) {
      const Bin& bin = pr.bins[i];
      const double A = best.amps[i].A, Bv = best.amps[i].B;
      if (!(A > 0.2 * (A + Bv))) continue;
      double prev = 1, prevT = 0;
      for (int t = 1; t < r.L; t++) {
        const double f = 1 - (r.D[(long)bin.b * r.L + t] - Bv) / A;
        if (f <= 0.5) {
          double tt;
          if (t == 1) {
            tt = 1;
          } else {
            const double l0 = std::log(prevT), l1 = std::log((double)t);
            tt = std::exp(l0 + (prev - 0.5) / (prev - f) * (l1 - l0));
          }
          sq.push_back(bin.q);
          st.push_back(tt * pr.dt);
          break;
        }
        prev = f;
        prevT = t;
      }
    }
    double mu = NAN;
    if (sq.size() >= 3) {
      double mx = 0, my = 0;
      for (size_t i = 0; i < sq.size(); i++) {
        mx += std::log(sq[i]);
        my += std::log(st[i]);
      }
      mx /= sq.size();
      my /= sq.size();
      double sxy = 0, sxx = 0;
      for (size_t i = 0; i < sq.size(); i++) {
        const double dx = std::log(sq[i]) - mx;
        sxy += dx * (std::log(st[i]) - my);
        sxx += dx * dx;
      }
      if (sxx > 0) mu = -sxy / sxx;
    }
    const bool flow = drift.mag > kFlowSigma * drift.sigma && drift.mag > kFlowMin;
    gate(2, drift.mag, !flow);
    res->gates[res->ngates - 1].sigma = drift.sigma;
    const bool g3 = dBIC > kDbicMin;
    gate(3, dBIC, g3);
    const bool g4 = swim.p.alpha > kAlphaMin && resolved > 0;
    gate(4, swim.p.alpha, g4);
    res->gates[res->ngates - 1].resolved_bins = resolved;
    if (!g1) verdict = DDM_VERDICT_NO_DYNAMICS;
    else if (flow) verdict = DDM_VERDICT_DIRECTED_TRANSPORT;
    else if (!g3 || !g4) verdict = DDM_VERDICT_NOT_MOTILE;
    else verdict = DDM_VERDICT_MOTILE;
    res->alpha = swim.p.alpha;
    res->alpha_resolved = resolved > 0;
    res->resolved_bins = resolved;
    res->vbar = swim.p.vbar;
    res->sigma_v = swim.p.vbar / std::sqrt(swim.p.Z + 1);
    res->Z = swim.p.Z;
    res->Ddiff = swim.p.Dd;
    res->Ddiff_brownian = brown.p.Dd;
    res->dBIC = dBIC;
    res->drift = drift;
    res->mu = mu;
    res->nData = nData;
    res->nEff = nEff;
    res->residual_rho = rho;
    res->swim_rss = swim.rss;
    res->swim_Vmag = swim.p.Vmag;
    res->brown_rss = brown.rss;
    res->brown_Vmag = brown.p.Vmag;
    res->nscaling = (int)sq.size();
    if (det) {
      const int nb = std::min((int)pr.bins.size(), det->cap);
      for (int i = 0; i < nb; i++) {
        const Bin& bin = pr.bins[i];
        if (det->bin) det->bin[i] = bin.b;
        if (det->q) det->q[i] = bin.q;
        if (det->swim_amps) {
          det->swim_amps[2 * i] = swim.amps[i].A;
          det->swim_amps[2 * i + 1] = swim.amps[i].B;
        }
        if (det->brown_amps) {
          det->brown_amps[2 * i] = brown.amps[i].A;
          det->brown_amps[2 * i + 1] = brown.amps[i].B;
        }
      }
      const int ns = std::min((int)sq.size(), det->cap);
      for (int i = 0; i < ns; i++) {
        if (det->scaling_q) det->scaling_q[i] = sq[i];
        if (det->scaling_tau_half) det->scaling_tau_half[i] = st[i];
      }
    }
  }
  gate(5, opt->exposure_locked ? 1 : 0, opt->exposure_locked ? 1 : 0);
  res->physics_verdict = verdict;
  res->verdict = opt->exposure_locked ? verdict : DDM_VERDICT_UNLOCKED_EXPOSURE;
  return DDM_OK;
}

int ddm_lag_grid(int32_t L, int32_t* lags, int32_t cap) {
  if (L < 2 || cap < 0 || (cap > 0 && !lags)) {
    g_err = "lagGrid: L must be >= 2";
    return DDM_EINVAL;
  }
  const std::vector<int> g = lag_grid(L);
  for (int i = 0; i < (int)g.size() && i < cap; i++) lags[i] = g[i];
  return (int)g.size();
}

int ddm_usable_bins(const ddm_radial_record* rad, int32_t* bins, int32_t cap) {
  if (!rad || rad->N < 2 || rad->NB < 1 || rad->L < 2 || !rad->kCenter || !rad->nModes || !rad->D ||
      rad->knodes < 0 || (rad->knodes > 0 && !rad->kNodes) || cap < 0 || (cap > 0 && !bins)) {
    g_err = "prepare: invalid radial record";
    return DDM_EINVAL;
  }
  const RadialView r{rad->N, rad->NB, rad->L, rad->knodes, rad->kCenter, rad->kNodes, rad->D, rad->drift, rad->nModes};
  const Prepared pr = prepare(r, 1.0, 1.0);   // which bins are usable depends on neither scale
  for (int i = 0; i < (int)pr.bins.size() && i < cap; i++) bins[i] = pr.bins[i].b;
  return (int)pr.bins.size();
}

int ddm_swimmer_isf(int32_t geometry, const double* x, int32_t n, double Z, double* g) {
  if ((geometry != DDM_GEOMETRY_3D && geometry != DDM_GEOMETRY_2D) || n < 0 || (n > 0 && (!x || !g)) ||
      !(Z > 0)) {
    g_err = "swimmer ISF: geometry must be 3D or 2D, Z > 0";
    return DDM_EINVAL;
  }
  const Kernel k = swimmer_kernel(geometry);
  for (int i = 0; i < n; i++) g[i] = k(x[i], Z);
  return DDM_OK;
}

int ddm_estimate_drift(const ddm_radial_record* rad, const ddm_motility_opts* opt, const double* swimmer,
                       ddm_drift_estimate* out) {
  if (!rad || !opt || !out || rad->N < 2 || rad->NB < 1 || rad->L < 3 || !rad->kCenter || !rad->nModes ||
      !rad->D || !rad->drift || rad->knodes < 0 || (rad->knodes > 0 && !rad->kNodes) || !(opt->dx_um > 0) ||
      !(opt->fps > 0) || (opt->geometry != DDM_GEOMETRY_3D && opt->geometry != DDM_GEOMETRY_2D)) {
    g_err = "estimateDrift: invalid radial record (drift sums required) or options";
    return DDM_EINVAL;
  }
  const RadialView r{rad->N, rad->NB, rad->L, rad->knodes, rad->kCenter, rad->kNodes, rad->D, rad->drift, rad->nModes};
  const Prepared pr = prepare(r, opt->dx_um, opt->fps);
  Params sw{};
  if (swimmer) {
    sw.vbar = swimmer[0];
    sw.Z = swimmer[1];
  }
  *out = estimate_drift(r, pr, opt->dx_um, opt->geometry, swimmer ? &sw : nullptr);
  return DDM_OK;
}

int ddm_analyze_motility_batch(const ddm_radial_record *r, const ddm_motility_opts *opt, int32_t n,
                               ddm_motility_result *res, ddm_motility_detail *detail, int32_t nthreads) {
  if (n < 0 || (n > 0 && (!r || !opt || !res))) {
    g_err = "analyzeMotility batch: null argument or negative count";
    return DDM_EINVAL;
  }
  int workers = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
  workers = std::min(workers, std::max(1, (int)n));
  std::atomic<int> next(0);
  std::mutex mu;
  int first_rc = DDM_OK;
  std::string first_err;
  auto work = [&] {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) return;
      const int rc = ddm_analyze_motility(r + i, opt + i, res + i, detail ? detail + i : nullptr);
      if (rc != DDM_OK) {
        std::lock_guard<std::mutex> lk(mu);
        if (first_rc == DDM_OK) {
          first_rc = rc;
          first_err = "record " + std::to_string(i) + ": " + g_err;
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; w++) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (first_rc != DDM_OK) g_err = first_err;
  return first_rc;
}

}  // extern "C"
