// Batched per-q Levenberg-Marquardt motility fits (gfx950).
//
// One wavefront per fit (one q-bin of one video); lanes own lag points.  The
// per-bin model is the reference's modelColumn (motility-ddm
// src/motility_fit.js:417-432) with explicit amplitudes:
//   f(tau) = A c(tau) + B,   c = (1/K) sum_i [ -expm1(-D q_i^2 tau) + E_i alpha (1 - g(q_i vbar tau; Z)) ]
// (= 1 - <E (1 - alpha + alpha g)> without the 1 - (~1) cancellation), g the
// Schulz-averaged 3D swimmer ISF (isf3d, :140-144) or the in-plane one
// (the isf2dFast table, :168-219).  chi^2 = w^2 sum (d - f)^2
// with w = 1/max d (:392-398, 469).  Parameters theta = (A, B, alpha, ln vbar,
// ln Z, ln D) (swimmer) or (A, B, ln D) (Brownian), box-constrained by
// projection with the reference's bounds (:480-491).
//
// Precision: the model and Jacobian are FP32, except 1 - g and its
// derivatives at small u = q vbar tau/(Z+1) (< 0.1), where 1 - g cancels in
// FP32 and is evaluated in FP64.  Normal equations are accumulated and solved
// (Cholesky) in FP64.  Marquardt scaling, Nielsen's gain-ratio damping.
//
// Seeding: "plain" = the reference's start grid per bin (49-point log-D
// Brownian grid; 16 vbar x Z{1.5, 8} x alpha{0.3, 0.8} at that D; amplitudes by
// 2-column NNLS), LM from the best `nstarts`; "seeded" = LM from one given
// start (the surrogate's), no grid.
#include "../../include/ddm.h"
#include "ddm_internal.h"
#include "ddm_fft.h"
#include "ddm_isf2d.h"

namespace ddm {

namespace {

// One fit (one wave) per workgroup: a fit's LM steps vary from a few to
// thousands, and a workgroup's slot frees only when its last wave ends; with
// four fits per workgroup three slots idled behind the slowest (131,072
// per-q Schulz fits from their start grids: 697 -> 504 ms).
constexpr int FNT = 64;
constexpr int MAXT = 128;         // lag points per fit
constexpr int MAXK = 8;           // q nodes per bin

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);        // one lane's sum for everybody: identical branches
}

// Wave sums of 2^S values at once by recursive halving: at offset o = 32,
// 16, ... each lane pair keeps one half of the current vector and sends the
// other, so 2^S values cost 2^S - 1 + (6 - S) shuffles instead of 6 each.
// Afterwards lane l holds the total of value k where bit 5 - j of l is bit
// S - 1 - j of k (split_lane); readlane_d broadcasts it (a wave-uniform
// result, the same in every lane).
template <int S>
__device__ __forceinline__ double wave_sum_split(double (&a)[1 << S], int lane) {
#pragma unroll
  for (int j = 0; j < S; j++) {
    const int o = 32 >> j, h = (1 << S) >> (j + 1);
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < h; i++) {
      const double send = up ? a[i] : a[i + h];
      const double keep = up ? a[i + h] : a[i];
      a[i] = keep + __shfl_xor(send, o, 64);
    }
  }
  double v = a[0];
#pragma unroll
  for (int o = 32 >> S; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__host__ __device__ constexpr int split_lane(int k, int S) {
  int l = 0;
  for (int j = 0; j < S; j++) l |= ((k >> (S - 1 - j)) & 1) << (5 - j);
  return l;
}

__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// 1 - g(x; Z) for the 3D Schulz ISF and its derivatives with respect to ln x
// and ln Z (x held fixed).  g = sin(Z atan u) / (Z u (1+u^2)^(Z/2)), u = x/(Z+1).
struct Omg {
  float v, dlx, dlz;
};

__device__ __noinline__ Omg one_minus_isf3d(float xf, float Zf) {
  const float uf = xf / (Zf + 1.f);
  Omg o;
  if (uf < 1e-4f) {
    // series of the reference's small-argument branch: g = 1 - x^2 (Z+2) / (6 (Z+1))
    const double x = xf, Z = Zf, u = x / (Z + 1);
    o.v = (float)(x * x * (Z + 2) / (6 * (Z + 1)));
    o.dlx = (float)(2 * x * x * (Z + 2) / (6 * (Z + 1)));          // d(1-g)/d ln x
    o.dlz = (float)(-Z * u * u / 6);                                 // d(1-g)/d ln Z at fixed x
    return o;
  }
  if (uf < 0.1f) {
    // 1 - g cancels in FP32 here: FP64
    const double Z = Zf, u = (double)xf / (Z + 1);
    const double phi = atan(u), s = sin(Z * phi), cz = cos(Z * phi);
    const double pre = exp(-0.5 * Z * log1p(u * u));
    const double g = s * pre / (Z * u);
    const double udg = pre / Z * (Z * cz / (1 + u * u) - s * (Z * u / (1 + u * u) + 1 / u));
    const double zdgz = pre / u * (phi * cz - s * (0.5 * log1p(u * u) + 1 / Z));
    o.v = (float)(1 - g);
    o.dlx = (float)(-udg);
    o.dlz = (float)(-(zdgz - Z / (Z + 1) * udg));
    return o;
  }
  const float u = uf, Z = Zf;
  const float phi = atanf(u);
  float s, cz;
  sincosf(Z * phi, &s, &cz);
  const float pre = expf(-0.5f * Z * log1pf(u * u));
  const float g = s * pre / (Z * u);
  const float udg = pre / Z * (Z * cz / (1.f + u * u) - s * (Z * u / (1.f + u * u) + 1.f / u));
  const float zdgz = pre / u * (phi * cz - s * (0.5f * log1pf(u * u) + 1.f / Z));
  o.v = 1.f - g;
  o.dlx = -udg;
  o.dlz = -(zdgz - Z / (Z + 1.f) * udg);
  return o;
}

struct FitIn {
  const float* tau;     // [nfits * MAXT]
  const float* d;       // [nfits * MAXT]
  const int* ntau;      // [nfits]
  const float* w;       // [nfits]
  const float* qn;      // [nfits * K]
  const float* bounds;  // [nfits * 4] vmin, vmax, Dmin, Dmax
  const float* seeds;   // [nfits * 6] or null
  const float* vmag;    // [nfits] uniform drift speed |V|, or null (none)
  int K;
  Isf2dDev t2d;         // in-plane swimmer model only
};

// 1 - g of the in-plane swimmer ISF from the reference's table (isf2dFast,
// src/motility_fit.js:168-219), FP64, with d(1-g)/d ln x and d(1-g)/d ln Z at
// fixed x: g = sum_k l_k(t) R_k(ln x - ln(Z_k+1)) over the three rows around
// Z (quadratic Lagrange weights l_k in t = (ln Z - ln Z_0)/dlz - i0), each row
// a Catmull-Rom spline R_k along ln Lambda (derivative: the spline's own), the
// series 1 - x^2 (Z_k+2)/(4(Z_k+1)) below Lambda = 1e-4 and 0 past the row.
// 1 - g = sum_k l_k (1 - R_k) (the weights sum to 1), so it does not cancel.
__device__ __noinline__ double one_minus_isf2d(double x, double Z, const Isf2dDev& T, double& dlx, double& dlz) {
  const double zc = fmin(100.0, fmax(0.5, Z));
  const double zf = (log(zc) - T.lz0) / T.dlz;
  const int i0 = min(max((int)floor(zf + 0.5) - 1, 0), 72 - 3);
  const double t = zf - i0;
  const double l[3] = {0.5 * (t - 1) * (t - 2), -t * (t - 2), 0.5 * t * (t - 1)};
  const double dl[3] = {t - 1.5, 2 - 2 * t, t - 0.5};
  // the interpolant's own slope also at and past the table ends: the LM box
  // holds Z in [0.5, 100], and a zero slope on a bound would pin Z there
  const double dtdz = 1 / T.dlz;
  double om = 0;
  dlx = dlz = 0;
  for (int k = 0; k < 3; k++) {
    const int i = i0 + k;
    const double Zi = T.Z[i], Lm = x / (Zi + 1);
    double omr, dR;   // 1 - R_k, dR_k / d ln x
    if (Lm < 1e-4) {
      omr = x * x * (Zi + 2) / (4 * (Zi + 1));
      dR = -2 * omr;
    } else {
      const double f = (log(Lm) - T.u0[i]) / T.du[i];
      if (f >= T.n[i] - 1) {
        omr = 1;
        dR = 0;
      } else {
        const int j = (int)floor(f);
        const double u = f - j;
        const double* v = T.v + T.off[i];
        const double p1 = v[j], p2 = v[j + 1], p3 = v[j + 2], p0 = j > 0 ? v[j - 1] : 2 * p1 - p2;
        const double c1 = p2 - p0, c2 = 2 * p0 - 5 * p1 + 4 * p2 - p3, c3 = 3 * (p1 - p2) + p3 - p0;
        omr = 1 - (p1 + 0.5 * u * (c1 + u * (c2 + u * c3)));
        dR = (0.5 * c1 + u * c2 + 1.5 * u * u * c3) / T.du[i];
      }
    }
    om += l[k] * omr;
    dlx -= l[k] * dR;
    dlz += dl[k] * dtdz * omr;
  }
  return om;
}

// J0 by the reference's rational approximation (besselJ0,
// src/motility_fit.js:54-68: a ratio of two degree-5 polynomials in x^2 below
// |x| = 8, the two-term asymptotic form above; absolute error < 1e-8), FP64.
// The coefficients are the reference's: the drift factor must be the same
// function.
__device__ __noinline__ double bessel_j0(double x) {
  const double ax = fabs(x);
  if (ax < 8) {
    const double y = x * x;
    const double p = 57568490574.0 +
                     y * (-13362590354.0 + y * (651619640.7 + y * (-11214424.18 + y * (77392.33017 + y * -184.9052456))));
    const double r = 57568490411.0 + y * (1029532985.0 + y * (9494680.718 + y * (59272.64853 + y * (267.8532712 + y))));
    return p / r;
  }
  const double z = 8 / ax, y = z * z, xx = ax - 0.785398164;
  const double p = 1 + y * (-0.1098628627e-2 + y * (0.2734510407e-4 + y * (-0.2073370639e-5 + y * 0.2093887211e-6)));
  const double r =
      -0.1562499995e-1 + y * (0.1430488765e-3 + y * (-0.6911147651e-5 + y * (0.7621095161e-6 - y * 0.934935152e-7)));
  return sqrt(0.636619772 / ax) * (cos(xx) * p - z * sin(xx) * r);
}

// Column value (and derivatives w.r.t. alpha, ln vbar, ln Z, ln D) at one lag;
// GEO 0: the 3D swimmer ISF, 1: the in-plane one (table T).  V: a uniform
// drift speed, E = exp(-D q^2 tau) J0(q V tau) (modelColumn), with
// 1 - E = -expm1(-s) + exp(-s) (1 - J0) so the small-lag values do not cancel.
template <int GEO>
__device__ __forceinline__ void column(float t, const float* q, int K, float alpha, float vbar, float Z, float Dd,
                                       float V, const Isf2dDev& T, bool swim, bool deriv, float& c, float& da,
                                       float& dlv, float& dlz, float& dld) {
  c = da = dlv = dlz = dld = 0.f;
#pragma unroll
  for (int i = 0; i < MAXK; i++) {
    if (i >= K) break;
    const float qi = q[i];
    const float s = Dd * qi * qi * t;
    float E = expf(-s), omE = -expm1f(-s);
    if (V > 0.f) {
      const double J = bessel_j0((double)qi * V * t);
      omE += (float)((double)E * (1.0 - J));
      E = (float)(E * J);
    }
    Omg og = {0.f, 0.f, 0.f};
    if (swim) {
      if constexpr (GEO == 1) {
        double gx, gz;
        og.v = (float)one_minus_isf2d((double)qi * vbar * t, Z, T, gx, gz);
        og.dlx = (float)gx;
        og.dlz = (float)gz;
      } else {
        og = one_minus_isf3d(qi * vbar * t, Z);
      }
    }
    c += omE + E * alpha * og.v;
    if (deriv) {
      da += E * og.v;
      dlv += E * alpha * og.dlx;
      dlz += E * alpha * og.dlz;
      dld += s * E * (1.f - alpha * og.v);
    }
  }
  const float inv = 1.f / K;
  c *= inv;
  da *= inv;
  dlv *= inv;
  dlz *= inv;
  dld *= inv;
}

// 1 - g(x; Z) in FP64 (value only), the series below u = 1e-4 as above.
__device__ __noinline__ double one_minus_isf3d_d(double x, double Z) {
  const double u = x / (Z + 1);
  if (u < 1e-4) return x * x * (Z + 2) / (6 * (Z + 1));
  const double phi = atan(u);
  return 1 - sin(Z * phi) * exp(-0.5 * Z * log1p(u * u)) / (Z * u);
}

// 1 - g(x; Z) in FP64 with its derivatives with respect to ln x and ln Z
// (x held fixed), the series below u = 1e-4 as above.
__device__ __noinline__ double one_minus_isf3d_dd(double x, double Z, double& dlx, double& dlz) {
  const double u = x / (Z + 1);
  if (u < 1e-4) {
    dlx = 2 * x * x * (Z + 2) / (6 * (Z + 1));
    dlz = -Z * u * u / 6;
    return x * x * (Z + 2) / (6 * (Z + 1));
  }
  const double phi = atan(u), s = sin(Z * phi), cz = cos(Z * phi);
  const double pre = exp(-0.5 * Z * log1p(u * u));
  const double udg = pre / Z * (Z * cz / (1 + u * u) - s * (Z * u / (1 + u * u) + 1 / u));
  const double zdgz = pre / u * (phi * cz - s * (0.5 * log1p(u * u) + 1 / Z));
  dlx = -udg;
  dlz = -(zdgz - Z / (Z + 1) * udg);
  return 1 - s * pre / (Z * u);
}

// The per-q LM's Jacobian column at one lag in FP64: the value c and its
// derivatives with respect to alpha, ln vbar, ln Z and ln D.  (With an FP32
// Jacobian the LM stalled in slow valleys, 2-4 % of the benchmark's bins
// stopping 1-15 % above the FP64 LM's chi^2 from the same start: its steps
// lost the valley direction below the FP32 derivatives' rounding.)
template <int GEO>
__device__ __forceinline__ void column_jd(double t, const float* q, int K, double alpha, double vbar, double Z,
                                          double Dd, double V, const Isf2dDev& T, bool swim, double& c, double& da,
                                          double& dlv, double& dlz, double& dld) {
  c = da = dlv = dlz = dld = 0;
#pragma unroll
  for (int i = 0; i < MAXK; i++) {
    if (i >= K) break;
    const double qi = q[i];
    const double s = Dd * qi * qi * t;
    double E = exp(-s), omE = -expm1(-s);
    if (V > 0) {
      const double J = bessel_j0(qi * V * t);
      omE += E * (1 - J);
      E *= J;
    }
    double ov = 0, ox = 0, oz = 0;
    if (swim) ov = GEO == 1 ? one_minus_isf2d(qi * vbar * t, Z, T, ox, oz) : one_minus_isf3d_dd(qi * vbar * t, Z, ox, oz);
    c += omE + E * alpha * ov;
    da += E * ov;
    dlv += E * alpha * ox;
    dlz += E * alpha * oz;
    dld += s * E * (1 - alpha * ov);
  }
  const double inv = 1.0 / K;
  c *= inv;
  da *= inv;
  dlv *= inv;
  dlz *= inv;
  dld *= inv;
}

// Column value at one lag in FP64 (the objective).
template <int GEO>
__device__ double column_d(double t, const float* q, int K, double alpha, double vbar, double Z, double Dd, double V,
                           const Isf2dDev& T, bool swim) {
  double c = 0;
#pragma unroll
  for (int i = 0; i < MAXK; i++) {
    if (i >= K) break;
    const double qi = q[i];
    const double s = Dd * qi * qi * t;
    double og = 0.0, gx, gz;
    if (swim) og = GEO == 1 ? one_minus_isf2d(qi * vbar * t, Z, T, gx, gz) : one_minus_isf3d_d(qi * vbar * t, Z);
    double E = exp(-s), omE = -expm1(-s);
    if (V > 0) {
      const double J = bessel_j0(qi * V * t);
      omE += E * (1 - J);
      E *= J;
    }
    c += omE + E * alpha * og;
  }
  return c / K;
}

// chi^2 at theta with the model in FP64 and an FP64 sum over the wave: the
// objective is resolved far below the 1e-6 relative level at which fits are
// compared, so LM accept/reject and termination decisions are not driven by
// FP32 rounding of the model.
// The lane's weighted residuals at theta go to res (they also form the LM's
// gradient, so that the gradient is the one of the FP64 objective: with an
// FP32 model value in the residual the LM would settle on the FP32 model's
// optimum, measurably above the FP64 one in slow valleys).
// The Jacobian columns at theta come from the same pass (the model's
// transcendentals are shared): jac[(5 r + k) * 64 + lane] holds the
// lane's lag slot r, k = 0..4 for c, dc/dalpha, dc/dln vbar, dc/dln Z,
// dc/dln D (a per-wave LDS buffer, so a trial that is accepted hands the LM
// its next Jacobian without a second model evaluation).
template <int GEO>
__device__ __noinline__ double chi2_jac_at(const double* th, const float* tv, const float* dv, int nt, float w,
                                           const float* q, int K, float V, const Isf2dDev& T, bool swim, int lane,
                                           double (&res)[MAXT / 64], double* __restrict__ jac) {
  double s = 0;
  const double alpha = swim ? th[2] : 0.0, vbar = exp(th[3]), Z = exp(th[4]), Dd = exp(th[5]);
#pragma unroll
  for (int r = 0; r < MAXT / 64; r++) {
    const int j = lane + 64 * r;
    res[r] = 0;
    if (j < nt) {
      double c, da, dlv, dlz, dld;
      column_jd<GEO>(tv[r], q, K, alpha, vbar, Z, Dd, V, T, swim, c, da, dlv, dlz, dld);
      res[r] = (double)w * ((double)dv[r] - (th[0] * c + th[1]));
      s += res[r] * res[r];
      double* jr = jac + 5 * 64 * r + lane;
      jr[0] = c;
      jr[64] = da;
      jr[128] = dlv;
      jr[192] = dlz;
      jr[256] = dld;
    }
  }
  return wave_sum_d(s);
}

// Exact 2-column NNLS (A, B >= 0) for a fixed column c, from the wave sums
// g11 = sum c^2, g12 = sum c, n, h1 = sum c d, h2 = sum d, dd = sum d^2.
__device__ double nnls2(double g11, double g12, double n, double h1, double h2, double dd, float& A, float& B) {
  double best = dd, bA = 0, bB = 0;
  if (g11 > 0) {  // A only
    const double a = h1 / g11;
    const double r = dd - a * h1;
    if (a >= 0 && r < best) {
      best = r;
      bA = a;
      bB = 0;
    }
  }
  if (n > 0) {  // B only
    const double b = h2 / n;
    const double r = dd - b * h2;
    if (b >= 0 && r < best) {
      best = r;
      bA = 0;
      bB = b;
    }
  }
  const double det = g11 * n - g12 * g12;
  if (det > 1e-11 * fmax(g11, n) * fmax(g11, n)) {
    const double a = (h1 * n - g12 * h2) / det, b = (g11 * h2 - g12 * h1) / det;
    const double r = dd - a * h1 - b * h2;
    if (a >= 0 && b >= 0 && r < best) {
      best = r;
      bA = a;
      bB = b;
    }
  }
  A = (float)bA;
  B = (float)bB;
  return best > 0 ? best : 0;
}

// One start-grid point: the model column in FP64 (as the objective), so the
// grid's ranking -- which starts the LM runs from -- is not decided by FP32
// rounding of near-equal grid chi^2 values.
template <int GEO>
__device__ double grid_point(const float* tv, const float* dv, int nt, const float* q, int K, double alpha, double vbar,
                             double Z, double Dd, float V, const Isf2dDev& T, bool swim, int lane, float& A, float& B) {
  double g11 = 0, g12 = 0, h1 = 0, h2 = 0, dd = 0;
#pragma unroll
  for (int r = 0; r < MAXT / 64; r++) {
    const int j = lane + 64 * r;
    if (j < nt) {
      const double c = column_d<GEO>(tv[r], q, K, alpha, vbar, Z, Dd, V, T, swim);
      g11 += c * c;
      g12 += c;
      h1 += (double)c * dv[r];
      h2 += dv[r];
      dd += (double)dv[r] * dv[r];
    }
  }
  double a[8] = {g11, g12, h1, h2, dd, 0.0, 0.0, 0.0};
  const double t = wave_sum_split<3>(a, lane);
  g11 = readlane_d(t, split_lane(0, 3));
  g12 = readlane_d(t, split_lane(1, 3));
  h1 = readlane_d(t, split_lane(2, 3));
  h2 = readlane_d(t, split_lane(3, 3));
  dd = readlane_d(t, split_lane(4, 3));
  return nnls2(g11, g12, (double)nt, h1, h2, dd, A, B);
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }
__device__ __forceinline__ double clampd(double x, float lo, float hi) {
  return fmin((double)hi, fmax((double)lo, x));
}

// LM from theta (in/out); returns chi^2, adds the steps taken to *steps.
template <bool SWIM, int GEO>
__device__ double lm_run(double* th, const float* lo, const float* hi, const float* tv, const float* dv, int nt, float w,
                         const float* q, int K, float V, const Isf2dDev& T, int maxit, int lane, int* steps,
                         double* jbuf) {
  constexpr bool swim = SWIM;
  constexpr int p = SWIM ? 6 : 3;
  constexpr int idx[6] = {0, 1, SWIM ? 2 : 5, 3, 4, 5};
  for (int k = 0; k < 6; k++) th[k] = clampd(th[k], lo[k], hi[k]);
  double rcur[MAXT / 64], rtry[MAXT / 64];   // FP64 residuals at theta and at the trial point
  // jcur: the Jacobian at theta, jtry: at the trial point (two halves of the wave's LDS buffer)
  double* jcur = jbuf;
  double* jtry = jbuf + 5 * MAXT;
  double chi = chi2_jac_at<GEO>(th, tv, dv, nt, w, q, K, V, T, swim, lane, rcur, jcur);
  double lam = 1e-3, nu = 2;
  int it = 0;
  bool done = false;
  while (it < maxit && !done) {
    double H[21], g[6];
#pragma unroll
    for (int k = 0; k < 21; k++) H[k] = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) g[k] = 0;
    // the Jacobian in FP64 at theta, from the evaluation that accepted it
#pragma unroll
    for (int r = 0; r < MAXT / 64; r++) {
      const int j = lane + 64 * r;
      if (j < nt) {
        const double* jr = jcur + 5 * 64 * r + lane;
        const double c = jr[0], da = jr[64], dlv = jr[128], dlz = jr[192], dld = jr[256];
        double F[6];
        F[0] = (double)w * c;
        F[1] = w;
        F[2] = (double)w * th[0] * da;
        F[3] = (double)w * th[0] * dlv;
        F[4] = (double)w * th[0] * dlz;
        F[5] = (double)w * th[0] * dld;
        const double res = rcur[r];
        int h = 0;
#pragma unroll
        for (int a = 0; a < 6; a++) {
          if (a < p) {
            g[a] += F[idx[a]] * res;
#pragma unroll
            for (int b = 0; b <= a; b++) H[h++] += F[idx[a]] * F[idx[b]];
          }
        }
      }
    }
    // the p(p+1)/2 packed entries of H and the p of g in one split reduction
    {
      constexpr int NH = p * (p + 1) / 2, NV = NH + p, S = NV > 16 ? 5 : 4;
      double a[1 << S];
#pragma unroll
      for (int k = 0; k < (1 << S); k++) a[k] = k < NH ? H[k] : (k < NV ? g[k - NH] : 0.0);
      const double t = wave_sum_split<S>(a, lane);
#pragma unroll
      for (int k = 0; k < NH; k++) H[k] = readlane_d(t, split_lane(k, S));
#pragma unroll
      for (int k = 0; k < p; k++) g[k] = readlane_d(t, split_lane(NH + k, S));
    }
    // Active set: a parameter on a bound whose gradient points out of the box
    // is held for this step (its row and column drop out of the system).
    // Marquardt scaling with a floor of 1e-12 x the largest diagonal, so a
    // parameter the data barely see cannot take an unbounded step.
    bool frozen[6];
    double diag[6], dmax = 0;
#pragma unroll
    for (int a = 0; a < p; a++) {
      const int k = idx[a];
      frozen[a] = (th[k] <= lo[k] && g[a] < 0) || (th[k] >= hi[k] && g[a] > 0);
      const double Haa = H[a * (a + 1) / 2 + a];
      dmax = Haa > dmax ? Haa : dmax;
    }
#pragma unroll
    for (int a = 0; a < p; a++) {
      const double Haa = H[a * (a + 1) / 2 + a], fl = fmax(1e-12 * dmax, 1e-300);
      diag[a] = Haa > fl ? Haa : fl;
      if (frozen[a]) g[a] = 0;
    }
    for (;;) {
      it++;
      // Cholesky of H + lam diag, packed lower triangle, FP64
      double Lc[21];
      int h = 0;
#pragma unroll
      for (int a = 0; a < p; a++)
#pragma unroll
        for (int b = 0; b <= a; b++, h++) {
          if (frozen[a] || frozen[b]) Lc[h] = a == b ? 1.0 : 0.0;
          else Lc[h] = H[h] + (a == b ? lam * diag[a] : 0.0);
        }
      bool ok = true;
#pragma unroll
      for (int a = 0; a < p; a++) {
        const int ra = a * (a + 1) / 2;
#pragma unroll
        for (int b = 0; b <= a; b++) {
          const int rb = b * (b + 1) / 2;
          double s = Lc[ra + b];
#pragma unroll
          for (int k = 0; k < b; k++) s -= Lc[ra + k] * Lc[rb + k];
          if (a == b) {
            ok = ok && (s > 0);
            Lc[ra + a] = sqrt(s > 0 ? s : 1.0);
          } else {
            Lc[ra + b] = s / Lc[rb + b];
          }
        }
      }
      if (ok) {
        double y[6], x[6];
#pragma unroll
        for (int a = 0; a < p; a++) {
          const int ra = a * (a + 1) / 2;
          double s = g[a];
#pragma unroll
          for (int k = 0; k < a; k++) s -= Lc[ra + k] * y[k];
          y[a] = s / Lc[ra + a];
        }
#pragma unroll
        for (int a = p - 1; a >= 0; a--) {
          double s = y[a];
#pragma unroll
          for (int k = a + 1; k < p; k++) s -= Lc[k * (k + 1) / 2 + a] * x[k];
          x[a] = s / Lc[a * (a + 1) / 2 + a];
        }
        double nth[6];
#pragma unroll
        for (int k = 0; k < 6; k++) nth[k] = th[k];
        double pred = 0;
#pragma unroll
        for (int a = 0; a < p; a++) {
          const int k = idx[a];
          nth[k] = clampd(th[k] + x[a], lo[k], hi[k]);
          const double dk = nth[k] - th[k];
          pred += dk * (lam * diag[a] * dk + g[a]);
        }
        const double nchi = chi2_jac_at<GEO>(nth, tv, dv, nt, w, q, K, V, T, swim, lane, rtry, jtry);
        const double rho = pred > 0 ? (chi - nchi) / pred : -1;
        if (nchi < chi && rho > 0) {
          const double rel = (chi - nchi) / (chi > 1e-300 ? chi : 1e-300);
          for (int k = 0; k < 6; k++) th[k] = nth[k];
#pragma unroll
          for (int r = 0; r < MAXT / 64; r++) rcur[r] = rtry[r];
          double* jt = jcur;
          jcur = jtry;
          jtry = jt;
          chi = nchi;
          const double f = 1 - (2 * rho - 1) * (2 * rho - 1) * (2 * rho - 1);
          lam *= f > 1.0 / 3 ? f : 1.0 / 3;
          if (lam < 1e-15) lam = 1e-15;
          nu = 2;
          // converged: the actual and the predicted relative decrease are both
          // below 1e-9 (MINPACK's ftol test), or the decrease vanished
          if ((rel < 1e-9 && pred < 1e-9 * chi) || rel < 1e-12 || chi < 1e-300) done = true;
          break;
        }
      }
      lam *= nu;
      nu *= 2;
      if (lam > 1e15 || it >= maxit) {
        done = true;
        break;
      }
    }
  }
  *steps += it;
  return chi;
}

}  // namespace

// A fit's intermediate results: its starts (at most LM_NS, ranked) and each
// start's LM result.  A grid-seeded fit runs its starts on separate waves: LM
// from one start takes up to maxit steps, and with the starts run one after
// another a batch of fewer fits than the device holds waited on its slowest
// fit's three runs in a row.
constexpr int LM_NS = 3;
struct LmWork {
  double* chi;     // [nfits][LM_NS]
  float* starts;   // [nfits][LM_NS][6]
  float* theta;    // [nfits][LM_NS][6]
  int* nrun;       // [nfits]
  int* steps;      // [nfits][LM_NS]
};
LmWork lm_work(void* ws, int nfits) {
  const size_t n = (size_t)nfits;
  char* p = static_cast<char*>(ws);
  LmWork w;
  w.chi = reinterpret_cast<double*>(p);
  w.starts = reinterpret_cast<float*>(p + n * LM_NS * 8);
  w.theta = w.starts + n * LM_NS * 6;
  w.nrun = reinterpret_cast<int*>(w.theta + n * LM_NS * 6);
  w.steps = w.nrun + n;
  return w;
}

// mode 0: LM from the seeds; 1: plain grid-seeded LM from the best `nstarts`
// starts; 2: nonlinear parameters from the seeds, A and B by NNLS there.
// Three launches: RUN = false sets each fit's starts (ranks the start grid in
// mode 1; one wave per fit), RUN = true runs LM from one start (one wave per
// fit and start), k_lm_pick keeps each fit's best start.  Without the start
// grid's code beside it the LM spills 35-53 registers instead of 264-285
// (131,072 plain per-q fits 502 -> 343 ms).
template <bool SWIM, int GEO, bool RUN>
// Three waves per SIMD: the FP64 normal equations and Cholesky would keep
// ~256 VGPRs live, and the kernel is latency-bound; capped at 168 VGPRs (the
// rest spills) it is fastest (10,240 Schulz fits: 52 ms at one wave per SIMD,
// 32 ms at two, 29 ms at three, 30 ms at four).
__global__ __launch_bounds__(FNT, 3) void k_lm_fit(FitIn in, int nfits, int mode, int nstarts, int maxit,
                                                LmWork ws) {
  const int lane = threadIdx.x & 63;
  const int f = RUN ? (int)blockIdx.x / nstarts : (int)blockIdx.x;
  if (f >= nfits) return;
  const int s_own = RUN ? (int)blockIdx.x % nstarts : 0;
  if (RUN && s_own >= ws.nrun[f]) return;
  const int nt = min(in.ntau[f], MAXT);
  const int K = min(in.K, MAXK);
  float tv[MAXT / 64], dv[MAXT / 64], q[MAXK];
#pragma unroll
  for (int r = 0; r < MAXT / 64; r++) {
    const int j = lane + 64 * r;
    tv[r] = j < nt ? in.tau[(long)f * MAXT + j] : 0.f;
    dv[r] = j < nt ? in.d[(long)f * MAXT + j] : 0.f;
  }
#pragma unroll
  for (int i = 0; i < MAXK; i++) q[i] = i < K ? in.qn[(long)f * in.K + i] : 0.f;
  const float w = in.w[f];
  const float V = in.vmag ? in.vmag[f] : 0.f;
  const float vmin = in.bounds[4 * f], vmax = in.bounds[4 * f + 1], Dmin = in.bounds[4 * f + 2],
              Dmax = in.bounds[4 * f + 3];
  const float lo[6] = {0.f, 0.f, 0.f, logf(vmin), logf(0.5f), logf(Dmin)};
  const float hi[6] = {3.4e38f, 3.4e38f, 1.f, logf(vmax), logf(100.f), logf(Dmax)};
  int steps = 0;
  if constexpr (RUN) {
    // the wave's Jacobian columns at theta and at the trial point (FP64, lane-indexed)
    __shared__ double lm_jac[2 * 5 * MAXT];
    double th[6];
#pragma unroll
    for (int k = 0; k < 6; k++) th[k] = ws.starts[((long)f * LM_NS + s_own) * 6 + k];
    const double chi = lm_run<SWIM, GEO>(th, lo, hi, tv, dv, nt, w, q, K, V, in.t2d, maxit, lane, &steps, lm_jac);
    if (lane == 0) {
      const long o = (long)f * LM_NS + s_own;
      for (int k = 0; k < 6; k++) ws.theta[o * 6 + k] = (float)th[k];
      ws.chi[o] = chi;
      ws.steps[o] = steps;
    }
    return;
  }
  float st[3][6];
  int nrun = 1;
  if (mode != 1) {
#pragma unroll
    for (int k = 0; k < 6; k++) st[0][k] = in.seeds[(long)f * 6 + k];
    if (mode == 2) {
      st[0][2] = fminf(1.f, fmaxf(0.f, st[0][2]));
#pragma unroll
      for (int k = 3; k < 6; k++) st[0][k] = fminf(hi[k], fmaxf(lo[k], st[0][k]));
      float A, B;
      grid_point<GEO>(tv, dv, nt, q, K, SWIM ? st[0][2] : 0.0, exp((double)st[0][3]), exp((double)st[0][4]),
                      exp((double)st[0][5]), V, in.t2d, SWIM, lane, A, B);
      st[0][0] = A;
      st[0][1] = B;
    }
  } else {
    // Brownian 49-point log-D grid (grid coordinates in FP64, as the ranking)
    const double lDmin = log((double)Dmin), lDmax = log((double)Dmax), lvmin = log((double)vmin),
                 lvmax = log((double)vmax);
    float bA = 0, bB = 0;
    double blD = lDmin;
    double bchi = 1e300;
    for (int i = 0; i <= 48; i++) {
      const double lD = lDmin + ((lDmax - lDmin) * i) / 48;
      float A, B;
      const double r = grid_point<GEO>(tv, dv, nt, q, K, 0.0, 1.0, 1.0, exp(lD), V, in.t2d, false, lane, A, B);
      if (r < bchi) {
        bchi = r;
        bA = A;
        bB = B;
        blD = lD;
      }
    }
    if (!SWIM) {
      st[0][0] = bA;
      st[0][1] = bB;
      st[0][2] = 0.f;
      st[0][3] = 0.5f * (lo[3] + hi[3]);
      st[0][4] = logf(4.f);
      st[0][5] = (float)blD;
    } else {
      double sf[3] = {1e300, 1e300, 1e300};
      const double lZs[2] = {log(1.5), log(8.0)}, as[2] = {0.3, 0.8};
      for (int i = 0; i < 16; i++) {
        const double lv = lvmin + ((lvmax - lvmin) * (i + 0.5)) / 16;
#pragma unroll
        for (int z = 0; z < 2; z++)
#pragma unroll
          for (int a = 0; a < 2; a++) {
            float A, B;
            const double r =
                grid_point<GEO>(tv, dv, nt, q, K, as[a], exp(lv), exp(lZs[z]), exp(blD), V, in.t2d, true, lane, A,
                                B) *
                w * w;
            // keep the best 3 (stable: earlier grid points win ties); an
            // insertion step with static register indices
            float cand[6] = {A, B, (float)as[a], (float)lv, (float)lZs[z], (float)blD};
            double cr = r;
#pragma unroll
            for (int k = 0; k < 3; k++) {
              if (cr < sf[k]) {
                const double tf = sf[k];
                sf[k] = cr;
                cr = tf;
#pragma unroll
                for (int m = 0; m < 6; m++) {
                  const float tm = st[k][m];
                  st[k][m] = cand[m];
                  cand[m] = tm;
                }
              }
            }
          }
      }
      nrun = min(nstarts, 3);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < 3; s++)
      if (s < nrun)
#pragma unroll
        for (int k = 0; k < 6; k++) ws.starts[((long)f * LM_NS + s) * 6 + k] = st[s][k];
    ws.nrun[f] = nrun;
  }
}

// Each fit's best start (the first of equal chi^2 in rank order) and its LM
// steps over all its starts.
__global__ __launch_bounds__(256) void k_lm_pick(LmWork ws, int nfits, float* __restrict__ theta_out,
                                                 double* __restrict__ chi_out, int* __restrict__ steps_out) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nfits) return;
  const int nr = ws.nrun[f];
  double best = 1e300;
  int bs = 0, steps = 0;
  for (int s = 0; s < nr; s++) {
    const double c = ws.chi[(long)f * LM_NS + s];
    steps += ws.steps[(long)f * LM_NS + s];
    if (c < best) {
      best = c;
      bs = s;
    }
  }
  for (int k = 0; k < 6; k++) theta_out[(long)f * 6 + k] = ws.theta[((long)f * LM_NS + bs) * 6 + k];
  chi_out[f] = best;
  steps_out[f] = steps;
}

// ----------------------------------------------------------------- in-plane ISF, FP64 quadrature
//
// The Schulz-averaged in-plane swimmer ISF g(x; Z) = P_Z(1/sqrt(1+L^2)) /
// (1+L^2)^((Z+1)/2), L = x/(Z+1) (isf2d, motility-ddm src/motility_fit.js:
// 152-159), with the Legendre function from Laplace's first integral
// P_nu(cos t) = (1/pi) int_0^pi (cos t + i sin t cos phi)^nu dphi by the
// trapezoid rule on 16 + ceil(nu t) panels (legendreP, :120-132): the
// integrand is smooth, even and periodic in phi, so the rule converges
// geometrically, and the phase nu arg(.) spans at most nu t.  In FP64
// throughout: near L -> 0 the terms cancel to 1 - O(L^2).  The device builds
// the fitter's isf2dFast table from it (every node of every row) and
// evaluates it for ddm_isf2d_eval.
__device__ double legendre_p_d(double nu, double c) {
  const double s = sqrt(fmax(0.0, 1 - c * c));
  const int n = 16 + (int)ceil(nu * atan2(s, c));
  double acc = 0;
  for (int i = 0; i <= n; i++) {
    const double im = s * cos(M_PI * i / n);
    const double val = exp(0.5 * nu * log(c * c + im * im)) * cos(nu * atan2(im, c));
    acc += (i == 0 || i == n) ? 0.5 * val : val;
  }
  return acc / n;
}

// g at L = x/(Z+1) >= 1e-4 (a table node); 0 once the envelope is below e^-25
__device__ double isf2d_lambda_d(double Lm, double Z) {
  const double u = 1 + Lm * Lm;
  const double ln_pre = -0.5 * (Z + 1) * log(u);
  if (ln_pre < -25) return 0;
  return legendre_p_d(Z, 1 / sqrt(u)) * exp(ln_pre);
}

__device__ double isf2d_exact_d(double x, double Z) {
  const double Lm = x / (Z + 1);
  if (Lm < 1e-4) return 1 - x * x * (Z + 2) / (4 * (Z + 1));
  return isf2d_lambda_d(Lm, Z);
}

// One thread per table node: v[off[i] + k] = g(exp(u0[i] + k du[i]); Z[i]) for
// k < n[i] + 2, the rows stored back to back (row i found by bisection on off).
__global__ __launch_bounds__(256) void k_isf2d_nodes(const double* __restrict__ Z, const double* __restrict__ u0,
                                                     const double* __restrict__ du, const int* __restrict__ off,
                                                     int nz, int total, double* __restrict__ v) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = nz - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  v[g] = isf2d_lambda_d(exp(u0[lo] + (g - off[lo]) * du[lo]), Z[lo]);
}

// exact = 1: isf2d by the quadrature; 0: isf2dFast, the table lookup the
// in-plane fits use (Catmull-Rom along ln L, quadratic Lagrange across Z)
__global__ __launch_bounds__(256) void k_isf2d_eval(const double* __restrict__ x, const double* __restrict__ Z, int n,
                                                    int exact, Isf2dDev T, double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double a, b;
  out[i] = exact ? isf2d_exact_d(x[i], Z[i]) : 1 - one_minus_isf2d(x[i], Z[i], T, a, b);
}

// The table of ddm_isf2d.h built on the device: its layout (Z, u0, du, off, n
// of each row) copied from the host, its values computed by k_isf2d_nodes.
// Synchronous (the caller's next launch may read it from another stream).
hipError_t build_isf2d_device(const Isf2dTable& T, double* v, double* z, double* u0, double* du, int* off, int* n,
                              hipStream_t st) {
  const size_t nz = T.Z.size();
  const int total = (int)T.v.size();
  hipError_t e;
  if ((e = hipMemcpyAsync(z, T.Z.data(), nz * 8, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(u0, T.u0.data(), nz * 8, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(du, T.du.data(), nz * 8, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(off, T.off.data(), nz * 4, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(n, T.n.data(), nz * 4, hipMemcpyHostToDevice, st)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_isf2d_nodes, dim3((total + 255) / 256), dim3(256), 0, st, z, u0, du, off, (int)nz, total, v);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

hipError_t launch_isf2d_eval(const double* x, const double* Z, int n, int exact, const Isf2dDev* T, double* out,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_isf2d_eval, dim3((n + 255) / 256), dim3(256), 0, st, x, Z, n, exact, *T, out);
  return hipGetLastError();
}

// ----------------------------------------------------------------- global-parameter fits
//
// One workgroup per video: (alpha, ln vbar, ln Z, ln D) shared by all its
// bins, the per-bin amplitudes A_b, B_b >= 0 eliminated exactly by 2-column
// NNLS at every evaluation -- the reference's variable projection (evalModel,
// motility-ddm src/motility_fit.js:450-474) and its objective RSS = sum_b
// w_b^2 rss_b, here with an FP64 model.  The nonlinear parameters take LM
// steps on the reduced Gauss-Newton matrix of the projected residual
// (Kaufman's approximation): per bin, the derivative columns v_k = w A_b
// dc/dtheta_k are projected off the span of the bin's active amplitude
// columns, H = sum_b (v^T v - m^T G^-1 m) with m = M^T v and G = M^T M, and
// g = sum_b v^T r (the NNLS residual is orthogonal to M).  A bin whose A_b is
// 0 does not depend on theta.  The four waves take bins round-robin, lanes
// take lags; every decision is made on workgroup sums, so all threads agree.
// The start grids are ranked with the FP32 model (113 points over every bin
// of the video; the per-q kernel ranks its few lags in FP64); LM accepts and
// rejects steps on the FP64 objective.
// Plain seeding follows fitBrownian / fitSwimmer (:500-536, :577-611): a
// 49-point ln D grid refined by LM in ln D, then the 16 vbar x 2 Z x 2 alpha
// start grid at that D and LM from the best `nstarts` (Nelder-Mead in the
// reference).
namespace {

struct GFitIn {
  const float* tau;      // [nvid * MAXT]
  const float* d;        // [nbins * MAXT]
  const float* w;        // [nbins]
  const float* qn;       // [nbins * K]
  const int* bin_off;    // [nvid + 1]
  const int* ntau;       // [nvid]
  const float* bounds;   // [nvid * 4] vmin, vmax, Dmin, Dmax
  const float* seeds;    // [nvid * 4] alpha, ln vbar, ln Z, ln D, or null
  const float* vmag;     // [nvid] or null
  int K;
  Isf2dDev t2d;
};
constexpr int GNT = 256;    // 4 waves per video (up to 512 registers per lane: the FP64 model spills below that)
constexpr int GW = GNT / 64;

// 2-column NNLS as nnls2, amplitudes in FP64
__device__ double nnls2_d(double g11, double g12, double n, double h1, double h2, double dd, double& A, double& B) {
  double best = dd, bA = 0, bB = 0;
  if (g11 > 0) {
    const double a = h1 / g11, r = dd - a * h1;
    if (a >= 0 && r < best) best = r, bA = a, bB = 0;
  }
  if (n > 0) {
    const double b = h2 / n, r = dd - b * h2;
    if (b >= 0 && r < best) best = r, bA = 0, bB = b;
  }
  const double det = g11 * n - g12 * g12;
  if (det > 1e-11 * fmax(g11, n) * fmax(g11, n)) {
    const double a = (h1 * n - g12 * h2) / det, b = (g11 * h2 - g12 * h1) / det, r = dd - a * h1 - b * h2;
    if (a >= 0 && b >= 0 && r < best) best = r, bA = a, bB = b;
  }
  A = bA;
  B = bB;
  return best > 0 ? best : 0;
}

struct GCtx {
  int b0, b1, nt, K, lane, wave;
  float tv[MAXT / 64];
  float V;
};

__device__ __forceinline__ void load_q(const GFitIn& in, int b, int K, float* q) {
#pragma unroll
  for (int i = 0; i < MAXK; i++) q[i] = i < K ? in.qn[(long)b * K + i] : 0.f;
}

// RSS = sum_b w_b^2 rss_b at th (FP64 model); the bins' NNLS amplitudes go to ab.
template <int GEO, bool SWIM>
__device__ double g_eval(const GFitIn& in, const GCtx& cx, const double* th, double* ab, double* red) {
  const double alpha = SWIM ? th[0] : 0.0, vbar = exp(th[1]), Z = exp(th[2]), Dd = exp(th[3]);
  double acc = 0;
  for (int b = cx.b0 + cx.wave; b < cx.b1; b += GW) {
    float q[MAXK];
    load_q(in, b, cx.K, q);
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < MAXT / 64; r++) {
      const int j = cx.lane + 64 * r;
      if (j < cx.nt) {
        const double c = column_d<GEO>(cx.tv[r], q, cx.K, alpha, vbar, Z, Dd, cx.V, in.t2d, SWIM);
        const double dv = in.d[(long)b * MAXT + j];
        a[0] += c * c;
        a[1] += c;
        a[2] += c * dv;
        a[3] += dv;
        a[4] += dv * dv;
      }
    }
    const double t = wave_sum_split<3>(a, cx.lane);
    double A, B;
    const double rss = nnls2_d(readlane_d(t, split_lane(0, 3)), readlane_d(t, split_lane(1, 3)), (double)cx.nt,
                               readlane_d(t, split_lane(2, 3)), readlane_d(t, split_lane(3, 3)),
                               readlane_d(t, split_lane(4, 3)), A, B);
    const double w = in.w[b];
    acc += w * w * rss;
    if (cx.lane == 0) {
      ab[2 * (b - cx.b0)] = A;
      ab[2 * (b - cx.b0) + 1] = B;
    }
  }
  if (cx.lane == 0) red[cx.wave] = acc;
  __syncthreads();
  double chi = 0;
#pragma unroll
  for (int k = 0; k < GW; k++) chi += red[k];
  __syncthreads();
  return chi;
}

// The same objective with the FP32 model (ranking the start grids only).
template <int GEO, bool SWIM>
__device__ double g_eval32(const GFitIn& in, const GCtx& cx, const double* th, double* red) {
  const float alpha = SWIM ? (float)th[0] : 0.f, vbar = (float)exp(th[1]), Z = (float)exp(th[2]),
              Dd = (float)exp(th[3]);
  double acc = 0;
  for (int b = cx.b0 + cx.wave; b < cx.b1; b += GW) {
    float q[MAXK];
    load_q(in, b, cx.K, q);
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < MAXT / 64; r++) {
      const int j = cx.lane + 64 * r;
      if (j < cx.nt) {
        float c, u0, u1, u2, u3;
        column<GEO>(cx.tv[r], q, cx.K, alpha, vbar, Z, Dd, cx.V, in.t2d, SWIM, false, c, u0, u1, u2, u3);
        const double dv = in.d[(long)b * MAXT + j];
        a[0] += (double)c * c;
        a[1] += c;
        a[2] += c * dv;
        a[3] += dv;
        a[4] += dv * dv;
      }
    }
    const double t = wave_sum_split<3>(a, cx.lane);
    double A, B;
    const double rss = nnls2_d(readlane_d(t, split_lane(0, 3)), readlane_d(t, split_lane(1, 3)), (double)cx.nt,
                               readlane_d(t, split_lane(2, 3)), readlane_d(t, split_lane(3, 3)),
                               readlane_d(t, split_lane(4, 3)), A, B);
    const double w = in.w[b];
    acc += w * w * rss;
  }
  if (cx.lane == 0) red[cx.wave] = acc;
  __syncthreads();
  double chi = 0;
#pragma unroll
  for (int k = 0; k < GW; k++) chi += red[k];
  __syncthreads();
  return chi;
}

// Reduced normal equations at th from the amplitudes ab of the evaluation at
// th: H (packed lower 4x4) and g.
template <int GEO, bool SWIM>
__device__ void g_normal(const GFitIn& in, const GCtx& cx, const double* th, const double* ab, double* H, double* g,
                         double* red) {
  const float alpha = SWIM ? (float)th[0] : 0.f, vbar = (float)exp(th[1]), Z = (float)exp(th[2]),
              Dd = (float)exp(th[3]);
  double Hw[10], gw[4];
#pragma unroll
  for (int k = 0; k < 10; k++) Hw[k] = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) gw[k] = 0;
  for (int b = cx.b0 + cx.wave; b < cx.b1; b += GW) {
    const double A = ab[2 * (b - cx.b0)], B = ab[2 * (b - cx.b0) + 1];
    if (!(A > 0)) continue;   // the bin's fit does not depend on theta
    float q[MAXK];
    load_q(in, b, cx.K, q);
    // S_kl (10), c dc_k (4), dc_k (4), d dc_k (4), c^2, c: 24 sums.  A lane
    // holds at most MAXT/64 = 2 lags, so its partial sums are FP32 (its terms
    // are FP32 products already); the wave sums are FP64, eight at a time.
    float a[24];
#pragma unroll
    for (int k = 0; k < 24; k++) a[k] = 0.f;
#pragma unroll
    for (int r = 0; r < MAXT / 64; r++) {
      const int j = cx.lane + 64 * r;
      if (j < cx.nt) {
        float c, dc[4];
        column<GEO>(cx.tv[r], q, cx.K, alpha, vbar, Z, Dd, cx.V, in.t2d, SWIM, true, c, dc[0], dc[1], dc[2], dc[3]);
        const float dv = in.d[(long)b * MAXT + j];
        int h = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
#pragma unroll
          for (int l = 0; l <= k; l++) a[h++] += dc[k] * dc[l];
          a[10 + k] += c * dc[k];
          a[14 + k] += dc[k];
          a[18 + k] += dv * dc[k];
        }
        a[22] += c * c;
        a[23] += c;
      }
    }
    double S[24];
#pragma unroll
    for (int grp = 0; grp < 3; grp++) {
      double t8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) t8[k] = a[8 * grp + k];
      const double t = wave_sum_split<3>(t8, cx.lane);
#pragma unroll
      for (int k = 0; k < 8; k++) S[8 * grp + k] = readlane_d(t, split_lane(k, 3));
    }
    const double w = in.w[b], w2 = w * w;
    // projection on the active amplitude columns: (w c, w) when B > 0, else w c alone
    const double G00 = w2 * S[22], G01 = w2 * S[23], G11 = w2 * cx.nt;
    const double det = G00 * G11 - G01 * G01;
    const bool two = B > 0 && det > 1e-11 * fmax(G00, G11) * fmax(G00, G11);
    double m0[4], m1[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      m0[k] = w2 * A * S[10 + k];
      m1[k] = w2 * A * S[14 + k];
      gw[k] += w2 * A * (S[18 + k] - A * S[10 + k] - B * S[14 + k]);
    }
    int h = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int l = 0; l <= k; l++, h++) {
        const double proj = two ? (m0[k] * (G11 * m0[l] - G01 * m1[l]) + m1[k] * (G00 * m1[l] - G01 * m0[l])) / det
                                : (G00 > 0 ? m0[k] * m0[l] / G00 : 0.0);
        Hw[h] += w2 * A * A * S[h] - proj;
      }
  }
  // the four waves' partials
  if (cx.lane == 0) {
#pragma unroll
    for (int k = 0; k < 10; k++) red[cx.wave * 16 + k] = Hw[k];
#pragma unroll
    for (int k = 0; k < 4; k++) red[cx.wave * 16 + 10 + k] = gw[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 14; k++) {
    double sum = 0;
#pragma unroll
    for (int w = 0; w < GW; w++) sum += red[w * 16 + k];
    if (k < 10) H[k] = sum;
    else g[k - 10] = sum;
  }
  __syncthreads();
}

// LM over (alpha, ln vbar, ln Z, ln D) (SWIM) or ln D alone, from th; the
// amplitudes of the returned point are left in *abp.  Returns RSS.
template <int GEO, bool SWIM>
__device__ double g_lm(const GFitIn& in, const GCtx& cx, double* th, const double* lo, const double* hi, int maxit,
                       double** abp, double** abq, double* red, int* steps) {
  for (int k = 0; k < 4; k++) th[k] = fmin(hi[k], fmax(lo[k], th[k]));
  double chi = g_eval<GEO, SWIM>(in, cx, th, *abp, red);
  double lam = 1e-3, nu = 2;
  int it = 0;
  bool done = false;
  while (it < maxit && !done) {
    double H[10], g[4];
    g_normal<GEO, SWIM>(in, cx, th, *abp, H, g, red);
    bool frozen[4];
    double diag[4], dmax = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      frozen[a] = (!SWIM && a < 3) || (th[a] <= lo[a] && g[a] < 0) || (th[a] >= hi[a] && g[a] > 0);
      const double Haa = H[a * (a + 1) / 2 + a];
      if (!frozen[a]) dmax = fmax(dmax, Haa);
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const double Haa = H[a * (a + 1) / 2 + a], fl = fmax(1e-12 * dmax, 1e-300);
      diag[a] = Haa > fl ? Haa : fl;
      if (frozen[a]) g[a] = 0;
    }
    for (;;) {
      it++;
      double Lc[10];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b <= a; b++) {
          const int h = a * (a + 1) / 2 + b;
          Lc[h] = (frozen[a] || frozen[b]) ? (a == b ? 1.0 : 0.0) : H[h] + (a == b ? lam * diag[a] : 0.0);
        }
      bool ok = true;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const int ra = a * (a + 1) / 2;
#pragma unroll
        for (int b = 0; b <= a; b++) {
          const int rb = b * (b + 1) / 2;
          double s = Lc[ra + b];
#pragma unroll
          for (int k = 0; k < b; k++) s -= Lc[ra + k] * Lc[rb + k];
          if (a == b) {
            ok = ok && s > 0;
            Lc[ra + a] = sqrt(s > 0 ? s : 1.0);
          } else {
            Lc[ra + b] = s / Lc[rb + b];
          }
        }
      }
      if (ok) {
        double y[4], x[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
          double s = g[a];
#pragma unroll
          for (int k = 0; k < a; k++) s -= Lc[a * (a + 1) / 2 + k] * y[k];
          y[a] = s / Lc[a * (a + 1) / 2 + a];
        }
#pragma unroll
        for (int a = 3; a >= 0; a--) {
          double s = y[a];
#pragma unroll
          for (int k = a + 1; k < 4; k++) s -= Lc[k * (k + 1) / 2 + a] * x[k];
          x[a] = s / Lc[a * (a + 1) / 2 + a];
        }
        double nth[4], pred = 0;
#pragma unroll
        for (int a = 0; a < 4; a++) {
          nth[a] = frozen[a] ? th[a] : fmin(hi[a], fmax(lo[a], th[a] + x[a]));
          const double dk = nth[a] - th[a];
          pred += dk * (lam * diag[a] * dk + g[a]);
        }
        const double nchi = g_eval<GEO, SWIM>(in, cx, nth, *abq, red);
        const double rho = pred > 0 ? (chi - nchi) / pred : -1;
        if (nchi < chi && rho > 0) {
          const double rel = (chi - nchi) / fmax(chi, 1e-300);
#pragma unroll
          for (int a = 0; a < 4; a++) th[a] = nth[a];
          chi = nchi;
          double* tmp = *abp;
          *abp = *abq;
          *abq = tmp;
          const double f = 1 - (2 * rho - 1) * (2 * rho - 1) * (2 * rho - 1);
          lam *= f > 1.0 / 3 ? f : 1.0 / 3;
          if (lam < 1e-15) lam = 1e-15;
          nu = 2;
          if ((rel < 1e-9 && pred < 1e-9 * chi) || rel < 1e-12 || chi < 1e-300) done = true;
          break;
        }
      }
      lam *= nu;
      nu *= 2;
      if (lam > 1e15 || it >= maxit) {
        done = true;
        break;
      }
    }
  }
  *steps += it;
  return chi;
}

// mode 0: LM from the seeds; 1: plain (the reference's grids).  swim 0: the
// Brownian model only.  Outputs per video: theta (alpha, ln vbar, ln Z, ln D),
// RSS, LM steps, the Brownian fit (D, RSS; plain mode), per bin (A, B).
// split > 0 (plain swimmer fits of small batches): `split` workgroups per
// video, workgroup s running the grids and the LM from start s only, each
// writing its own output slot (blockIdx.x); the caller keeps the best start,
// which is what the sequential loop over starts returns.
template <int GEO>
__global__ __launch_bounds__(GNT) void k_lm_global(GFitIn in, int mode, int swim, int nstarts, int maxit, int split,
                                                   double* __restrict__ theta_out, double* __restrict__ chi_out,
                                                   int* __restrict__ steps_out, double* __restrict__ brown_out,
                                                   double* __restrict__ ab_out) {
  __shared__ double ab0[2 * DDM_GLOBAL_MAX_BINS], ab1[2 * DDM_GLOBAL_MAX_BINS], red[16 * GW];
  const int v = split > 0 ? (int)blockIdx.x / split : (int)blockIdx.x;
  const int s_only = split > 0 ? (int)blockIdx.x % split : -1;
  GCtx cx;
  cx.lane = threadIdx.x & 63;
  cx.wave = threadIdx.x >> 6;
  cx.b0 = in.bin_off[v];
  cx.b1 = in.bin_off[v + 1];
  cx.nt = min(in.ntau[v], MAXT);
  cx.K = min(in.K, MAXK);
  cx.V = in.vmag ? in.vmag[v] : 0.f;
#pragma unroll
  for (int r = 0; r < MAXT / 64; r++) {
    const int j = cx.lane + 64 * r;
    cx.tv[r] = j < cx.nt ? in.tau[(long)v * MAXT + j] : 0.f;
  }
  const double lo[4] = {0.0, log((double)in.bounds[4 * v]), log(0.5), log((double)in.bounds[4 * v + 2])};
  const double hi[4] = {1.0, log((double)in.bounds[4 * v + 1]), log(100.0), log((double)in.bounds[4 * v + 3])};
  double* abp = ab0;
  double* abq = ab1;
  int steps = 0;
  double th[4], chi;
  if (mode == 0) {
    for (int k = 0; k < 4; k++) th[k] = in.seeds[4 * v + k];
    chi = swim ? g_lm<GEO, true>(in, cx, th, lo, hi, maxit, &abp, &abq, red, &steps)
               : g_lm<GEO, false>(in, cx, th, lo, hi, maxit, &abp, &abq, red, &steps);
  } else {
    // fitBrownian: 49-point ln D grid, then LM in ln D
    double tb[4] = {0.0, lo[1], lo[2], lo[3]}, cb = 1e300;
    for (int i = 0; i <= 48; i++) {
      const double t[4] = {0.0, lo[1], lo[2], lo[3] + (hi[3] - lo[3]) * i / 48};
      const double c = g_eval32<GEO, false>(in, cx, t, red);
      if (c < cb) {
        cb = c;
        tb[3] = t[3];
      }
    }
    cb = g_lm<GEO, false>(in, cx, tb, lo, hi, maxit, &abp, &abq, red, &steps);
    if (s_only > 0) steps = 0;   // the Brownian steps count once per video
    if (threadIdx.x == 0 && brown_out && s_only <= 0) {
      brown_out[2 * v] = exp(tb[3]);
      brown_out[2 * v + 1] = cb;
    }
    if (!swim) {
      for (int k = 0; k < 4; k++) th[k] = tb[k];
      chi = cb;
    } else {
      // fitSwimmer's start grid at the Brownian D: the best nstarts (<= 3),
      // kept as grid indices (i, zi, ai) in a sorted insertion where an
      // earlier start wins ties, as the reference's stable sort keeps it first
      int si[3] = {-1, -1, -1};
      double sc[3] = {1e300, 1e300, 1e300};
      const int ns = nstarts < 1 ? 1 : (nstarts > 3 ? 3 : nstarts);
      for (int gi = 0; gi < 64; gi++) {
        const int i = gi >> 2, zi = (gi >> 1) & 1, ai = gi & 1;
        const double t[4] = {ai ? 0.8 : 0.3, lo[1] + (hi[1] - lo[1]) * (i + 0.5) / 16, log(zi ? 8.0 : 1.5), tb[3]};
        const double c = g_eval32<GEO, true>(in, cx, t, red);
        if (c < sc[2]) {
          if (c < sc[1]) {
            sc[2] = sc[1];
            si[2] = si[1];
            if (c < sc[0]) {
              sc[1] = sc[0];
              si[1] = si[0];
              sc[0] = c;
              si[0] = gi;
            } else {
              sc[1] = c;
              si[1] = gi;
            }
          } else {
            sc[2] = c;
            si[2] = gi;
          }
        }
      }
      chi = 1e300;
      for (int k = 0; k < 4; k++) th[k] = tb[k];
      for (int s = 0; s < ns; s++) {
        const int gi = s == 0 ? si[0] : (s == 1 ? si[1] : si[2]);
        if (gi < 0 || (s_only >= 0 && s != s_only)) continue;
        const int i = gi >> 2, zi = (gi >> 1) & 1, ai = gi & 1;
        double t[4] = {ai ? 0.8 : 0.3, lo[1] + (hi[1] - lo[1]) * (i + 0.5) / 16, log(zi ? 8.0 : 1.5), tb[3]};
        const double c = g_lm<GEO, true>(in, cx, t, lo, hi, maxit, &abp, &abq, red, &steps);
        if (c < chi) {
          chi = c;
          for (int k = 0; k < 4; k++) th[k] = t[k];
        }
      }
    }
  }
  // the amplitudes at the reported point
  chi = swim ? g_eval<GEO, true>(in, cx, th, abp, red) : g_eval<GEO, false>(in, cx, th, abp, red);
  const int slot = blockIdx.x;
  if (threadIdx.x == 0) {
    for (int k = 0; k < 4; k++) theta_out[4 * slot + k] = th[k];
    chi_out[slot] = chi;
    steps_out[slot] = steps;
  }
  const long ab0ff = split > 0 ? 2L * ((long)cx.b0 * split + (long)s_only * (cx.b1 - cx.b0)) : 2L * cx.b0;
  for (int i = threadIdx.x; i < 2 * (cx.b1 - cx.b0); i += GNT) ab_out[ab0ff + i] = abp[i];
}

}  // namespace

hipError_t launch_lm_global(const float* tau, const float* d, const float* w, const float* qn, int K,
                            const int* bin_off, const int* ntau, const float* bounds, const float* seeds,
                            const float* vmag, int nvid, int model, const Isf2dDev* t2d, int mode, int nstarts,
                            int maxit, int split, double* theta, double* chi, int* steps, double* brown, double* ab,
                            hipStream_t st) {
  if (K > MAXK || model < 0 || model > 2 || (model == 2 && !t2d) || nvid < 1 || split < 0 || split > 3)
    return hipErrorInvalidValue;
  GFitIn in{tau, d, w, qn, bin_off, ntau, bounds, seeds, vmag, K, model == 2 ? *t2d : Isf2dDev{}};
  const dim3 grid(split > 0 ? nvid * split : nvid);
  if (model == 2)
    hipLaunchKernelGGL(k_lm_global<1>, grid, dim3(GNT), 0, st, in, mode, 1, nstarts, maxit, split, theta, chi, steps,
                       brown, ab);
  else
    hipLaunchKernelGGL(k_lm_global<0>, grid, dim3(GNT), 0, st, in, mode, model == 1 ? 1 : 0, nstarts, maxit, split,
                       theta, chi, steps, brown, ab);
  return hipGetLastError();
}

size_t lm_fit_workspace_bytes(int nfits) {
  return (size_t)nfits * (LM_NS * 8 + 2 * LM_NS * 6 * 4 + 4 + LM_NS * 4);
}

template <bool SWIM, int GEO>
static hipError_t launch_lm_fit_t(const FitIn& in, int nfits, int mode, int nstarts, int maxit, float* theta,
                                  double* chi, int* steps, void* ws, hipStream_t st) {
  static_assert(FNT == 64, "one wave per workgroup: the LM launch numbers its workgroups by (fit, start)");
  if (!ws) return hipErrorInvalidValue;
  const LmWork w = lm_work(ws, nfits);
  const int ns = SWIM && mode == 1 ? nstarts : 1;   // seeded fits and the Brownian grid: one start
  hipLaunchKernelGGL((k_lm_fit<SWIM, GEO, false>), dim3(nfits), dim3(FNT), 0, st, in, nfits, mode, nstarts, maxit, w);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_lm_fit<SWIM, GEO, true>), dim3(nfits * ns), dim3(FNT), 0, st, in, nfits, mode, ns, maxit, w);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_lm_pick, dim3((nfits + 255) / 256), dim3(256), 0, st, w, nfits, theta, chi, steps);
  return hipGetLastError();
}

hipError_t launch_lm_fit(const float* tau, const float* d, const int* ntau, const float* w, const float* qn, int K,
                         const float* bounds, const float* seeds, const float* vmag, int nfits, int model,
                         const Isf2dDev* t2d, int seed_mode, int nstarts, int maxit, float* theta, double* chi,
                         int* steps, void* ws, hipStream_t st) {
  // seed_mode: 0 grid, 1 given, 2 nonlinear given (the ABI's DDM_SEED_*) -> kernel mode
  const int mode = seed_mode == 0 ? 1 : (seed_mode == 2 ? 2 : 0);
  if (K > MAXK || model < 0 || model > 2 || (model == 2 && !t2d) || nstarts < 1 || nstarts > LM_NS || nfits < 1)
    return hipErrorInvalidValue;
  FitIn in{tau, d, ntau, w, qn, bounds, seeds, vmag, K, model == 2 ? *t2d : Isf2dDev{}};
  if (model == 2) return launch_lm_fit_t<true, 1>(in, nfits, mode, nstarts, maxit, theta, chi, steps, ws, st);
  if (model == 1) return launch_lm_fit_t<true, 0>(in, nfits, mode, nstarts, maxit, theta, chi, steps, ws, st);
  return launch_lm_fit_t<false, 0>(in, nfits, mode, nstarts, maxit, theta, chi, steps, ws, st);
}

}  // namespace ddm

// ----------------------------------------------------------------- allocation-free fits
//
// ddm_fit_device_noalloc (include/ddm.h): the LM above on device-resident
// arrays with every buffer beyond the caller's inputs and outputs in one
// fixed device array, `scratch`, so a call allocates nothing (for code under
// HIP graph capture and serving loops that must not call hipMalloc).
// Layout in bytes: the in-plane ISF table in [0, kNaTab) (uploaded once per
// device at set-up), then per-fit arrays of kNaFits entries each.  Every
// call on a GPU shares that GPU's copy: calls that overlap on one GPU race.
#include "../../include/ddm.h"
#include "ddm_isf2d.h"

#include <mutex>

namespace ddm {
namespace {
constexpr size_t kNaFits = DDM_FIT_NOALLOC_MAX;
constexpr size_t kNaTabV = 98304;                 // table values (doubles)
constexpr size_t kNaTabZ = kNaTabV * 8;           // Z, u0, du (doubles), off, n (ints): 72 rows each
constexpr size_t kNaTab = 1 << 20;
constexpr size_t kNaTau = kNaTab;
constexpr size_t kNaD = kNaTau + kNaFits * MAXT * 4;
constexpr size_t kNaW = kNaD + kNaFits * MAXT * 4;
constexpr size_t kNaQn = kNaW + kNaFits * 4;
constexpr size_t kNaBounds = kNaQn + kNaFits * MAXK * 4;
constexpr size_t kNaSeeds = kNaBounds + kNaFits * 4 * 4;
constexpr size_t kNaVmag = kNaSeeds + kNaFits * 6 * 4;
constexpr size_t kNaNtau = kNaVmag + kNaFits * 4;
constexpr size_t kNaValid = kNaNtau + kNaFits * 4;
constexpr size_t kNaTheta = kNaValid + kNaFits * 4;
constexpr size_t kNaChi = kNaTheta + kNaFits * 6 * 4;
constexpr size_t kNaSteps = kNaChi + kNaFits * 8;
constexpr size_t kNaWork = (kNaSteps + kNaFits * 4 + 255) / 256 * 256;   // the LM's workspace
constexpr size_t kNaBytes = kNaWork + kNaFits * (LM_NS * 8 + 2 * LM_NS * 6 * 4 + 4 + LM_NS * 4);
static_assert(kNaChi % 8 == 0, "FP64 chi^2 slots");
static_assert(kNaTabZ + Isf2dTable::nz * (3 * 8 + 2 * 4) <= kNaTab, "table region");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// per-device address of `scratch` (hipGetSymbolAddress; the array's name is a
// host-side stand-in) and the device view of the table inside it
constexpr int kNaMaxDev = 64;
std::mutex g_na_mu;
unsigned char* g_na_ptr[kNaMaxDev] = {};
Isf2dDev g_na_t2d[kNaMaxDev];

// One workgroup of MAXT threads per fit: the caller's FP64 rows -> the f32
// rows k_lm_fit reads, with the checks ddm_fit makes on the host (a failing
// fit gets a safe placeholder problem and valid = 0).
__global__ __launch_bounds__(MAXT) void k_fit_pack(ddm_fit_batch b, unsigned char* __restrict__ s) {
  const int i = blockIdx.x, j = threadIdx.x;
  const int K = b.knodes;
  const bool seeded = b.seed_mode != DDM_SEED_GRID;
  const int m = b.ntau[i];
  const bool mok = m >= 1 && m <= b.ntau_max;
  const int mm = mok ? m : 1;
  float* tau = reinterpret_cast<float*>(s + kNaTau) + (size_t)i * MAXT;
  float* d = reinterpret_cast<float*>(s + kNaD) + (size_t)i * MAXT;
  tau[j] = j < mm ? (float)b.tau[(size_t)i * b.ntau_max + j] : 0.f;
  d[j] = j < mm ? (float)b.data[(size_t)i * b.ntau_max + j] : 0.f;
  if (j < K) reinterpret_cast<float*>(s + kNaQn)[(size_t)i * K + j] = (float)b.qnodes[(size_t)i * K + j];
  if (j < 6)
    reinterpret_cast<float*>(s + kNaSeeds)[(size_t)i * 6 + j] = seeded ? (float)b.seeds[(size_t)i * 6 + j] : 0.f;
  if (j == 0) {
    float bd[4];
    for (int k = 0; k < 4; k++) bd[k] = (float)b.bounds[(size_t)i * 4 + k];
    const bool bok = bd[0] > 0 && bd[1] > bd[0] && bd[2] > 0 && bd[3] > bd[2];
    float vm = 0.f;
    bool vok = true;
    if (b.vmag) {
      const double v = b.vmag[i];
      vok = v >= 0 && isfinite(v);
      vm = vok ? (float)v : 0.f;
    }
    float* bo = reinterpret_cast<float*>(s + kNaBounds) + (size_t)i * 4;
    for (int k = 0; k < 4; k++) bo[k] = bok ? bd[k] : (k & 1 ? 2.f : 1.f);
    reinterpret_cast<float*>(s + kNaVmag)[i] = vm;
    reinterpret_cast<float*>(s + kNaW)[i] = (float)b.weight[i];
    reinterpret_cast<int*>(s + kNaNtau)[i] = mm;
    reinterpret_cast<int*>(s + kNaValid)[i] = (mok && bok && vok) ? 1 : 0;
  }
}

// One thread per fit: FP64 outputs, NaN / -1 for the fits that failed a check.
__global__ __launch_bounds__(256) void k_fit_unpack(const unsigned char* __restrict__ s, int nfits,
                                                    double* __restrict__ theta, double* __restrict__ chi2,
                                                    int* __restrict__ iters) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nfits) return;
  const bool ok = reinterpret_cast<const int*>(s + kNaValid)[i] != 0;
  const float* th = reinterpret_cast<const float*>(s + kNaTheta) + (size_t)i * 6;
  for (int k = 0; k < 6; k++) theta[(size_t)i * 6 + k] = ok ? (double)th[k] : __builtin_nan("");
  chi2[i] = ok ? reinterpret_cast<const double*>(s + kNaChi)[i] : __builtin_nan("");
  iters[i] = ok ? reinterpret_cast<const int*>(s + kNaSteps)[i] : -1;
}
}  // namespace

// Set-up on the current device, once: the address of `scratch` and the
// in-plane ISF table built into it (by the FP64 quadrature on the device).
hipError_t noalloc_setup() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kNaMaxDev) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr[dev]) return hipSuccess;
  void* p = nullptr;
  if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
  const Isf2dTable& T = isf2d_table();
  const size_t nz = T.Z.size();
  if (T.v.size() > kNaTabV || nz != (size_t)Isf2dTable::nz) return hipErrorInvalidValue;
  unsigned char* b = static_cast<unsigned char*>(p);
  double* v = reinterpret_cast<double*>(b);
  double* z = reinterpret_cast<double*>(b + kNaTabZ);
  double* u0 = z + nz;
  double* du = u0 + nz;
  int* off = reinterpret_cast<int*>(du + nz);
  int* n = off + nz;
  if ((e = build_isf2d_device(T, v, z, u0, du, off, n, nullptr)) != hipSuccess) return e;
  g_na_t2d[dev] = Isf2dDev{v, z, u0, du, off, n, T.lz0, T.dlz};
  g_na_ptr[dev] = b;
  return hipSuccess;
}

// Enqueues pack -> LM -> unpack on `st` (a stream of device `dev`); b's arrays
// and the outputs are device pointers, b's sizes already checked.
hipError_t launch_fit_noalloc(int dev, const ddm_fit_batch* b, double* theta, double* chi2, int* iters,
                              hipStream_t st) {
  if (dev < 0 || dev >= kNaMaxDev) return hipErrorInvalidDevice;
  unsigned char* s = nullptr;
  Isf2dDev t2d{};
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    s = g_na_ptr[dev];
    t2d = g_na_t2d[dev];
  }
  if (!s) return hipErrorNotInitialized;
  if (b->nfits < 1 || (size_t)b->nfits > kNaFits) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fit_pack, dim3(b->nfits), dim3(MAXT), 0, st, *b, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_lm_fit(reinterpret_cast<const float*>(s + kNaTau), reinterpret_cast<const float*>(s + kNaD),
                    reinterpret_cast<const int*>(s + kNaNtau), reinterpret_cast<const float*>(s + kNaW),
                    reinterpret_cast<const float*>(s + kNaQn), b->knodes,
                    reinterpret_cast<const float*>(s + kNaBounds), reinterpret_cast<const float*>(s + kNaSeeds),
                    b->vmag ? reinterpret_cast<const float*>(s + kNaVmag) : nullptr, b->nfits, b->model, &t2d,
                    b->seed_mode, b->nstarts < 1 ? 1 : (b->nstarts > 3 ? 3 : b->nstarts),
                    b->max_iter < 0 ? 0 : b->max_iter, reinterpret_cast<float*>(s + kNaTheta),
                    reinterpret_cast<double*>(s + kNaChi), reinterpret_cast<int*>(s + kNaSteps), s + kNaWork, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fit_unpack, dim3((b->nfits + 255) / 256), dim3(256), 0, st, s, b->nfits, theta, chi2, iters);
  return hipGetLastError();
}

}  // namespace ddm
