/*
 * ddm_oracle.c -- FP64 CPU restatement of the motility-ddm hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is held
 * to; it is never linked into, loaded by, or called from the product library
 * (hackathon_amd/libddm.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function names the reference lines it restates (paths relative to
 * motility-ddm/ in mehrinsoha77/Hackathon).  Generalisations beyond the
 * reference (any power-of-two N, non power-of-two T via padding to the next
 * power of two >= 2T, any L <= T) are marked as such; they reduce to the
 * reference exactly at the reference geometry.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   orc_fft, orc_dft            direct DFT to 1e-12 (src test t_fft.mjs:23-40), numpy.fft
 *   orc_blackman_harris         w(0)=6e-5, w(N/2)=1, symmetry (t_fft.mjs:71-77)
 *   orc_spectrum2d_half         direct 2D DFT (t_fft.mjs:42-69), numpy.fft.rfft2
 *   orc_temporal_mode(_direct)  WK vs direct sums with masks/shuffle, phase rotation,
 *                               D(0)=0 (t_fft.mjs:79-137)
 *   orc_v8_hypot                node's Math.hypot on 2.1M integer pairs (fixture)
 *   orc_make_bins               brute-force integer enumeration (indices, counts);
 *                               kCenter/kNodes against vectors written by importing
 *                               src/ddm_ref.js (tests/golden/make_golden.mjs)
 *   special functions           reference KAT values (t_kernels.mjs:39-51), scipy
 *   isf3d / isf2d               Simpson quadrature over the Schulz pdf (t_kernels.mjs:60-72)
 *   orc_isf2d_fast              |table - exact| < 2e-5 (t_kernels.mjs:85-102); the exact
 *                               table values against the reference's isf2dFast outputs
 *                               (tests/golden/fits_reference.json, make_golden_fits.mjs)
 *   orc_nnls3                   KKT conditions (t_kernels.mjs:107-134)
 *   orc_fit_brownian/_swimmer   the reference's fitBrownian / fitSwimmer outputs on six
 *                               noisy oracle records (same fixture file)
 *   orc_simulate                the reference's simulate() movies, byte for byte (same file)
 *   orc_lm_fit                  noise-free recovery, closed-form Brownian data; with a
 *                               drift speed, recovery on orc_model_column data
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off; contraction would change
 * V8-hypot bin edges).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------ window */

/* 4-term periodic Blackman-Harris; src/fft.js:126-133 */
void orc_blackman_harris(int n, double *w) {
  for (int i = 0; i < n; i++) {
    double x = (2 * ORC_PI * i) / n;
    w[i] = 0.35875 - 0.48829 * cos(x) + 0.14128 * cos(2 * x) - 0.01168 * cos(3 * x);
  }
}

/* ------------------------------------------------------------------ FFTs */

static int is_pow2(long n) { return n >= 1 && (n & (n - 1)) == 0; }

/* Radix-2 Stockham autosort, natural order, unnormalised both ways;
 * src/fft.js:59-93 (twiddles :18-29).  Interleaved re/im, length 2n. */
int orc_fft(int n, double *data, int inverse) {
  if (!is_pow2(n)) return -1;
  if (n == 1) return 0;
  int half = n >> 1;
  double sgn = inverse ? -1.0 : 1.0;
  double *tw = (double *)malloc(sizeof(double) * n);
  double *a = (double *)malloc(sizeof(double) * 2 * n);
  double *b = (double *)malloc(sizeof(double) * 2 * n);
  for (int m = 0; m < half; m++) {
    double ang = (-2 * ORC_PI * m) / n;
    tw[2 * m] = cos(ang);
    tw[2 * m + 1] = sin(ang);
  }
  memcpy(a, data, sizeof(double) * 2 * n);
  double *src = a, *dst = b;
  for (int Ns = 1; Ns < n; Ns <<= 1) {
    int step = half / Ns;
    for (int j = 0; j < half; j++) {
      int k = j % Ns;
      double wr = tw[2 * k * step];
      double wi = sgn * tw[2 * k * step + 1];
      int i0 = 2 * j, i1 = 2 * (j + half);
      double xr = src[i1] * wr - src[i1 + 1] * wi;
      double xi = src[i1] * wi + src[i1 + 1] * wr;
      int o = 2 * ((j - k) * 2 + k);
      int o2 = o + 2 * Ns;
      dst[o] = src[i0] + xr;
      dst[o + 1] = src[i0 + 1] + xi;
      dst[o2] = src[i0] - xr;
      dst[o2 + 1] = src[i0 + 1] - xi;
    }
    double *t = src;
    src = dst;
    dst = t;
  }
  memcpy(data, src, sizeof(double) * 2 * n);
  free(tw);
  free(a);
  free(b);
  return 0;
}

/* Direct O(n^2) DFT; src/fft.js:101-119.  Works for any n (used for the
 * radix-3 geometry N=384, which the reference's fft() rejects). */
void orc_dft(int n, const double *in, double *out, int inverse) {
  double sgn = inverse ? 1.0 : -1.0;
  /* the n distinct angles sgn 2 pi m / n, m = (k t) mod n, evaluated once */
  double *cs = (double *)malloc(sizeof(double) * 2 * n);
  for (int m = 0; m < n; m++) {
    double ang = (sgn * 2 * ORC_PI * (double)m) / n;
    cs[2 * m] = cos(ang);
    cs[2 * m + 1] = sin(ang);
  }
  for (int k = 0; k < n; k++) {
    double sr = 0, si = 0;
    long m = 0;
    for (int t = 0; t < n; t++) {
      double c = cs[2 * m], s = cs[2 * m + 1];
      sr += in[2 * t] * c - in[2 * t + 1] * s;
      si += in[2 * t] * s + in[2 * t + 1] * c;
      m += k;
      if (m >= n) m -= n;
    }
    out[2 * k] = sr;
    out[2 * k + 1] = si;
  }
  free(cs);
}

/* FFT for any n: Stockham when n is a power of two, direct DFT otherwise. */
static void fft_any(int n, double *data, int inverse) {
  if (is_pow2(n)) {
    orc_fft(n, data, inverse);
  } else {
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
    orc_dft(n, data, tmp, inverse);
    memcpy(data, tmp, sizeof(double) * 2 * n);
    free(tmp);
  }
}

/* Half-plane 2D spectrum: mean -> window -> rows -> cols -> x 1/N;
 * mode m = ky*NH + kx; src/fft.js:149-183. */
int orc_spectrum2d_half(const double *img, int N, const double *win, double *out) {
  int NH = (N >> 1) + 1;
  double mean = 0;
  for (long i = 0; i < (long)N * N; i++) mean += img[i];
  mean /= (double)N * N;
  double *row = (double *)malloc(sizeof(double) * 2 * N);
  double *rows = (double *)malloc(sizeof(double) * 2 * (size_t)N * NH);
  for (int y = 0; y < N; y++) {
    double wy = win[y];
    for (int x = 0; x < N; x++) {
      row[2 * x] = (img[(long)y * N + x] - mean) * win[x] * wy;
      row[2 * x + 1] = 0;
    }
    fft_any(N, row, 0);
    for (int kx = 0; kx < NH; kx++) {
      rows[2 * ((long)y * NH + kx)] = row[2 * kx];
      rows[2 * ((long)y * NH + kx) + 1] = row[2 * kx + 1];
    }
  }
  double s = 1.0 / N;
  for (int kx = 0; kx < NH; kx++) {
    for (int y = 0; y < N; y++) {
      row[2 * y] = rows[2 * ((long)y * NH + kx)];
      row[2 * y + 1] = rows[2 * ((long)y * NH + kx) + 1];
    }
    fft_any(N, row, 0);
    for (int ky = 0; ky < N; ky++) {
      out[2 * ((long)ky * NH + kx)] = row[2 * ky] * s;
      out[2 * ((long)ky * NH + kx) + 1] = row[2 * ky + 1] * s;
    }
  }
  free(row);
  free(rows);
  return 0;
}

/* History.ingest: spectrum of one frame into ring slot `slot`,
 * hist[(m*T + slot)*2 + {0,1}]; src/ddm_ref.js:196-205. */
void orc_history_ingest(double *hist, int N, int T, const double *img, const double *win, int slot) {
  int NH = (N >> 1) + 1;
  long modes = (long)N * NH;
  double *spec = (double *)malloc(sizeof(double) * 2 * modes);
  orc_spectrum2d_half(img, N, win, spec);
  for (long m = 0; m < modes; m++) {
    long o = (m * T + slot) * 2;
    hist[o] = spec[2 * m];
    hist[o + 1] = spec[2 * m + 1];
  }
  free(spec);
}

/* Batch form used by tests and the CPU baseline: `frames` are nframes
 * integer/float frames (dtype 0=u8, 1=u16, 2=f32) scaled by `scale`, stored
 * into consecutive slots (first_slot + i) mod T. */
void orc_history_ingest_stack(double *hist, int N, int T, const void *frames, int dtype, double scale,
                              int nframes, int first_slot) {
  long npx = (long)N * N;
  double *img = (double *)malloc(sizeof(double) * npx);
  double *win = (double *)malloc(sizeof(double) * N);
  orc_blackman_harris(N, win);
  for (int f = 0; f < nframes; f++) {
    for (long i = 0; i < npx; i++) {
      double v;
      if (dtype == 0) v = ((const uint8_t *)frames)[f * npx + i];
      else if (dtype == 1) v = ((const uint16_t *)frames)[f * npx + i];
      else v = ((const float *)frames)[f * npx + i];
      img[i] = v * scale;
    }
    orc_history_ingest(hist, N, T, img, win, (first_slot + f) % T);
  }
  free(img);
  free(win);
}

/* ------------------------------------------------------------------ temporal */

static int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

/* Masked Wiener-Khinchin structure function of one mode;
 * src/ddm_ref.js:229-288.  The reference pads to n = 2T (T a power of two);
 * here n = next power of two >= 2T, which gives the same linear correlation
 * for every T.  Returns P (static power). */
double orc_temporal_mode(const double *hist, long mode, int T, const uint32_t *order, const float *mask, int L,
                         double *outD, double *outC) {
  int n = next_pow2(2 * T);
  double *z = (double *)calloc(2 * n, sizeof(double));
  double *y = (double *)calloc(2 * n, sizeof(double));
  double *S = (double *)calloc(2 * n, sizeof(double));
  long base = mode * T * 2;
  for (int j = 0; j < T; j++) {
    if (!mask[j]) continue;
    uint32_t s = order[j];
    double xr = hist[base + 2 * s], xi = hist[base + 2 * s + 1];
    z[2 * j] = 1;
    z[2 * j + 1] = xr * xr + xi * xi;
    y[2 * j] = xr;
    y[2 * j + 1] = xi;
  }
  orc_fft(n, z, 0);
  for (int k = 0; k < n; k++) {
    int kk = (n - k) % n;
    double zr = z[2 * k], zi = z[2 * k + 1];
    double cr = z[2 * kk], ci = -z[2 * kk + 1];
    double Mr = 0.5 * (zr + cr), Mi = 0.5 * (zi + ci);
    double dr = zr - cr, di = zi - ci;
    double Pr = 0.5 * di, Pi = -0.5 * dr;
    S[2 * k] = Mr * Pr + Mi * Pi;
    S[2 * k + 1] = Mr * Pi - Mi * Pr + (Mr * Mr + Mi * Mi);
  }
  orc_fft(n, S, 1);
  orc_fft(n, y, 0);
  for (int k = 0; k < n; k++) {
    double yr = y[2 * k], yi = y[2 * k + 1];
    y[2 * k] = yr * yr + yi * yi;
    y[2 * k + 1] = 0;
  }
  orc_fft(n, y, 1);
  double inv = 1.0 / n;
  double s1 = S[1] * inv;
  double P = S[0] * inv / (s1 > 1e-300 ? s1 : 1e-300);
  for (int tau = 0; tau < L; tau++) {
    double A = S[2 * tau] * inv;
    double B = S[2 * ((n - tau) % n)] * inv;
    double Np = round(S[2 * tau + 1] * inv); /* Math.round; the argument is never a half-integer */
    if (Np > 0) {
      double Gr = y[2 * tau] * inv, Gi = y[2 * tau + 1] * inv;
      outD[tau] = (A + B - 2 * Gr) / Np;
      if (outC) {
        outC[2 * tau] = Gr / Np;
        outC[2 * tau + 1] = Gi / Np;
      }
    } else {
      outD[tau] = 0;
      if (outC) outC[2 * tau] = outC[2 * tau + 1] = 0;
    }
  }
  free(z);
  free(y);
  free(S);
  return s1 > 0.5 ? P : 0;
}

/* The same WK formulas with the mask and p transformed separately (three
 * FP64 FFTs instead of the reference's packed z = m + i p): packing the O(1)
 * mask with p = |x|^2 in one transform limits p's accuracy to ~ulp x (the
 * mask's energy), which at T = 1024 and |x|^2 ~ 1e-9 is a 1e-4 relative error
 * at the last lags; the split form keeps each term's error relative to its
 * own magnitude.  It is the FP64 truth the engine is held to at full size;
 * orc_temporal_mode keeps the reference's algorithm. */
double orc_temporal_mode_split(const double *hist, long mode, int T, const uint32_t *order, const float *mask, int L,
                               double *outD) {
  int n = next_pow2(2 * T);
  double *M = (double *)calloc(2 * n, sizeof(double));
  double *Pp = (double *)calloc(2 * n, sizeof(double));
  double *y = (double *)calloc(2 * n, sizeof(double));
  long base = mode * T * 2;
  for (int j = 0; j < T; j++) {
    if (!mask[j]) continue;
    uint32_t s = order[j];
    double xr = hist[base + 2 * s], xi = hist[base + 2 * s + 1];
    M[2 * j] = 1;
    Pp[2 * j] = xr * xr + xi * xi;
    y[2 * j] = xr;
    y[2 * j + 1] = xi;
  }
  orc_fft(n, M, 0);
  orc_fft(n, Pp, 0);
  orc_fft(n, y, 0);
  /* A(tau) = sum_t m(t) p(t + tau) = IFFT(conj(M) P); B(tau) = A(-tau); N(tau) = IFFT(|M|^2) */
  double *A = (double *)calloc(2 * n, sizeof(double));
  double *Nn = (double *)calloc(2 * n, sizeof(double));
  for (int k = 0; k < n; k++) {
    double mr = M[2 * k], mi = M[2 * k + 1], pr = Pp[2 * k], pi = Pp[2 * k + 1];
    A[2 * k] = mr * pr + mi * pi;
    A[2 * k + 1] = mr * pi - mi * pr;
    Nn[2 * k] = mr * mr + mi * mi;
    double yr = y[2 * k], yi = y[2 * k + 1];
    y[2 * k] = yr * yr + yi * yi;
    y[2 * k + 1] = 0;
  }
  orc_fft(n, A, 1);
  orc_fft(n, Nn, 1);
  orc_fft(n, y, 1);
  double inv = 1.0 / n;
  double n0 = round(Nn[0] * inv), P = 0;
  if (n0 > 0.5) P = A[0] * inv / n0;
  for (int tau = 0; tau < L; tau++) {
    double Av = A[2 * tau] * inv, Bv = A[2 * ((n - tau) % n)] * inv;
    double Np = round(Nn[2 * tau] * inv);
    outD[tau] = Np > 0 ? (Av + Bv - 2 * y[2 * tau] * inv) / Np : 0;
  }
  free(M);
  free(Pp);
  free(y);
  free(A);
  free(Nn);
  return P;
}

/* Direct O(T*L) masked sums; src/ddm_ref.js:306-333. */
double orc_temporal_mode_direct(const double *hist, long mode, int T, const uint32_t *order, const float *mask, int L,
                                double *outD, double *outC) {
  long base = mode * T * 2;
  double *xr = (double *)calloc(T, sizeof(double));
  double *xi = (double *)calloc(T, sizeof(double));
  double pSum = 0, nSum = 0;
  for (int j = 0; j < T; j++) {
    if (!mask[j]) continue;
    xr[j] = hist[base + 2 * order[j]];
    xi[j] = hist[base + 2 * order[j] + 1];
    pSum += xr[j] * xr[j] + xi[j] * xi[j];
    nSum++;
  }
  for (int tau = 0; tau < L; tau++) {
    double d = 0, gr = 0, gi = 0, cnt = 0;
    for (int t = 0; t + tau < T; t++) {
      if (!mask[t] || !mask[t + tau]) continue;
      double dr = xr[t + tau] - xr[t], di = xi[t + tau] - xi[t];
      d += dr * dr + di * di;
      gr += xr[t] * xr[t + tau] + xi[t] * xi[t + tau];
      gi += xr[t] * xi[t + tau] - xi[t] * xr[t + tau];
      cnt++;
    }
    outD[tau] = cnt ? d / cnt : 0;
    if (outC) {
      outC[2 * tau] = cnt ? gr / cnt : 0;
      outC[2 * tau + 1] = cnt ? gi / cnt : 0;
    }
  }
  free(xr);
  free(xi);
  return nSum ? pSum / nSum : 0;
}

/* Per-analysis metadata: exact pair counts N(tau) and the FP64 FFT of the
 * zero-padded mask rounded to f32; src/ddm_ref.js:538-560 (generalised T, L). */
void orc_temporal_meta(int T, int L, const uint32_t *order, const float *mask, float *npairs, float *mspec) {
  (void)order;
  int n = next_pow2(2 * T);
  for (int tau = 0; tau < L; tau++) {
    int c = 0;
    for (int t = 0; t + tau < T; t++)
      if (mask[t] && mask[t + tau]) c++;
    npairs[tau] = (float)c;
  }
  if (mspec) {
    double *z = (double *)calloc(2 * n, sizeof(double));
    for (int j = 0; j < T; j++) z[2 * j] = mask[j] ? 1 : 0;
    orc_fft(n, z, 0);
    for (int i = 0; i < 2 * n; i++) mspec[i] = (float)z[i];
    free(z);
  }
}

/* ------------------------------------------------------------------ bins */

/* V8's Math.hypot for two arguments: scale by the max, Kahan-sum the squared
 * ratios in argument order, sqrt * max.  Not correctly rounded; reproducing it
 * is what makes the reference's bin edges bit-exact (src/ddm_ref.js:93). */
double orc_v8_hypot(double a, double b) {
  double aa = fabs(a), ab = fabs(b);
  if (isinf(aa) || isinf(ab)) return INFINITY;
  if (isnan(aa) || isnan(ab)) return NAN;
  double mx = aa > ab ? aa : ab;
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  double v[2] = {aa, ab};
  for (int i = 0; i < 2; i++) {
    double r = v[i] / mx;
    double summand = r * r - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return sqrt(sum) * mx;
}

/* Math.round: nearest integer, ties toward +infinity. */
static double js_round(double x) {
  double r = floor(x);
  if (x - r >= 0.5) r += 1;
  return r;
}

/* windowLeakage: |W(d)|^2 of the window at fractional offsets d in
 * [-4, 4] step 0.5, normalised to sum 1; src/ddm_ref.js:40-55. */
static int window_leakage(int N, double *d, double *p) {
  double *w = (double *)malloc(sizeof(double) * N);
  orc_blackman_harris(N, w);
  int cnt = 0;
  for (double x = -4; x <= 4 + 1e-9; x += 0.5) {
    double re = 0, im = 0;
    for (int n = 0; n < N; n++) {
      double a = (-2 * ORC_PI * x * n) / N;
      re += w[n] * cos(a);
      im += w[n] * sin(a);
    }
    d[cnt] = x;
    p[cnt] = re * re + im * im;
    cnt++;
  }
  double s = 0;
  for (int i = 0; i < cnt; i++) s += p[i];
  for (int i = 0; i < cnt; i++) p[i] /= s;
  free(w);
  return cnt;
}

/* makeBins; src/ddm_ref.js:84-170.  hypot_mode 0 = V8 Math.hypot (reference
 * parity), 1 = correctly rounded libm hypot.  binModes holds 4 int32 per entry
 * (mode, kx, ky, 0).  Returns the number of binned modes, or -1 if `cap` is
 * too small. */
long orc_make_bins(int N, int NB, double kmin, double kmax, int hypot_mode, int K, uint32_t *binOffsets,
                   int32_t *binModes, long cap, double *kCenter, uint32_t *nModes, double *kNodes) {
  int NH = (N >> 1) + 1;
  long modes = (long)N * NH;
  double width = (kmax - kmin) / NB;
  double (*hyp)(double, double) = hypot_mode ? hypot : orc_v8_hypot;
  int32_t *bin_of = (int32_t *)malloc(sizeof(int32_t) * modes);
  long *count = (long *)calloc(NB, sizeof(long));
  for (long m = 0; m < modes; m++) {
    int kx = (int)(m % NH);
    int kyi = (int)((m - kx) / NH);
    int ky = kyi < N / 2 ? kyi : kyi - N;
    bin_of[m] = -1;
    if (kx == 0 && ky <= 0) continue;
    if (kx == N / 2 || ky == -N / 2) continue;
    double kr = hyp((double)kx, (double)ky);
    double b = floor((kr - kmin) / width);
    if (b >= 0 && b < NB) {
      bin_of[m] = (int32_t)b;
      count[(int)b]++;
    }
  }
  long total = 0;
  for (int b = 0; b < NB; b++) {
    binOffsets[b] = (uint32_t)total;
    total += count[b];
  }
  binOffsets[NB] = (uint32_t)total;
  if (total > cap) {
    free(bin_of);
    free(count);
    return -1;
  }
  long *fill = (long *)calloc(NB, sizeof(long));
  for (long m = 0; m < modes; m++) {
    int b = bin_of[m];
    if (b < 0) continue;
    int kx = (int)(m % NH);
    int kyi = (int)((m - kx) / NH);
    int ky = kyi < N / 2 ? kyi : kyi - N;
    long o = 4 * (binOffsets[b] + fill[b]++);
    binModes[o] = (int32_t)m;
    binModes[o + 1] = kx;
    binModes[o + 2] = ky;
    binModes[o + 3] = 0;
  }
  double ld[64], lp[64];
  int nl = window_leakage(N, ld, lp);
  double h0 = kmin - 6 > 0 ? kmin - 6 : 0, hres = 0.01;
  long hlen = (long)ceil((kmax + 6 - h0) / hres) + 1;
  double *hist = (double *)malloc(sizeof(double) * hlen);
  double *hk = (double *)malloc(sizeof(double) * hlen);
  for (int b = 0; b < NB; b++) {
    long i0 = binOffsets[b], i1 = binOffsets[b + 1];
    double sk = 0;
    for (long i = i0; i < i1; i++) sk += hyp((double)binModes[4 * i + 1], (double)binModes[4 * i + 2]);
    nModes[b] = (uint32_t)(i1 - i0);
    kCenter[b] = (i1 > i0) ? sk / (double)(i1 - i0) : kmin + (b + 0.5) * width;
    if (!kNodes) continue;
    memset(hist, 0, sizeof(double) * hlen);
    memset(hk, 0, sizeof(double) * hlen);
    for (long i = i0; i < i1; i++) {
      int kx = binModes[4 * i + 1], ky = binModes[4 * i + 2];
      for (int a = 0; a < nl; a++) {
        double ax = kx + ld[a];
        for (int c = 0; c < nl; c++) {
          double k = hyp(ax, ky + ld[c]);
          double wt = lp[a] * lp[c];
          double cc = js_round((k - h0) / hres);
          if (cc < 0) cc = 0;
          if (cc > hlen - 1) cc = (double)(hlen - 1);
          long ci = (long)cc;
          hist[ci] += wt;
          hk[ci] += wt * k;
        }
      }
    }
    double wsum = 0;
    for (long c = 0; c < hlen; c++) wsum += hist[c];
    if (!(wsum > 0)) {
      for (int n = 0; n < K; n++) kNodes[b * K + n] = kCenter[b];
      continue;
    }
    double acc = 0, sw = 0, swk = 0;
    int n = 0;
    for (long c = 0; c < hlen && n < K; c++) {
      double rem = hist[c];
      double kc = hist[c] > 0 ? hk[c] / hist[c] : 0;
      while (rem > 0 && n < K) {
        double lim = ((double)(n + 1) / K) * wsum - acc;
        double take = rem < lim ? rem : lim;
        sw += take;
        swk += take * kc;
        acc += take;
        rem -= take;
        if (acc >= ((double)(n + 1) / K) * wsum * (1 - 1e-12)) {
          kNodes[b * K + n] = sw > 0 ? swk / sw : kc;
          n++;
          sw = 0;
          swk = 0;
        }
      }
    }
    for (; n < K; n++) kNodes[b * K + n] = kmax;
  }
  free(hist);
  free(hk);
  free(fill);
  free(bin_of);
  free(count);
  return total;
}

/* analyzeRadial: temporalMode over binned modes, accumulate, divide by the
 * bin's mode count; src/ddm_ref.js:360-437 (the drift sums are not part of
 * this path and are not computed).  D[b*L+tau], P[b]. */
void orc_analyze_radial_split(const double *hist, int T, const uint32_t *order, const float *mask,
                              const uint32_t *binOffsets, const int32_t *binModes, int NB, int L, double *D,
                              double *P) {
  double *d = (double *)malloc(sizeof(double) * L);
  memset(D, 0, sizeof(double) * NB * L);
  memset(P, 0, sizeof(double) * NB);
  for (int b = 0; b < NB; b++) {
    for (uint32_t i = binOffsets[b]; i < binOffsets[b + 1]; i++) {
      double p = orc_temporal_mode_split(hist, binModes[4 * i], T, order, mask, L, d);
      P[b] += p;
      for (int t = 0; t < L; t++) D[b * L + t] += d[t];
    }
    uint32_t n = binOffsets[b + 1] - binOffsets[b];
    if (!n) continue;
    P[b] /= n;
    for (int t = 0; t < L; t++) D[b * L + t] /= n;
  }
  free(d);
}

void orc_analyze_radial(const double *hist, int T, const uint32_t *order, const float *mask, const uint32_t *binOffsets,
                        const int32_t *binModes, int NB, int L, double *D, double *P) {
  double *d = (double *)malloc(sizeof(double) * L);
  memset(D, 0, sizeof(double) * NB * L);
  memset(P, 0, sizeof(double) * NB);
  for (int b = 0; b < NB; b++) {
    for (uint32_t i = binOffsets[b]; i < binOffsets[b + 1]; i++) {
      long m = binModes[4 * i];
      double p = orc_temporal_mode(hist, m, T, order, mask, L, d, NULL);
      P[b] += p;
      for (int t = 0; t < L; t++) D[b * L + t] += d[t];
    }
    uint32_t n = binOffsets[b + 1] - binOffsets[b];
    if (!n) continue;
    P[b] /= n;
    for (int t = 0; t < L; t++) D[b * L + t] /= n;
  }
  free(d);
}

/* Phase-drift sums of accumulateMode (src/ddm_ref.js:360-381): per bin b and
 * lag tau < L-1, over the bin's modes with (kx, ky) = binModes[4i+1], [4i+2]:
 * z = C(tau+1) conj C(tau), w = |z|, phi = atan2(z); drift[(b*L+tau)*8 + c] =
 * sum of w, w kx^2, w kx ky, w ky^2, w kx phi, w ky phi, w phi^2, [w > 0]
 * (Math.hypot of the reference is the correctly rounded hypot here). */
static void drift_accumulate(const double *c, int L, double kx, double ky, double *bin_drift) {
  for (int tau = 0; tau < L - 1; tau++) {
    double ar = c[2 * tau], ai = c[2 * tau + 1], br = c[2 * tau + 2], bi = c[2 * tau + 3];
    double zr = br * ar + bi * ai, zi = bi * ar - br * ai;
    double w = hypot(zr, zi), phi = atan2(zi, zr);
    double *o = bin_drift + (long)tau * 8;
    o[0] += w;
    o[1] += w * kx * kx;
    o[2] += w * kx * ky;
    o[3] += w * ky * ky;
    o[4] += w * kx * phi;
    o[5] += w * ky * phi;
    o[6] += w * phi * phi;
    o[7] += w > 0 ? 1 : 0;
  }
}

void orc_radial_drift(const double *hist, int T, const uint32_t *order, const float *mask, const uint32_t *binOffsets,
                      const int32_t *binModes, int NB, int L, double *drift) {
  double *d = (double *)malloc(sizeof(double) * L);
  double *c = (double *)malloc(sizeof(double) * 2 * L);
  memset(drift, 0, sizeof(double) * NB * L * 8);
  for (int b = 0; b < NB; b++) {
    for (uint32_t i = binOffsets[b]; i < binOffsets[b + 1]; i++) {
      long m = binModes[4 * i];
      orc_temporal_mode(hist, m, T, order, mask, L, d, c);
      drift_accumulate(c, L, binModes[4 * i + 1], binModes[4 * i + 2], drift + (long)b * L * 8);
    }
  }
  free(d);
  free(c);
}

/* The same sums from given per-mode C(tau) (complex, C[(m*L + tau)*2 + {0,1}]),
 * as radialFromFull (src/ddm_ref.js:474-487) does for per-mode arrays. */
void orc_drift_from_modes(const double *C, const uint32_t *binOffsets, const int32_t *binModes, int NB, int L,
                          double *drift) {
  memset(drift, 0, sizeof(double) * NB * L * 8);
  for (int b = 0; b < NB; b++)
    for (uint32_t i = binOffsets[b]; i < binOffsets[b + 1]; i++)
      drift_accumulate(C + (long)binModes[4 * i] * L * 2, L, binModes[4 * i + 1], binModes[4 * i + 2],
                       drift + (long)b * L * 8);
}

/* Same reduction from direct O(T*L) sums (cancellation-free reference for the
 * accuracy check of D at small lags). */
void orc_analyze_radial_direct(const double *hist, int T, const uint32_t *order, const float *mask,
                               const uint32_t *binOffsets, const int32_t *binModes, int NB, int L, double *D,
                               double *P) {
  double *d = (double *)malloc(sizeof(double) * L);
  memset(D, 0, sizeof(double) * NB * L);
  memset(P, 0, sizeof(double) * NB);
  for (int b = 0; b < NB; b++) {
    for (uint32_t i = binOffsets[b]; i < binOffsets[b + 1]; i++) {
      long m = binModes[4 * i];
      double p = orc_temporal_mode_direct(hist, m, T, order, mask, L, d, NULL);
      P[b] += p;
      for (int t = 0; t < L; t++) D[b * L + t] += d[t];
    }
    uint32_t n = binOffsets[b + 1] - binOffsets[b];
    if (!n) continue;
    P[b] /= n;
    for (int t = 0; t < L; t++) D[b * L + t] /= n;
  }
  free(d);
}

/* ------------------------------------------------------------------ special functions */

/* Bessel J0, rational approximations; src/motility_fit.js:54-68 */
double orc_bessel_j0(double x) {
  double ax = fabs(x);
  if (ax < 8) {
    double y = x * x;
    double a1 = 57568490574.0 + y * (-13362590354.0 + y * (651619640.7 + y * (-11214424.18 + y * (77392.33017 + y * -184.9052456))));
    double a2 = 57568490411.0 + y * (1029532985.0 + y * (9494680.718 + y * (59272.64853 + y * (267.8532712 + y))));
    return a1 / a2;
  }
  double z = 8 / ax, y = z * z, xx = ax - 0.785398164;
  double a1 = 1 + y * (-0.1098628627e-2 + y * (0.2734510407e-4 + y * (-0.2073370639e-5 + y * 0.2093887211e-6)));
  double a2 = -0.1562499995e-1 + y * (0.1430488765e-3 + y * (-0.6911147651e-5 + y * (0.7621095161e-6 - y * 0.934935152e-7)));
  return sqrt(0.636619772 / ax) * (cos(xx) * a1 - z * sin(xx) * a2);
}

/* ln Gamma, Lanczos g=7 n=9; src/motility_fit.js:74-86 */
double orc_lngamma(double x) {
  static const double c[9] = {0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                              771.32342877765313, -176.61502916214059, 12.507343278686905,
                              -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
  if (x < 0.5) return log(ORC_PI / fabs(sin(ORC_PI * x))) - orc_lngamma(1 - x);
  x -= 1;
  double a = c[0], t = x + 7.5;
  for (int i = 1; i < 9; i++) a += c[i] / (x + i);
  return 0.5 * log(2 * ORC_PI) + (x + 0.5) * log(t) - t + log(a);
}

/* Schulz speed pdf; src/motility_fit.js:95-99 */
double orc_schulz_pdf(double v, double vbar, double Z) {
  if (v <= 0) return 0;
  double b = (Z + 1) / vbar;
  return exp(Z * log(v) - orc_lngamma(Z + 1) + (Z + 1) * log(b) - b * v);
}

/* Legendre P_nu(c) by Laplace's integral, trapezoid rule;
 * src/motility_fit.js:120-132 */
double orc_legendre_p(double nu, double c) {
  double s = sqrt(fmax(0, 1 - c * c));
  int n = 16 + (int)ceil(nu * atan2(s, c));
  double sum = 0;
  for (int i = 0; i <= n; i++) {
    double phi = (ORC_PI * i) / n;
    double im = s * cos(phi);
    double r2 = c * c + im * im;
    double val = exp(0.5 * nu * log(r2)) * cos(nu * atan2(im, c));
    sum += (i == 0 || i == n) ? 0.5 * val : val;
  }
  return sum / n;
}

/* 3D swimmer ISF closed form; src/motility_fit.js:140-144 */
double orc_isf3d(double x, double Z) {
  double L = x / (Z + 1);
  if (L < 1e-4) return 1 - (x * x * (Z + 2)) / (6 * (Z + 1));
  return sin(Z * atan(L)) / (Z * L * exp(0.5 * Z * log1p(L * L)));
}

/* 2D swimmer ISF closed form; src/motility_fit.js:152-159 */
double orc_isf2d(double x, double Z) {
  double L = x / (Z + 1);
  if (L < 1e-4) return 1 - (x * x * (Z + 2)) / (4 * (Z + 1));
  double u = 1 + L * L;
  double lnPre = -0.5 * (Z + 1) * log(u);
  if (lnPre < -25) return 0;
  return orc_legendre_p(Z, 1 / sqrt(u)) * exp(lnPre);
}

/* Tabulated 2D ISF (Catmull-Rom in ln Lambda, quadratic Lagrange in ln Z);
 * src/motility_fit.js:161-220. */
#define T2D_NZ 72
typedef struct {
  double Z, u0, du;
  int n;
  double *v;
} t2d_row;
static t2d_row t2d_rows[T2D_NZ];
static int t2d_built = 0;
static double t2d_lz0, t2d_dlz;

static double isf2d_exact_lambda(double L, double Z) {
  double u = 1 + L * L;
  double lnPre = -0.5 * (Z + 1) * log(u);
  if (lnPre < -25) return 0;
  return orc_legendre_p(Z, 1 / sqrt(u)) * exp(lnPre);
}

static void build_t2d(void) {
  t2d_lz0 = log(0.5);
  double lz1 = log(100.0);
  t2d_dlz = (lz1 - t2d_lz0) / (T2D_NZ - 1);
  for (int i = 0; i < T2D_NZ; i++) {
    double Z = exp(t2d_lz0 + i * t2d_dlz);
    double Lc = fmin(1e4, sqrt(expm1(50 / (Z + 1))));
    double u0 = log(1e-4), u1 = log(Lc);
    double du = fmin(0.04, 0.1 / ((Z + 0.5) / 2));
    int n = (int)ceil((u1 - u0) / du) + 1;
    double *v = (double *)malloc(sizeof(double) * (n + 2));
    for (int k = 0; k < n + 2; k++) v[k] = isf2d_exact_lambda(exp(u0 + k * du), Z);
    t2d_rows[i] = (t2d_row){Z, u0, du, n, v};
  }
  t2d_built = 1;
}

static double t2d_row_eval(const t2d_row *r, double x) {
  double L = x / (r->Z + 1);
  if (L < 1e-4) return 1 - (x * x * (r->Z + 2)) / (4 * (r->Z + 1));
  double f = (log(L) - r->u0) / r->du;
  if (f >= r->n - 1) return 0;
  int i = (int)floor(f);
  double t = f - i;
  const double *v = r->v;
  double p1 = v[i], p2 = v[i + 1], p3 = v[i + 2];
  double p0 = i > 0 ? v[i - 1] : 2 * p1 - p2;
  return p1 + 0.5 * t * (p2 - p0 + t * (2 * p0 - 5 * p1 + 4 * p2 - p3 + t * (3 * (p1 - p2) + p3 - p0)));
}

double orc_isf2d_fast(double x, double Z) {
  if (!t2d_built) build_t2d();
  double zc = fmin(100, fmax(0.5, Z));
  double zf = (log(zc) - t2d_lz0) / t2d_dlz;
  int i0 = (int)js_round(zf) - 1;
  if (i0 < 0) i0 = 0;
  if (i0 > T2D_NZ - 3) i0 = T2D_NZ - 3;
  double t = zf - i0;
  double g0 = t2d_row_eval(&t2d_rows[i0], x), g1 = t2d_row_eval(&t2d_rows[i0 + 1], x),
         g2 = t2d_row_eval(&t2d_rows[i0 + 2], x);
  return 0.5 * (t - 1) * (t - 2) * g0 - t * (t - 2) * g1 + 0.5 * t * (t - 1) * g2;
}

/* ------------------------------------------------------------------ fit building blocks */

/* lagGrid: 1..10 then x1.08 steps; src/motility_fit.js:358-369 */
int orc_lag_grid(int L, int *lags, int cap) {
  int n = 0;
  for (int t = 1; t <= (10 < L - 1 ? 10 : L - 1); t++)
    if (n < cap) lags[n++] = t;
  double x = 10;
  for (;;) {
    x *= 1.08;
    int t = (int)js_round(x);
    if (t > L - 1) break;
    if (n == 0 || t > lags[n - 1]) {
      if (n < cap) lags[n++] = t;
    }
  }
  return n;
}

/* Gaussian elimination with partial pivoting, k <= 3; src/motility_fit.js:270-291 */
static int solve_small(int k, double A[3][3], const double *b, double *x) {
  double M[3][4];
  double scale = 1e-300;
  for (int i = 0; i < k; i++) {
    for (int j = 0; j < k; j++) M[i][j] = A[i][j];
    M[i][k] = b[i];
    if (fabs(A[i][i]) > scale) scale = fabs(A[i][i]);
  }
  for (int c = 0; c < k; c++) {
    int p = c;
    for (int r = c + 1; r < k; r++)
      if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
    if (fabs(M[p][c]) < 1e-11 * scale) return 0;
    for (int j = 0; j <= k; j++) {
      double t = M[c][j];
      M[c][j] = M[p][j];
      M[p][j] = t;
    }
    for (int r = c + 1; r < k; r++) {
      double f = M[r][c] / M[c][c];
      for (int j = c; j <= k; j++) M[r][j] -= f * M[c][j];
    }
  }
  for (int r = k - 1; r >= 0; r--) {
    double s = M[r][k];
    for (int j = r + 1; j < k; j++) s -= M[r][j] * x[j];
    x[r] = s / M[r][r];
  }
  return 1;
}

/* Exact NNLS for <= 3 columns by support enumeration; src/motility_fit.js:243-263 */
void orc_nnls3(const double *G, const double *h, double dd, double *a_out, double *rss_out) {
  double best_a[3] = {0, 0, 0}, best_rss = dd;
  for (int mask = 1; mask < 8; mask++) {
    int idx[3], k = 0;
    for (int i = 0; i < 3; i++)
      if (mask & (1 << i)) idx[k++] = i;
    double A[3][3], b[3], x[3];
    for (int i = 0; i < k; i++) {
      for (int j = 0; j < k; j++) A[i][j] = G[3 * idx[i] + idx[j]];
      b[i] = h[idx[i]];
    }
    if (!solve_small(k, A, b, x)) continue;
    int neg = 0;
    for (int i = 0; i < k; i++)
      if (x[i] < 0) neg = 1;
    if (neg) continue;
    double rss = dd;
    for (int i = 0; i < k; i++) rss -= x[i] * b[i];
    if (rss < best_rss) {
      best_a[0] = best_a[1] = best_a[2] = 0;
      for (int i = 0; i < k; i++) best_a[idx[i]] = x[i];
      best_rss = rss > 0 ? rss : 0;
    }
  }
  for (int i = 0; i < 3; i++) a_out[i] = best_a[i];
  *rss_out = best_rss;
}

/* modelColumn: c(tau) = 1 - <E (1 - alpha + alpha g(q vbar tau; Z))>_nodes,
 * E = exp(-Dd q^2 tau) [* J0(q |V| tau)]; src/motility_fit.js:417-432.
 * geometry 0 = 3D kernel (isf3d), 1 = 2D exact (isf2d), 2 = 2D table (isf2dFast). */
void orc_model_column(int ntau, const double *tau, int nq, const double *qn, double vbar, double Z, double Dd,
                      double Vmag, double alpha, int swim, int geometry, double *out) {
  double inv = 1.0 / nq;
  double al = swim ? alpha : 0;
  for (int j = 0; j < ntau; j++) {
    double t = tau[j], f = 0;
    for (int i = 0; i < nq; i++) {
      double q = qn[i];
      double E = exp(-Dd * q * q * t);
      if (Vmag > 0) E *= orc_bessel_j0(q * Vmag * t);
      if (al > 0) {
        double x = q * vbar * t;
        double g = geometry == 0 ? orc_isf3d(x, Z) : geometry == 1 ? orc_isf2d(x, Z) : orc_isf2d_fast(x, Z);
        f += E * (1 - al + al * g);
      } else {
        f += E;
      }
    }
    out[j] = 1 - f * inv;
  }
}

/* ------------------------------------------------------------------ global VarPro fit (reference) */

/* evalModel: per bin the model column c(tau), exact 2-column NNLS for (A, B),
 * rss = sum_b rss_b w_b^2; src/motility_fit.js:450-474.  amps (2*nb) may be NULL. */
double orc_eval_model(int nb, int ntau, int K, const double *tau, const double *d, const double *w, const double *qn,
                      double vbar, double Z, double Dd, double Vmag, double alpha, int swim, int geometry,
                      double *amps) {
  double rss = 0;
  double *c = (double *)malloc(sizeof(double) * ntau);
  for (int b = 0; b < nb; b++) {
    orc_model_column(ntau, tau, K, qn + (long)b * K, vbar, Z, Dd, Vmag, alpha, swim, geometry, c);
    double g11 = 0, g12 = 0, h1 = 0, h2 = 0, dd = 0;
    const double *db = d + (long)b * ntau;
    for (int j = 0; j < ntau; j++) {
      g11 += c[j] * c[j];
      g12 += c[j];
      h1 += c[j] * db[j];
      h2 += db[j];
      dd += db[j] * db[j];
    }
    double G[9] = {g11, g12, 0, g12, (double)ntau, 0, 0, 0, 0}, h[3] = {h1, h2, 0}, a[3], r;
    orc_nnls3(G, h, dd, a, &r);
    rss += r * w[b] * w[b];
    if (amps) {
      amps[2 * b] = a[0];
      amps[2 * b + 1] = a[1];
    }
  }
  free(c);
  return rss;
}

/* bounds: src/motility_fit.js:480-491 (GATES.SWIM_MIN_UM_S = 2). out: vmin, vmax, Dmin, Dmax, Zmin, Zmax */
void orc_bounds(int nb, const double *q, double dt, double tauMax, double *out) {
  double qmin = q[0], qmax = q[0];
  for (int b = 1; b < nb; b++) {
    if (q[b] < qmin) qmin = q[b];
    if (q[b] > qmax) qmax = q[b];
  }
  double v0 = 0.3 / (qmax * tauMax);
  out[0] = v0 > 2 ? v0 : 2;
  out[1] = 3 / (qmin * dt);
  out[2] = 0.01 / (qmax * qmax * tauMax);
  out[3] = 10 / (qmin * qmin * dt);
  out[4] = 0.5;
  out[5] = 100;
}

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

typedef double (*orc_objective)(const double *u, void *ctx);

/* Bounded Nelder-Mead (bounds by clamping inside the objective);
 * src/motility_fit.js:302-350.  x0/x (n <= 8). Returns f at the best vertex. */
double orc_nelder_mead(orc_objective f, void *ctx, int n, const double *x0, const double *step, int maxIter, double tol,
                       double *xbest) {
  double S[9][8], fs[9];
  for (int j = 0; j < n; j++) S[0][j] = x0[j];
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) S[i + 1][j] = x0[j];
    S[i + 1][i] += step[i];
  }
  for (int i = 0; i <= n; i++) fs[i] = f(S[i], ctx);
  for (int it = 0; it < maxIter; it++) {
    /* stable sort by f (insertion sort keeps equal elements in order, as Array.sort does) */
    for (int i = 1; i <= n; i++) {
      double fv = fs[i], xv[8];
      for (int j = 0; j < n; j++) xv[j] = S[i][j];
      int k = i - 1;
      while (k >= 0 && fs[k] > fv) {
        fs[k + 1] = fs[k];
        for (int j = 0; j < n; j++) S[k + 1][j] = S[k][j];
        k--;
      }
      fs[k + 1] = fv;
      for (int j = 0; j < n; j++) S[k + 1][j] = xv[j];
    }
    if (fabs(fs[n] - fs[0]) <= tol * (fabs(fs[0]) + 1e-30)) break;
    double c[8] = {0};
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) c[j] += S[i][j] / n;
    double xr[8], xe[8], xc[8];
    for (int j = 0; j < n; j++) xr[j] = c[j] + -1 * (S[n][j] - c[j]);
    double fr = f(xr, ctx);
    if (fr < fs[0]) {
      for (int j = 0; j < n; j++) xe[j] = c[j] + -2 * (S[n][j] - c[j]);
      double fe = f(xe, ctx);
      if (fe < fr) {
        for (int j = 0; j < n; j++) S[n][j] = xe[j];
        fs[n] = fe;
      } else {
        for (int j = 0; j < n; j++) S[n][j] = xr[j];
        fs[n] = fr;
      }
    } else if (fr < fs[n - 1]) {
      for (int j = 0; j < n; j++) S[n][j] = xr[j];
      fs[n] = fr;
    } else {
      int outside = fr < fs[n];
      double t = outside ? -0.5 : 0.5;
      for (int j = 0; j < n; j++) xc[j] = c[j] + t * (S[n][j] - c[j]);
      double fc = f(xc, ctx);
      if (fc < (outside ? fr : fs[n])) {
        for (int j = 0; j < n; j++) S[n][j] = xc[j];
        fs[n] = fc;
      } else {
        for (int i = 1; i <= n; i++) {
          for (int j = 0; j < n; j++) S[i][j] = S[0][j] + 0.5 * (S[i][j] - S[0][j]);
          fs[i] = f(S[i], ctx);
        }
      }
    }
  }
  int ib = 0;
  for (int i = 1; i <= n; i++)
    if (fs[i] < fs[ib]) ib = i;
  for (int j = 0; j < n; j++) xbest[j] = S[ib][j];
  return fs[ib];
}

typedef struct {
  int nb, ntau, K, geometry;
  const double *tau, *d, *w, *qn;
  double B[6], Vmag;
} fit_ctx;

static void swim_unpack(const fit_ctx *c, const double *u, double *p) {
  p[0] = exp(clampd(u[0], log(c->B[0]), log(c->B[1])));
  p[1] = exp(clampd(u[1], log(c->B[4]), log(c->B[5])));
  p[2] = exp(clampd(u[2], log(c->B[2]), log(c->B[3])));
  p[3] = clampd(u[3], 0, 1);
}

static double swim_obj(const double *u, void *vc) {
  fit_ctx *c = (fit_ctx *)vc;
  double p[4];
  swim_unpack(c, u, p);
  return orc_eval_model(c->nb, c->ntau, c->K, c->tau, c->d, c->w, c->qn, p[0], p[1], p[2], c->Vmag, p[3], 1,
                        c->geometry, NULL);
}

static double brown_obj(double lnD, fit_ctx *c) {
  return orc_eval_model(c->nb, c->ntau, c->K, c->tau, c->d, c->w, c->qn, 0, 1, exp(lnD), c->Vmag, 0, 0, 0, NULL);
}

/* fitBrownian: 49-point log-D grid + golden section; src/motility_fit.js:500-536.
 * bounds6 from orc_bounds.  Returns Dd; *rss the final rss. */
double orc_fit_brownian(int nb, int ntau, int K, const double *tau, const double *d, const double *w,
                        const double *qn, const double *bounds6, double Vmag, double *rss) {
  fit_ctx c = {nb, ntau, K, 0, tau, d, w, qn, {0}, Vmag};
  for (int i = 0; i < 6; i++) c.B[i] = bounds6[i];
  double lo = log(c.B[2]), hi = log(c.B[3]);
  int n = 48, bi = 0;
  double bf = INFINITY, xs[49];
  for (int i = 0; i <= n; i++) {
    double x = lo + ((hi - lo) * i) / n;
    xs[i] = x;
    double v = brown_obj(x, &c);
    if (v < bf) {
      bf = v;
      bi = i;
    }
  }
  double a = xs[bi > 0 ? bi - 1 : 0], b = xs[bi < n ? bi + 1 : n];
  double gr = (sqrt(5.0) - 1) / 2;
  double cc = b - gr * (b - a), dd = a + gr * (b - a);
  double fc = brown_obj(cc, &c), fd = brown_obj(dd, &c);
  for (int it = 0; it < 60 && b - a > 1e-6; it++) {
    if (fc < fd) {
      b = dd;
      dd = cc;
      fd = fc;
      cc = b - gr * (b - a);
      fc = brown_obj(cc, &c);
    } else {
      a = cc;
      cc = dd;
      fc = fd;
      dd = a + gr * (b - a);
      fd = brown_obj(dd, &c);
    }
  }
  double lnD = fc < fd ? cc : dd;
  double Dd = exp(lnD);
  if (rss) *rss = orc_eval_model(nb, ntau, K, tau, d, w, qn, 0, 1, Dd, Vmag, 0, 0, 0, NULL);
  return Dd;
}

/* fitSwimmer: 64 grid starts (16 vbar x Z{1.5, 8} x alpha{0.3, 0.8}), Nelder-Mead
 * from the best 3 (<= 400 iterations, tol 1e-10); src/motility_fit.js:577-611.
 * out4 = (vbar, Z, Dd, alpha); returns rss. */
double orc_fit_swimmer(int nb, int ntau, int K, const double *tau, const double *d, const double *w, const double *qn,
                       const double *bounds6, int geometry, double Vmag, double Dstart, double *out4) {
  fit_ctx c = {nb, ntau, K, geometry, tau, d, w, qn, {0}, Vmag};
  for (int i = 0; i < 6; i++) c.B[i] = bounds6[i];
  double D0 = Dstart > 0 ? clampd(Dstart, c.B[2], c.B[3]) : sqrt(c.B[2] * c.B[3]);
  double su[64][4], sf[64];
  int ns = 0;
  for (int i = 0; i < 16; i++) {
    double lv = log(c.B[0]) + ((log(c.B[1]) - log(c.B[0])) * (i + 0.5)) / 16;
    double lZs[2] = {log(1.5), log(8.0)}, as[2] = {0.3, 0.8};
    for (int z = 0; z < 2; z++)
      for (int q = 0; q < 2; q++) {
        su[ns][0] = lv;
        su[ns][1] = lZs[z];
        su[ns][2] = log(D0);
        su[ns][3] = as[q];
        sf[ns] = swim_obj(su[ns], &c);
        ns++;
      }
  }
  /* stable sort of the starts by f */
  int idx[64];
  for (int i = 0; i < 64; i++) idx[i] = i;
  for (int i = 1; i < 64; i++) {
    int v = idx[i], k = i - 1;
    while (k >= 0 && sf[idx[k]] > sf[v]) {
      idx[k + 1] = idx[k];
      k--;
    }
    idx[k + 1] = v;
  }
  double best[4], bestf = INFINITY, step[4] = {0.3, 0.7, 0.7, 0.2};
  for (int s = 0; s < 3; s++) {
    double x[4];
    double fr = orc_nelder_mead(swim_obj, &c, 4, su[idx[s]], step, 400, 1e-10, x);
    if (fr < bestf) {
      bestf = fr;
      for (int j = 0; j < 4; j++) best[j] = x[j];
    }
  }
  swim_unpack(&c, best, out4);
  /* out4 order: vbar, Z, Dd, alpha */
  return orc_eval_model(nb, ntau, K, tau, d, w, qn, out4[0], out4[1], out4[2], Vmag, out4[3], 1, geometry, NULL);
}

/* ------------------------------------------------------------------ per-q Levenberg-Marquardt (FP64) */
/*
 * The per-bin model of the batched GPU fitter, in FP64: for one q-bin with K
 * effective-|q| nodes q_i and lags tau_j,
 *   f_j = A c_j + B,  c_j = (1/K) sum_i [ -expm1(-D q_i^2 tau_j) + E_ij alpha (1 - g(q_i vbar tau_j; Z)) ]
 * which equals the reference's modelColumn 1 - <E (1 - alpha + alpha g)> (src/motility_fit.js:417-432)
 * written without the 1 - (~1) cancellation.  Parameters theta = (A, B, alpha,
 * ln vbar, ln Z, ln D) (swimmer) or (A, B, ln D) (Brownian); chi^2 =
 * w^2 sum_j (d_j - f_j)^2 with w = 1/max_j d_j (the reference's rss weighting,
 * src/motility_fit.js:392-398,469).  Box constraints by projection.  This
 * oracle uses the exact 2D kernel (isf2d) where the GPU uses the tabulated one.
 */
static double one_minus_isf(double x, double Z, int geometry) {
  double L = x / (Z + 1);
  if (geometry == 0) {
    if (L < 1e-4) return (x * x * (Z + 2)) / (6 * (Z + 1));
    return 1 - sin(Z * atan(L)) / (Z * L * exp(0.5 * Z * log1p(L * L)));
  }
  if (L < 1e-4) return (x * x * (Z + 2)) / (4 * (Z + 1));
  return 1 - (geometry == 1 ? orc_isf2d(x, Z) : orc_isf2d_fast(x, Z));
}

/* model column and analytic/numeric derivatives; dc[k][j] for k over the
 * nonlinear parameters (alpha, ln vbar, ln Z, ln D) */
static void lm_column(int ntau, const double *tau, int K, const double *qn, double alpha, double vbar, double Z,
                      double Dd, int swim, int geometry, double vmag, double *c, double (*dc)[128]) {
  for (int j = 0; j < ntau; j++) {
    double t = tau[j], cj = 0, da = 0, dlv = 0, dlz = 0, dld = 0;
    for (int i = 0; i < K; i++) {
      double q = qn[i], s = Dd * q * q * t;
      double E = exp(-s), omE = -expm1(-s);
      if (vmag > 0) {   /* uniform drift: E = exp(-D q^2 tau) J0(q |V| tau) (modelColumn :417-432) */
        double J = orc_bessel_j0(q * vmag * t);
        omE += E * (1 - J);
        E *= J;
      }
      double x = q * vbar * t;
      double omg = swim ? one_minus_isf(x, Z, geometry) : 0;
      cj += omE + E * alpha * omg;
      da += E * omg;
      dld += s * E * (1 - alpha * omg);   /* d/dlnD of [1 - E(1 - alpha omg)] = s E (1 - alpha omg) */
      if (swim) {
        double hv = 1e-5, hz = 1e-5;
        double omgp = one_minus_isf(x * exp(hv), Z, geometry), omgm = one_minus_isf(x * exp(-hv), Z, geometry);
        dlv += E * alpha * (omgp - omgm) / (2 * hv);
        double omzp = one_minus_isf(x, Z * exp(hz), geometry), omzm = one_minus_isf(x, Z * exp(-hz), geometry);
        dlz += E * alpha * (omzp - omzm) / (2 * hz);
      }
    }
    c[j] = cj / K;
    if (dc) {
      dc[0][j] = da / K;
      dc[1][j] = dlv / K;
      dc[2][j] = dlz / K;
      dc[3][j] = dld / K;
    }
  }
}

static int cholesky_solve(int p, double *H, const double *g, double *x) {
  double Lm[8][8] = {{0}};
  for (int i = 0; i < p; i++) {
    for (int j = 0; j <= i; j++) {
      double s = H[i * p + j];
      for (int k = 0; k < j; k++) s -= Lm[i][k] * Lm[j][k];
      if (i == j) {
        if (!(s > 0)) return 0;
        Lm[i][i] = sqrt(s);
      } else {
        Lm[i][j] = s / Lm[j][j];
      }
    }
  }
  double y[8];
  for (int i = 0; i < p; i++) {
    double s = g[i];
    for (int k = 0; k < i; k++) s -= Lm[i][k] * y[k];
    y[i] = s / Lm[i][i];
  }
  for (int i = p - 1; i >= 0; i--) {
    double s = y[i];
    for (int k = i + 1; k < p; k++) s -= Lm[k][i] * x[k];
    x[i] = s / Lm[i][i];
  }
  return 1;
}

/* theta layout (6): A, B, alpha, ln vbar, ln Z, ln D.  Brownian uses A, B, ln D
 * (alpha = 0; vbar, Z unused).  lo/hi: box (6). */
static double lm_chi2(int ntau, const double *tau, const double *d, double w, int K, const double *qn,
                      const double *th, int swim, int geometry, double vmag, double *c) {
  lm_column(ntau, tau, K, qn, swim ? th[2] : 0, exp(th[3]), exp(th[4]), exp(th[5]), swim, geometry, vmag, c, NULL);
  double s = 0;
  for (int j = 0; j < ntau; j++) {
    double r = w * (d[j] - (th[0] * c[j] + th[1]));
    s += r * r;
  }
  return s;
}

/* 2-column NNLS (A, B >= 0) for a fixed column c: weighted LS in closed form */
static double nnls_ab(int ntau, const double *c, const double *d, double w, double *A, double *B) {
  double g11 = 0, g12 = 0, h1 = 0, h2 = 0, dd = 0;
  for (int j = 0; j < ntau; j++) {
    g11 += c[j] * c[j];
    g12 += c[j];
    h1 += c[j] * d[j];
    h2 += d[j];
    dd += d[j] * d[j];
  }
  double G[9] = {g11, g12, 0, g12, (double)ntau, 0, 0, 0, 0}, h[3] = {h1, h2, 0}, a[3], r;
  orc_nnls3(G, h, dd, a, &r);
  *A = a[0];
  *B = a[1];
  return r * w * w;
}

/* Grid seed, the reference's start grid applied to one bin: Brownian 49-point
 * log-D grid, then (swimmer) 16 vbar x Z{1.5, 8} x alpha{0.3, 0.8} at that D;
 * A, B by NNLS at every grid point. */
void orc_lm_seed(int ntau, const double *tau, const double *d, double w, int K, const double *qn,
                 const double *bounds4, int swim, int geometry, double vmag, double *th) {
  double c[128];
  double lo = log(bounds4[2]), hi = log(bounds4[3]);
  double best = INFINITY;
  for (int i = 0; i <= 48; i++) {
    double lD = lo + ((hi - lo) * i) / 48, A, B;
    lm_column(ntau, tau, K, qn, 0, 1, 1, exp(lD), 0, geometry, vmag, c, NULL);
    double r = nnls_ab(ntau, c, d, w, &A, &B);
    if (r < best) {
      best = r;
      th[0] = A;
      th[1] = B;
      th[2] = 0;
      th[3] = log(sqrt(bounds4[0] * bounds4[1]));
      th[4] = log(4.0);
      th[5] = lD;
    }
  }
  if (!swim) return;
  double lD = th[5];
  best = INFINITY;
  double lZs[2] = {log(1.5), log(8.0)}, as[2] = {0.3, 0.8};
  for (int i = 0; i < 16; i++) {
    double lv = log(bounds4[0]) + ((log(bounds4[1]) - log(bounds4[0])) * (i + 0.5)) / 16;
    for (int z = 0; z < 2; z++)
      for (int q = 0; q < 2; q++) {
        double A, B;
        lm_column(ntau, tau, K, qn, as[q], exp(lv), exp(lZs[z]), exp(lD), 1, geometry, vmag, c, NULL);
        double r = nnls_ab(ntau, c, d, w, &A, &B);
        if (r < best) {
          best = r;
          th[0] = A;
          th[1] = B;
          th[2] = as[q];
          th[3] = lv;
          th[4] = lZs[z];
          th[5] = lD;
        }
      }
  }
}

/* One per-q LM fit (ntau <= 128, K <= 16).  bounds4: vmin, vmax, Dmin, Dmax.
 * th: in = seed, out = fit.  Marquardt scaling D = diag(F^T F), Nielsen's
 * gain-ratio damping update.  Returns chi^2; *iters = LM steps taken. */
double orc_lm_fit(int ntau, const double *tau, const double *d, double w, int K, const double *qn,
                  const double *bounds4, int swim, int geometry, double vmag, int maxit, double *th, int *iters) {
  double lo[6] = {0, 0, 0, log(bounds4[0]), log(0.5), log(bounds4[2])};
  double hi[6] = {INFINITY, INFINITY, 1, log(bounds4[1]), log(100.0), log(bounds4[3])};
  int idx[6] = {0, 1, 2, 3, 4, 5}, p = 6;
  if (!swim) {
    idx[2] = 5;
    p = 3;
  }
  for (int k = 0; k < 6; k++) th[k] = clampd(th[k], lo[k], hi[k]);
  double c[128], dc[4][128], cn[128];
  double chi = lm_chi2(ntau, tau, d, w, K, qn, th, swim, geometry, vmag, c);
  double lam = 1e-3, nu = 2;
  int it = 0, done = 0;
  while (it < maxit && !done) {
    lm_column(ntau, tau, K, qn, swim ? th[2] : 0, exp(th[3]), exp(th[4]), exp(th[5]), swim, geometry, vmag, c, dc);
    double H[36] = {0}, g[6] = {0};
    for (int j = 0; j < ntau; j++) {
      double F[6];
      F[0] = w * c[j];
      F[1] = w;
      F[2] = w * th[0] * dc[0][j];
      F[3] = w * th[0] * dc[1][j];
      F[4] = w * th[0] * dc[2][j];
      F[5] = w * th[0] * dc[3][j];
      double r = w * (d[j] - (th[0] * c[j] + th[1]));
      for (int a = 0; a < p; a++) {
        g[a] += F[idx[a]] * r;
        for (int b = 0; b <= a; b++) H[a * p + b] += F[idx[a]] * F[idx[b]];
      }
    }
    for (int a = 0; a < p; a++)
      for (int b = 0; b < a; b++) H[b * p + a] = H[a * p + b];
    /* active set: a parameter on a bound whose gradient points out of the box
       is held for this step (its row and column drop out of the system);
       Marquardt scaling with a floor of 1e-12 x the largest diagonal */
    int frozen[6];
    double diag[6], dmax = 0;
    for (int a = 0; a < p; a++) {
      int k = idx[a];
      frozen[a] = (th[k] <= lo[k] && g[a] < 0) || (th[k] >= hi[k] && g[a] > 0);
      if (H[a * p + a] > dmax) dmax = H[a * p + a];
    }
    for (int a = 0; a < p; a++) {
      double f = 1e-12 * dmax;
      diag[a] = H[a * p + a] > f ? H[a * p + a] : (f > 1e-300 ? f : 1e-300);
    }
    /* one damped step (retried with larger damping until it lowers chi^2) */
    for (;;) {
      it++;
      double Hd[36], delta[6], ntheta[6], gf[6];
      memcpy(Hd, H, sizeof(Hd));
      for (int a = 0; a < p; a++) {
        gf[a] = frozen[a] ? 0 : g[a];
        for (int b = 0; b < p; b++)
          if (frozen[a] || frozen[b]) Hd[a * p + b] = a == b ? 1 : 0;
        if (!frozen[a]) Hd[a * p + a] += lam * diag[a];
      }
      int ok = cholesky_solve(p, Hd, gf, delta);
      if (ok) {
        memcpy(ntheta, th, sizeof(ntheta));
        double pred = 0;
        for (int a = 0; a < p; a++) {
          int k = idx[a];
          ntheta[k] = clampd(th[k] + delta[a], lo[k], hi[k]);
          double dk = ntheta[k] - th[k];
          pred += dk * (lam * diag[a] * dk + gf[a]);
        }
        double nchi = lm_chi2(ntau, tau, d, w, K, qn, ntheta, swim, geometry, vmag, cn);
        double rho = pred > 0 ? (chi - nchi) / pred : -1;
        if (nchi < chi && rho > 0) {
          double rel = (chi - nchi) / (chi > 1e-300 ? chi : 1e-300);
          memcpy(th, ntheta, sizeof(ntheta));
          chi = nchi;
          double f = 1 - (2 * rho - 1) * (2 * rho - 1) * (2 * rho - 1);
          lam *= f > 1.0 / 3 ? f : 1.0 / 3;
          if (lam < 1e-15) lam = 1e-15;
          nu = 2;
          /* converged: actual and predicted relative decrease below 1e-9 */
          if ((rel < 1e-9 && pred < 1e-9 * chi) || rel < 1e-13 || chi < 1e-300) done = 1;
          break;
        }
      }
      lam *= nu;
      nu *= 2;
      if (lam > 1e15 || it >= maxit) {
        done = 1;
        break;
      }
    }
  }
  if (iters) *iters = it;
  return chi;
}

/* Plain grid-seeded per-q LM: the 64 grid starts of orc_lm_seed are ranked by
 * chi^2 at NNLS amplitudes, LM runs from the best `nstarts`, the lowest final
 * chi^2 wins.  out6 = theta; returns chi^2; *iters = total LM steps. */
/* The start grid of orc_lm_plain, ranked: starts[s*6..] (A, B, alpha, ln v, ln Z, ln D) in increasing
 * grid chi^2 (sf[s]); returns the number of starts (64 for the swimmer, 1 for Brownian). */
int orc_lm_starts(int ntau, const double *tau, const double *d, double w, int K, const double *qn,
                  const double *bounds4, int swim, int geometry, double vmag, double *starts_out, double *sf_out);

double orc_lm_plain(int ntau, const double *tau, const double *d, double w, int K, const double *qn,
                    const double *bounds4, int swim, int geometry, double vmag, int nstarts, int maxit,
                    double *out6, int *iters) {
  double starts[64 * 6], sf[64];
  int ns = orc_lm_starts(ntau, tau, d, w, K, qn, bounds4, swim, geometry, vmag, starts, sf);
  double best = INFINITY;
  int tot = 0;
  for (int s = 0; s < (nstarts < ns ? nstarts : ns); s++) {
    double th[6];
    int it;
    memcpy(th, starts + 6 * s, sizeof(th));
    double chi = orc_lm_fit(ntau, tau, d, w, K, qn, bounds4, swim, geometry, vmag, maxit, th, &it);
    tot += it;
    if (chi < best) {
      best = chi;
      memcpy(out6, th, sizeof(th));
    }
  }
  if (iters) *iters = tot;
  return best;
}

int orc_lm_starts(int ntau, const double *tau, const double *d, double w, int K, const double *qn,
                  const double *bounds4, int swim, int geometry, double vmag, double *starts_out, double *sf_out) {
  double c[128];
  double starts[64][6], sf[64];
  int ns = 0;
  double th0[6];
  orc_lm_seed(ntau, tau, d, w, K, qn, bounds4, 0, geometry, vmag, th0);   /* Brownian seed: D, A, B */
  if (!swim) {
    starts[0][0] = th0[0];
    for (int k = 0; k < 6; k++) starts[0][k] = th0[k];
    sf[0] = 0;
    ns = 1;
  } else {
    double lZs[2] = {log(1.5), log(8.0)}, as[2] = {0.3, 0.8};
    for (int i = 0; i < 16; i++) {
      double lv = log(bounds4[0]) + ((log(bounds4[1]) - log(bounds4[0])) * (i + 0.5)) / 16;
      for (int z = 0; z < 2; z++)
        for (int q = 0; q < 2; q++) {
          double A, B;
          lm_column(ntau, tau, K, qn, as[q], exp(lv), exp(lZs[z]), exp(th0[5]), 1, geometry, vmag, c, NULL);
          sf[ns] = nnls_ab(ntau, c, d, w, &A, &B);
          double s[6] = {A, B, as[q], lv, lZs[z], th0[5]};
          memcpy(starts[ns], s, sizeof(s));
          ns++;
        }
    }
  }
  int idx[64];
  for (int i = 0; i < ns; i++) idx[i] = i;
  for (int i = 1; i < ns; i++) {
    int v = idx[i], k = i - 1;
    while (k >= 0 && sf[idx[k]] > sf[v]) {
      idx[k + 1] = idx[k];
      k--;
    }
    idx[k + 1] = v;
  }
  for (int s = 0; s < ns; s++) {
    memcpy(starts_out + 6 * s, starts[idx[s]], sizeof(double) * 6);
    sf_out[s] = sf[idx[s]];
  }
  return ns;
}

/* ---------------------------------------------------------------- scenario simulator
 * test/sim.mjs:13-205 (simulate, makeRng, mulberry32): Gaussian blobs on a
 * periodic field, Schulz-speed swimmers (3D-isotropic headings seen in
 * projection, or in-plane), Brownian and fixed particles, a uniform drift,
 * additive Gaussian camera noise, 8-bit quantisation, dropped frames.  The
 * draws are taken in the reference's order from the same generators, so the
 * movies are the reference's up to libm's last-ulp differences in exp/log. */
typedef struct {
  uint32_t a;
  int have_spare;
  double spare;
} orc_rng;

static double rng_u(orc_rng *r) {   /* mulberry32 */
  r->a += 0x6D2B79F5u;
  uint32_t t = r->a;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}

static double rng_normal(orc_rng *r) {   /* polar Box-Muller with one spare */
  if (r->have_spare) {
    r->have_spare = 0;
    return r->spare;
  }
  double x, y, s;
  do {
    x = 2 * rng_u(r) - 1;
    y = 2 * rng_u(r) - 1;
    s = x * x + y * y;
  } while (s >= 1 || s == 0);
  const double f = sqrt(-2 * log(s) / s);
  r->spare = y * f;
  r->have_spare = 1;
  return x * f;
}

static double rng_gamma(orc_rng *r, double k, double theta) {   /* Marsaglia-Tsang */
  if (k < 1) return rng_gamma(r, k + 1, theta) * pow(rng_u(r), 1 / k);
  const double d = k - 1.0 / 3, c = 1 / sqrt(9 * d);
  for (;;) {
    double x, v;
    do {
      x = rng_normal(r);
      v = 1 + c * x;
    } while (v <= 0);
    v = v * v * v;
    const double uu = rng_u(r);
    if (uu < 1 - 0.0331 * x * x * x * x) return d * v * theta;
    if (log(uu) < 0.5 * x * x + d * (1 - v + log(v))) return d * v * theta;
  }
}

typedef struct {
  int N, T, seed, n_swim, swim_2d, n_passive, n_fixed, drop_exact;
  double fps, dx_um, vbar, sigma_v, D_um2s, drift_speed, drift_angle_deg, drop_frac, blob_sigma, amp, background,
      noise;
} orc_sim_params;

/* frames[T * N * N] u8, valid[T]; returns the mean swimmer speed */
double orc_simulate(const orc_sim_params *p, uint8_t *frames, uint8_t *valid) {
  const int N = p->N, T = p->T;
  const double dt = 1 / p->fps, Lf = N * p->dx_um, sig = p->blob_sigma;
  orc_rng rng = {(uint32_t)p->seed, 0, 0};
  const double Vx = p->drift_speed * cos(p->drift_angle_deg * ORC_PI / 180);
  const double Vy = p->drift_speed * sin(p->drift_angle_deg * ORC_PI / 180);
  const int np_ = p->n_swim + p->n_passive + p->n_fixed;
  double *px = malloc(sizeof(double) * np_), *py = malloc(sizeof(double) * np_);
  double *vx = calloc(np_, sizeof(double)), *vy = calloc(np_, sizeof(double));
  int *mob = malloc(sizeof(int) * np_);
  double speed_sum = 0;
  int k = 0;
  if (p->n_swim > 0) {
    const double Z = (p->vbar / p->sigma_v) * (p->vbar / p->sigma_v) - 1;
    for (int i = 0; i < p->n_swim; i++, k++) {
      const double v = rng_gamma(&rng, Z + 1, p->vbar / (Z + 1));
      speed_sum += v;
      const double phi = 2 * ORC_PI * rng_u(&rng);
      double s = 1;
      if (!p->swim_2d) {
        const double cz = 2 * rng_u(&rng) - 1;
        s = sqrt(1 - cz * cz);
      }
      px[k] = rng_u(&rng) * Lf;
      py[k] = rng_u(&rng) * Lf;
      vx[k] = v * s * cos(phi);
      vy[k] = v * s * sin(phi);
      mob[k] = 1;
    }
  }
  for (int i = 0; i < p->n_passive; i++, k++) {
    px[k] = rng_u(&rng) * Lf;
    py[k] = rng_u(&rng) * Lf;
    mob[k] = 1;
  }
  for (int i = 0; i < p->n_fixed; i++, k++) {
    px[k] = rng_u(&rng) * Lf;
    py[k] = rng_u(&rng) * Lf;
    mob[k] = 0;
  }
  const int R = (int)ceil(4 * sig), W = 2 * R + 2;
  double *gx = malloc(sizeof(double) * W), *gy = malloc(sizeof(double) * W);
  double *img = malloc(sizeof(double) * N * N);
  const double step = sqrt(2 * p->D_um2s * dt);
  for (int t = 0; t < T; t++) {
    for (long i = 0; i < (long)N * N; i++) img[i] = p->background;
    for (int q = 0; q < np_; q++) {
      const double cx = px[q] / p->dx_um, cy = py[q] / p->dx_um;
      const int x0 = (int)floor(cx) - R, y0 = (int)floor(cy) - R;
      for (int i = 0; i < W; i++) {
        const double ddx = x0 + i - cx, ddy = y0 + i - cy;
        gx[i] = exp(-ddx * ddx / (2 * sig * sig));
        gy[i] = p->amp * exp(-ddy * ddy / (2 * sig * sig));
      }
      for (int j = 0; j < W; j++) {
        const int yy = (((y0 + j) % N) + N) % N;
        for (int i = 0; i < W; i++) {
          const int xx = (((x0 + i) % N) + N) % N;
          img[(long)yy * N + xx] += gx[i] * gy[j];
        }
      }
    }
    uint8_t *f = frames + (long)t * N * N;
    for (long i = 0; i < (long)N * N; i++) {
      const double v = js_round((img[i] + p->noise * rng_normal(&rng)) * 255);
      f[i] = v < 0 ? 0 : v > 255 ? 255 : (uint8_t)v;
    }
    for (int q = 0; q < np_; q++) {
      if (!mob[q]) continue;
      px[q] += (vx[q] + Vx) * dt + step * rng_normal(&rng);
      py[q] += (vy[q] + Vy) * dt + step * rng_normal(&rng);
      px[q] = fmod(fmod(px[q], Lf) + Lf, Lf);
      py[q] = fmod(fmod(py[q], Lf) + Lf, Lf);
    }
  }
  for (int t = 0; t < T; t++) valid[t] = 1;
  orc_rng drng = {(uint32_t)(p->seed * 7919 + 13), 0, 0};
  if (p->drop_exact) {
    int *idx = malloc(sizeof(int) * (T - 1));
    for (int i = 0; i < T - 1; i++) idx[i] = i + 1;
    for (int i = T - 2; i > 0; i--) {
      const int j = (int)floor(rng_u(&drng) * (i + 1));
      const int tmp = idx[i];
      idx[i] = idx[j];
      idx[j] = tmp;
    }
    for (int i = 0; i < p->drop_exact; i++) valid[idx[i]] = 0;
    free(idx);
  } else if (p->drop_frac > 0) {
    for (int t = 1; t < T; t++)
      if (rng_u(&drng) < p->drop_frac) valid[t] = 0;
  }
  free(px), free(py), free(vx), free(vy), free(mob), free(gx), free(gy), free(img);
  return p->n_swim ? speed_sum / p->n_swim : 0;
}
