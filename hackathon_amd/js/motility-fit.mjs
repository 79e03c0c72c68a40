// Per-q motility fits for the TypeScript-facing API (ES module, ES2019).
//
// prepare / bounds / lagGrid keep the reference's data preparation
// (motility-ddm src/motility_fit.js:358-369, :383-401, :480-491): usable bins
// (nModes >= 3, positive maximum), the quasi-logarithmic lag grid, q and the
// q-nodes from kCenter / kNodes, weight w = 1 / max D.  analyzeMotilityLM runs
// every usable bin of a Radial record as one batched Levenberg-Marquardt launch
// on the MI355X (libddm.so ddm_fit through the N-API addon) and returns the
// fitted per-q parameters.  estimateDrift turns the radial record's drift sums
// into a uniform drift velocity (src/motility_fit.js:671-725); analyzeMotility
// is the reference's verdict layer (:808-990) run by the library's host code.
// Types are declared in motility-fit.d.ts.
import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import path from 'path';
import { isMainThread, parentPort } from 'worker_threads';

const here = path.dirname(fileURLToPath(import.meta.url));
const require = createRequire(import.meta.url);
const addon = require(path.join(here, '..', 'napi', 'ddm_napi.node'));

export const MODEL = Object.freeze({ BROWNIAN: 0, SWIMMER_3D: 1, SWIMMER_2D: 2 });
const SEED = Object.freeze({ GRID: 0, GIVEN: 1, NONLINEAR: 2 });
/** GATES.SWIM_MIN_UM_S (src/motility_fit.js:37) */
export const SWIM_MIN_UM_S = 2;
const pick = (a, b) => (a === undefined || a === null ? b : a);

/** Lags 1..10, then steps of x1.08 (src/motility_fit.js:358-369). */
export function lagGrid(L) {
  const lags = [];
  for (let t = 1; t <= Math.min(10, L - 1); t++) lags.push(t);
  let x = 10;
  for (;;) {
    x *= 1.08;
    const t = Math.round(x);
    if (t > L - 1) break;
    if (t > lags[lags.length - 1]) lags.push(t);
  }
  return lags;
}

/** prepare(radial, dxUm, fps) (src/motility_fit.js:383-401). */
export function prepare(radial, dxUm, fps) {
  const { N, NB, L } = radial;
  const dt = 1 / fps;
  const lags = lagGrid(L);
  const tau = Float64Array.from(lags, (t) => t * dt);
  const qs = (2 * Math.PI) / (N * dxUm);
  const K = radial.kNodes && radial.kNodes.length ? radial.kNodes.length / NB : 0;
  const bins = [];
  for (let b = 0; b < NB; b++) {
    if (radial.nModes[b] < 3) continue;
    const d = Float64Array.from(lags, (t) => radial.D[b * L + t]);
    let scale = -Infinity;
    for (const v of d) scale = Math.max(scale, v);
    if (!(scale > 0) || !Number.isFinite(scale)) continue;
    const q = qs * radial.kCenter[b];
    const qn = K ? Float64Array.from(radial.kNodes.subarray(b * K, (b + 1) * K), (k) => k * qs) : Float64Array.of(q);
    bins.push({ b, q, qn, d, w: 1 / scale });
  }
  return { bins, lags, tau, dt, L, tauMax: (L - 1) * dt };
}

/** vmin, vmax, Dmin, Dmax (src/motility_fit.js:480-491). */
export function bounds(pr) {
  let qmin = Infinity, qmax = -Infinity;
  for (const b of pr.bins) { qmin = Math.min(qmin, b.q); qmax = Math.max(qmax, b.q); }
  return [Math.max(SWIM_MIN_UM_S, 0.3 / (qmax * pr.tauMax)), 3 / (qmin * pr.dt),
    0.01 / (qmax * qmax * pr.tauMax), 10 / (qmin * qmin * pr.dt)];
}

let fitter = null;

/**
 * Per-q LM fits of every usable bin of one Radial record.
 * opt: {dxUm, fps, model (MODEL.*), nstarts, maxIter, seeds (Float64Array 6 per usable bin), nnlsAmplitudes,
 *      vmag (a uniform drift speed in um/s, e.g. estimateDrift's mag: E = exp(-D q^2 tau) J0(q vmag tau))}
 * -> {pr, fits: [{b, q, A, B, alpha, vbar, Z, D, chi2, iters}], kernelMs}
 */
export function analyzeMotilityLM(radial, opt = {}) {
  const pr = prepare(radial, pick(opt.dxUm, 1), pick(opt.fps, 30));
  const n = pr.bins.length;
  if (!n) return { pr, fits: [], kernelMs: 0 };
  const ntauMax = pr.tau.length;
  const K = Math.max(...pr.bins.map((b) => b.qn.length));
  const bd = bounds(pr);
  const batch = {
    nfits: n, ntauMax, knodes: K, model: pick(opt.model, MODEL.SWIMMER_3D),
    seedMode: opt.seeds ? (opt.nnlsAmplitudes ? SEED.NONLINEAR : SEED.GIVEN) : SEED.GRID,
    nstarts: pick(opt.nstarts, 3), maxIter: pick(opt.maxIter, 100),
    ntau: new Int32Array(n).fill(ntauMax), tau: new Float64Array(n * ntauMax), data: new Float64Array(n * ntauMax),
    weight: new Float64Array(n), qnodes: new Float64Array(n * K), bounds: new Float64Array(4 * n),
    seeds: opt.seeds || null,
    vmag: opt.vmag > 0 ? new Float64Array(n).fill(opt.vmag) : null,
  };
  pr.bins.forEach((bin, i) => {
    batch.tau.set(pr.tau, i * ntauMax);
    batch.data.set(bin.d, i * ntauMax);
    batch.weight[i] = bin.w;
    for (let k = 0; k < K; k++) batch.qnodes[i * K + k] = bin.qn[k % bin.qn.length];
    batch.bounds.set(bd, 4 * i);
  });
  if (!fitter) fitter = addon.fitterCreate();
  const r = addon.fit(fitter, batch);
  const fits = pr.bins.map((bin, i) => {
    const t = r.theta.subarray(6 * i, 6 * i + 6);
    return { b: bin.b, q: bin.q, A: t[0], B: t[1], alpha: t[2], vbar: Math.exp(t[3]), Z: Math.exp(t[4]),
      D: Math.exp(t[5]), chi2: r.chi2[i], iters: r.iters[i] };
  });
  return { pr, fits, kernelMs: r.kernelMs };
}

// ---------------------------------------------------------------- drift

/** The drift gates of src/motility_fit.js:28-38 used here. */
export const GATES = Object.freeze({ FLOW_SIGMA: 3, FLOW_MIN_UM_S: 2, DRIFT_ISF_MAX: 0.2 });

/** 3D Schulz-averaged swimmer self-ISF (src/motility_fit.js:135-144). */
export function isf3d(x, Z) {
  const u = x / (Z + 1);
  if (u < 1e-4) return 1 - (x * x * (Z + 2)) / (6 * (Z + 1));
  return Math.sin(Z * Math.atan(u)) / (Z * u * Math.exp(0.5 * Z * Math.log1p(u * u)));
}

// P_nu(c), |c| <= 1: trapezoid rule on Laplace's integral (src/motility_fit.js:120-132)
function legendreP(nu, c) {
  const s = Math.sqrt(Math.max(0, 1 - c * c));
  const n = 16 + Math.ceil(nu * Math.atan2(s, c));
  let acc = 0;
  for (let i = 0; i <= n; i++) {
    const im = s * Math.cos((Math.PI * i) / n);
    const v = Math.exp(0.5 * nu * Math.log(c * c + im * im)) * Math.cos(nu * Math.atan2(im, c));
    acc += i === 0 || i === n ? 0.5 * v : v;
  }
  return acc / n;
}

/** In-plane Schulz-averaged swimmer self-ISF (src/motility_fit.js:146-159). */
export function isf2d(x, Z) {
  const u = x / (Z + 1);
  if (u < 1e-4) return 1 - (x * x * (Z + 2)) / (4 * (Z + 1));
  const w = 1 + u * u;
  const lnPre = -0.5 * (Z + 1) * Math.log(w);
  return lnPre < -25 ? 0 : legendreP(Z, 1 / Math.sqrt(w)) * Math.exp(lnPre);
}

/**
 * estimateDrift(radial, pr, opt, swimmer) (src/motility_fit.js:671-725): uniform
 * drift V from the radial record's phase sums (readRadial with drift on), by
 * weighted least squares of phi = -q.V dt over the usable bins' lag pairs; with
 * a swimmer component {vbar, Z}, only pairs where its ISF is below
 * GATES.DRIFT_ISF_MAX.  opt: {dxUm = 1, geometry = '3d'}.
 */
export function estimateDrift(radial, pr, opt = {}, swimmer = null) {
  const { L, N } = radial;
  if (!radial.drift || radial.drift.length < radial.NB * L * 8) throw new Error('estimateDrift: Radial record has no drift sums');
  const g = pick(opt.geometry, '3d') === '2d' ? isf2d : isf3d;
  const S = new Float64Array(8);
  let modesUsed = 0;
  let pairs = 0;
  for (const bin of pr.bins) {
    let used = false;
    for (let tau = 1; tau < L - 1; tau++) {
      if (swimmer && !(Math.abs(g(bin.q * swimmer.vbar * tau * pr.dt, swimmer.Z)) < GATES.DRIFT_ISF_MAX)) continue;
      const o = (bin.b * L + tau) * 8;
      if (!(radial.drift[o] > 0)) continue;
      for (let c = 0; c < 8; c++) S[c] += radial.drift[o + c];
      used = true;
      pairs++;
    }
    if (used) modesUsed += radial.nModes[bin.b];
  }
  const [Sw, Sxx, Sxy, Syy, Rx, Ry, Rpp, n] = S;
  const det = Sxx * Syy - Sxy * Sxy;
  if (n < 10 || !(det > 1e-12 * Sxx * Syy) || !(Sw > 0)) {
    return { Vx: 0, Vy: 0, mag: 0, sigma: Infinity, angleDeg: 0, significant: false, pairs, samples: n };
  }
  const bx = (Syy * Rx - Sxy * Ry) / det;
  const by = (Sxx * Ry - Sxy * Rx) / det;
  const rss = Math.max(Rpp - bx * Rx - by * Ry, 0);
  const s2 = (rss / (n - 2)) * Math.max(1, n / Math.max(modesUsed, 1));
  const conv = -(N * pick(opt.dxUm, 1)) / (2 * Math.PI * pr.dt);
  const Vx = conv * bx;
  const Vy = conv * by;
  const mag = Math.hypot(Vx, Vy);
  // variance along V (isotropic mean when V = 0) of the covariance s2 [Sxx Sxy; Sxy Syy]^-1
  const cxx = (s2 * Syy) / det, cyy = (s2 * Sxx) / det, cxy = (-s2 * Sxy) / det;
  const ux = mag > 0 ? Vx / mag : 0, uy = mag > 0 ? Vy / mag : 0;
  const varAlong = conv * conv * (mag > 0 ? ux * ux * cxx + 2 * ux * uy * cxy + uy * uy * cyy : (cxx + cyy) / 2);
  const sigma = Math.sqrt(Math.max(varAlong, 0));
  return { Vx, Vy, mag, sigma, angleDeg: (Math.atan2(Vy, Vx) * 180) / Math.PI,
    significant: mag > GATES.FLOW_SIGMA * sigma, pairs, samples: n };
}

// ---------------------------------------------------------------- verdict

export const VERDICTS = Object.freeze(['MOTILE', 'NOT_MOTILE', 'DIRECTED_TRANSPORT', 'NO_DYNAMICS', 'UNLOCKED_EXPOSURE']);
const GATE_TEXT = {
  0: ['G0', 'Usable q-bins', '>= 4'],
  1: ['G1', 'Temporal structure index', '>= 0.15'],
  2: ['G2', 'Uniform flow', 'not (|V| > 3 sigma and |V| > 2 um/s)'],
  3: ['G3', 'Swimmer vs Brownian Delta BIC', '> 10'],
  4: ['G4', 'Motile contrast share alpha', '> 0.05, swimmer decay resolved in >= 1 bin'],
  5: ['G5', 'Exposure locked', 'manual exposure readback'],
};
const f64 = (a) => (a instanceof Float64Array ? a : Float64Array.from(a));

/**
 * analyzeMotility(radial, opt) (src/motility_fit.js:808-990): fits, drift,
 * gates and verdict of one Radial record (readRadial with its drift sums), in
 * the library's host code.  opt: {dxUm, fps, geometry = '3d', exposureLocked = true}.
 * Returns the reference's MotilityResult fields.
 */
export function analyzeMotility(radial, opt = {}) {
  const geometry = pick(opt.geometry, '3d');
  if (geometry !== '3d' && geometry !== '2d') throw new Error(`analyzeMotility: geometry must be '3d' or '2d'`);
  const r = addon.analyzeMotility({
    N: radial.N, NB: radial.NB, L: radial.L, kCenter: f64(radial.kCenter),
    kNodes: radial.kNodes && radial.kNodes.length ? f64(radial.kNodes) : undefined,
    nModes: radial.nModes instanceof Uint32Array ? radial.nModes : Uint32Array.from(radial.nModes),
    D: f64(radial.D), drift: f64(radial.drift || []),
  }, { dxUm: pick(opt.dxUm, 1), fps: pick(opt.fps, 30), geometry: geometry === '2d' ? 1 : 0,
    exposureLocked: pick(opt.exposureLocked, true) ? 1 : 0 });
  const gates = [];
  for (let i = 0; i < r.ngates; i++) {
    const g = r.gates.subarray(5 * i, 5 * i + 5);
    const [id, name, threshold] = GATE_TEXT[g[0]];
    const o = { id, name, value: g[2], threshold, pass: g[1] !== 0 };
    if (g[0] === 2) o.sigma = g[3];
    if (g[0] === 4) o.resolvedBins = g[4];
    gates.push(o);
  }
  const out = { verdict: VERDICTS[r.verdict], physicsVerdict: VERDICTS[r.physicsVerdict], gates, TSI: r.TSI };
  if (r.insufficient) return out;
  const nb = r.nbins, ns = r.nscaling;
  const amps = (a, alpha) => Array.from({ length: nb }, (_, i) =>
    ({ A: a[2 * i], B: a[2 * i + 1], W: alpha * a[2 * i], P: (1 - alpha) * a[2 * i] }));
  const lags = lagGrid(radial.L);
  const dt = 1 / pick(opt.fps, 30);
  return Object.assign(out, {
    alpha: r.alpha, alphaResolved: r.alphaResolved !== 0, resolvedBins: r.resolvedBins, vbar: r.vbar,
    sigmaV: r.sigmaV, Z: r.Z, Ddiff: r.Ddiff, DdiffBrownian: r.DdiffBrownian, dBIC: r.dBIC,
    drift: { Vx: r.driftVx, Vy: r.driftVy, mag: r.driftMag, sigma: r.driftSigma, angleDeg: r.driftAngleDeg,
      significant: r.driftSignificant !== 0, pairs: r.driftPairs, samples: r.driftSamples },
    mu: r.mu,
    scaling: { q: Array.from(r.scalingQ.subarray(0, ns)), tauHalf: Array.from(r.scalingTauHalf.subarray(0, ns)), mu: r.mu },
    geometry,
    fit: {
      swimmer: { vbar: r.vbar, Z: r.Z, Dd: r.Ddiff, Vmag: r.swimVmag, alpha: r.alpha, rss: r.swimRss, amps: amps(r.swimAmps, r.alpha) },
      brownian: { vbar: 0, Z: 1, Dd: r.DdiffBrownian, Vmag: r.brownVmag, alpha: 0, rss: r.brownRss, amps: amps(r.brownAmps, 0) },
    },
    q: Array.from(r.q.subarray(0, nb)), bins: Array.from(r.bins.subarray(0, nb)), lags, tau: lags.map((l) => l * dt),
    nData: r.nData, nEff: r.nEff, residualRho: r.residualRho,
  });
}

// ---------------------------------------------------------------- worker entry
// The reference runs analyzeMotility in a module Worker built from its fitting
// module, so a fit never blocks frame delivery (src/motility_fit.js:992-1005,
// caller src/main.ts:325-353): {id, radial, opt} in, {id, result} or
// {id, error: stack text} out.  The same protocol in a Node worker thread
// started on this file (new Worker(new URL('./motility-fit.mjs', ...)));
// inert on the main thread.
if (!isMainThread && parentPort) {
  parentPort.on('message', (msg) => {
    const { id, radial, opt } = msg || {};
    try {
      parentPort.postMessage({ id, result: analyzeMotility(radial, opt) });
    } catch (e) {
      parentPort.postMessage({ id, error: String((e && e.stack) || e) });
    }
  });
}
