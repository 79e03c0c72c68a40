// Per-q motility fits for the TypeScript-facing API (ES module, ES2019).
//
// prepare / bounds / lagGrid give the reference's data preparation
// (motility-ddm src/motility_fit.js:358-369, :383-401, :480-491) from the
// library's own routines (the lag grid and the usable-bin rule are
// ddm_lag_grid / ddm_usable_bins, shared with the verdict layer and Python).
// analyzeMotilityLM runs every usable bin of a Radial record as one batched
// Levenberg-Marquardt launch on the MI355X (libddm.so ddm_fit through the
// N-API addon) and returns the fitted per-q parameters.  estimateDrift
// (src/motility_fit.js:671-725) and analyzeMotility (:808-990) are the
// library's host code.  Types are declared in motility-fit.d.ts.
import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import path from 'path';

const here = path.dirname(fileURLToPath(import.meta.url));
const require = createRequire(import.meta.url);
const addon = require(path.join(here, '..', 'napi', 'ddm_napi.node'));

export const MODEL = Object.freeze({ BROWNIAN: 0, SWIMMER_3D: 1, SWIMMER_2D: 2 });
const SEED = Object.freeze({ GRID: 0, GIVEN: 1, NONLINEAR: 2 });
/** GATES.SWIM_MIN_UM_S (src/motility_fit.js:37) */
export const SWIM_MIN_UM_S = 2;
const pick = (a, b) => (a === undefined || a === null ? b : a);
const f64 = (a) => (a instanceof Float64Array ? a : Float64Array.from(a));
const GEOMETRY = { '3d': 0, '2d': 1 };

// the record's typed views as the addon takes them
function record(radial) {
  return {
    N: radial.N, NB: radial.NB, L: radial.L, kCenter: f64(radial.kCenter),
    kNodes: radial.kNodes && radial.kNodes.length ? f64(radial.kNodes) : undefined,
    nModes: radial.nModes instanceof Uint32Array ? radial.nModes : Uint32Array.from(radial.nModes),
    D: f64(radial.D), drift: radial.drift && radial.drift.length ? f64(radial.drift) : undefined,
  };
}

/** The lags the fits sample (src/motility_fit.js:358-369): 1..10, then roughly 8 % apart, below L. */
export function lagGrid(L) { return Array.from(addon.lagGrid(L)); }

/**
 * prepare(radial, dxUm, fps) (src/motility_fit.js:383-401): the usable bins (the library's rule) with
 * their data on the lag grid, q = 2 pi kCenter / (N dx), q-nodes from kNodes and weight 1 / max D.
 */
export function prepare(radial, dxUm, fps) {
  const dt = 1 / fps;
  const lags = lagGrid(radial.L);
  const tau = Float64Array.from(lags, (l) => l * dt);
  const perK = (2 * Math.PI) / (radial.N * dxUm);
  const K = radial.kNodes && radial.kNodes.length ? radial.kNodes.length / radial.NB : 0;
  const bins = Array.from(addon.usableBins(record(radial)), (b) => {
    const d = Float64Array.from(lags, (l) => radial.D[b * radial.L + l]);
    const qn = K ? Float64Array.from({ length: K }, (_, k) => radial.kNodes[b * K + k] * perK)
      : Float64Array.of(radial.kCenter[b] * perK);
    return { b, q: radial.kCenter[b] * perK, qn, d, w: 1 / d.reduce((m, v) => (v > m ? v : m), -Infinity) };
  });
  return { bins, lags, tau, dt, L: radial.L, tauMax: (radial.L - 1) * dt };
}

/** vmin, vmax, Dmin, Dmax (src/motility_fit.js:480-491). */
export function bounds(pr) {
  const qs = pr.bins.map((b) => b.q);
  const qmin = Math.min(...qs), qmax = Math.max(...qs);
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
    nstarts: pick(opt.nstarts, 3), maxIter: pick(opt.maxIter, 400),
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

const isfOf = (geometry) => (x, Z) => addon.swimmerIsf(geometry, Float64Array.of(x), Z)[0];
/** The 3D Schulz-averaged swimmer self-ISF of the fits (src/motility_fit.js:135-144), library code. */
export const isf3d = isfOf(GEOMETRY['3d']);
/** The in-plane swimmer self-ISF of the fits (isf2dFast, src/motility_fit.js:168-219), library code. */
export const isf2d = isfOf(GEOMETRY['2d']);

/**
 * estimateDrift(radial, pr, opt, swimmer) (src/motility_fit.js:671-725): the uniform drift V from the
 * radial record's phase sums (readRadial with drift on), in the library's host code; a swimmer component
 * {vbar, Z} gates the lag pairs by its ISF.  opt: {dxUm = 1, geometry = '3d'}; pr gives the frame interval.
 */
export function estimateDrift(radial, pr, opt = {}, swimmer = null) {
  if (!radial.drift || radial.drift.length < radial.NB * radial.L * 8) {
    throw new Error('estimateDrift: Radial record has no drift sums');
  }
  const geometry = pick(opt.geometry, '3d');
  if (!(geometry in GEOMETRY)) throw new Error(`estimateDrift: geometry must be '3d' or '2d'`);
  return addon.estimateDrift(record(radial), { dxUm: pick(opt.dxUm, 1), fps: 1 / pr.dt, geometry: GEOMETRY[geometry] },
    swimmer ? { vbar: swimmer.vbar, Z: swimmer.Z } : null);
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
/**
 * analyzeMotility(radial, opt) (src/motility_fit.js:808-990): fits, drift,
 * gates and verdict of one Radial record (readRadial with its drift sums), in
 * the library's host code.  opt: {dxUm, fps, geometry = '3d', exposureLocked = true}.
 * Returns the reference's MotilityResult fields.
 */
export function analyzeMotility(radial, opt = {}) {
  const geometry = pick(opt.geometry, '3d');
  if (geometry !== '3d' && geometry !== '2d') throw new Error(`analyzeMotility: geometry must be '3d' or '2d'`);
  const r = addon.analyzeMotility(Object.assign(record(radial), { drift: f64(radial.drift || []) }), { dxUm: pick(opt.dxUm, 1), fps: pick(opt.fps, 30), geometry: geometry === '2d' ? 1 : 0,
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
