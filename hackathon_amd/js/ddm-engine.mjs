// TypeScript-facing facade of the MI355X DDM engine (ES module, ES2019).
//
// Keeps the method names, argument meaning, return values and thrown errors of
// the reference's WebGPU engine class (motility-ddm src/ddm-engine.ts:62-426)
// so callers switch by changing the import.  Every call goes through the N-API
// addon (../napi/ddm_napi.node) to libddm.so: HIP kernels on one MI355X.  Types
// are declared in ddm-engine.d.ts.
import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import path from 'path';

const here = path.dirname(fileURLToPath(import.meta.url));
const require = createRequire(import.meta.url);
const addon = require(path.join(here, '..', 'napi', 'ddm_napi.node'));

/** The reference geometry (src/ddm_ref.js:18-31). */
export const DDM = Object.freeze({ N: 256, T: 256, L: 128, NB: 32, KMIN: 4, KMAX: 100, KNODES: 4, DRIFT_STRIDE: 8 });

/** a ?? b for Node 12 (no nullish coalescing). */
const pick = (a, b) => (a === undefined || a === null ? b : a);

export class DdmEngine {
  constructor(handle, opts) {
    this._h = handle;
    this._destroyed = false;
    this._temporalBusy = false;
    this._onError = opts.onError;
    const s = addon.sizes(handle);
    this.sizes = s;
    this.N = s.N;
    this.T = s.T;
    this.L = s.L;
    this.NB = s.NB;
    /** makeBins() tables of this geometry (src/ddm_ref.js:84-170). */
    this.bins = Object.assign({ N: s.N, NB: s.NB, kmin: opts.kmin, kmax: opts.kmax }, addon.bins(handle));
    this.adapterInfo = { vendor: 'AMD', architecture: 'gfx950', description: 'AMD Instinct MI355X (HIP)' };
    /** Frames dropped because of back-pressure: ingest is stream-ordered and never drops. */
    this.throttled = 0;
  }

  /**
   * DdmEngine.create (src/ddm-engine.ts:125-149).  opts: geometry {N, T, L, NB,
   * kmin, kmax, hypotMode, knodes} (defaults = the reference's DDM), plus the
   * reference's onLost/onError callbacks.  shaderCode/powerPreference are accepted
   * and ignored (the kernels are compiled HIP).
   */
  static async create(opts = {}) {
    const g = {
      N: pick(opts.N, DDM.N), T: pick(opts.T, DDM.T), L: pick(opts.L, DDM.L), NB: pick(opts.NB, DDM.NB),
      kmin: pick(opts.kmin, DDM.KMIN), kmax: pick(opts.kmax, DDM.KMAX), hypotMode: pick(opts.hypotMode, 0),
      knodes: pick(opts.knodes, DDM.KNODES),
    };
    const h = addon.create(g);
    return new DdmEngine(h, Object.assign({}, opts, g));
  }

  get pendingSubmits() { return 0; }
  get saturated() { return false; }
  async waitForCapacity() {}

  /** ingestLuma (src/ddm-engine.ts:308-324): one N*N Float32Array frame into slot % T. */
  ingestLuma(luma, slot) {
    if (this._destroyed) { this.throttled++; return false; }
    if (luma.length !== this.N * this.N) throw new Error(`ingestLuma expects ${this.N * this.N} values`);
    return addon.ingest(this._h, luma, 1, slot % this.T);
  }

  /** Frame stack in: nframes frames (Uint8Array | Uint16Array | Float32Array) into slots firstSlot + i. */
  ingestFrames(frames, nframes, firstSlot = 0) {
    if (this._destroyed) { this.throttled++; return false; }
    return addon.ingest(this._h, frames, nframes, firstSlot);
  }

  /** computeTemporal (src/ddm-engine.ts:331-365). */
  async computeTemporal(order, mask, opts = {}) {
    if (this._temporalBusy) throw new Error('computeTemporal is already running');
    this._temporalBusy = true;
    try {
      addon.computeTemporal(this._h, order, mask);
      addon.idle(this._h);
    } finally {
      this._temporalBusy = false;
    }
  }

  /**
   * readRadial (src/ddm-engine.ts:391-395): the Radial record of src/ddm_ref.js:336-348, phase-drift
   * sums included as in the reference (opts.drift = false skips them: they take a per-mode pass).
   */
  async readRadial(opts = {}) {
    const r = addon.readRadial(this._h, pick(opts.drift, true));
    return {
      N: this.N, NB: this.NB, L: this.L,
      kCenter: Float64Array.from(this.bins.kCenter),
      kNodes: Float64Array.from(this.bins.kNodes),
      nModes: Uint32Array.from(this.bins.nModes),
      D: Float64Array.from(r.D),
      P: Float64Array.from(r.P),
      drift: Float64Array.from(r.drift),
    };
  }

  /** readTemporalFull (src/ddm-engine.ts:403-408): per-mode {D, C, P} of every mode (opt-in, computed on demand). */
  async readTemporalFull() { return addon.readTemporalFull(this._h); }

  /** readHistory (src/ddm-engine.ts:397-400): hist[(mode*T + slot)*2 + {0,1}]. */
  async readHistory() { return addon.readHistory(this._h); }

  async idle() { addon.idle(this._h); }

  destroy() {
    if (this._destroyed) return;
    this._destroyed = true;
    addon.destroy(this._h);
  }
}

/** mulberry32 seeded uniform RNG on [0, 1) (the reference's test/sim.mjs:14-23 generator). */
export function mulberry32(seed) {
  let a = seed >>> 0;
  return function () {
    a = (a + 0x6d2b79f5) >>> 0;
    let t = a;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

/** shuffleOrder (src/ddm_ref.js:504-519): seeded Fisher-Yates over the (slot, mask) pairs. */
export function shuffleOrder(o, rand) {
  const order = Uint32Array.from(o.order);
  const mask = Float32Array.from(o.mask);
  for (let i = order.length - 1; i > 0; i--) {
    const j = Math.floor(rand() * (i + 1));
    const to = order[i]; order[i] = order[j]; order[j] = to;
    const tm = mask[i]; mask[i] = mask[j]; mask[j] = tm;
  }
  return { order, mask };
}

/** linearOrder (src/ddm_ref.js:493-502). */
export function linearOrder(valid) {
  const T = valid.length;
  const order = new Uint32Array(T);
  const mask = new Float32Array(T);
  for (let j = 0; j < T; j++) { order[j] = j; mask[j] = valid[j] ? 1 : 0; }
  return { order, mask };
}
