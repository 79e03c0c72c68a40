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

  /** readMean (src/ddm-engine.ts:410-414): mean luma of the last frame ingested (u8/255, u16/65535). */
  async readMean() { return addon.readMean(this._h); }

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

/** shuffleOrder (src/ddm_ref.js:510-519): seeded Fisher-Yates over the (slot, mask) pairs. */
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

/**
 * FrameClock (src/camera.ts:306-425): frame n = round((t - t0) / dt) goes to
 * slot n mod T of the history ring; each slot remembers the frame index it
 * holds, so snapshot() masks dropped, late, duplicated and overwritten frames.
 * T is the engine's ring length (the reference fixes 256).  Time sources
 * 'captureTime' and 'watchdog' share a timebase; switching to or from
 * 'mediaTime' restarts the clock.
 */
export class FrameClock {
  constructor(fps, T = DDM.T) {
    if (!(fps > 0) || !(T >= 1)) throw new Error('FrameClock: fps must be positive and T >= 1');
    this.T = T;
    this.dtMs = 1000 / fps;
    this.frames = 0;
    this.missed = 0;
    this.duplicates = 0;
    this.resets = 0;
    this._slotN = new Float64Array(T).fill(-1);
    this._t0 = null;
    this._lastN = -1;
    this._src = null;
    this._lastTs = NaN;
    this._dts = [];
  }

  get timeSource() { return this._src; }

  get newestN() { return this._lastN; }

  reset() {
    this._t0 = null;
    this._lastN = -1;
    this._slotN.fill(-1);
    this._src = null;
    this._lastTs = NaN;
    this.resets++;
  }

  /** {slot, n} for a frame captured at tMs, or null for a duplicate / late frame. */
  place(tMs, source = 'captureTime') {
    if (this._src !== null && (source === 'mediaTime') !== (this._src === 'mediaTime')) this.reset();
    this._src = source;
    if (this._t0 === null) this._t0 = tMs;
    let n = Math.round((tMs - this._t0) / this.dtMs);
    if (n <= this._lastN) {
      this.duplicates++;
      return null;
    }
    if (this._lastN >= 0 && n - this._lastN > this.T) {   // gap longer than the ring: restart here
      this.reset();
      this._src = source;
      this._t0 = tMs;
      n = 0;
    }
    if (this._lastN >= 0) this.missed += n - this._lastN - 1;
    if (Number.isFinite(this._lastTs)) {
      this._dts.push(tMs - this._lastTs);
      if (this._dts.length > 120) this._dts.shift();
    }
    this._lastTs = tMs;
    const slot = n % this.T;
    this._slotN[slot] = n;
    this._lastN = n;
    this.frames++;
    return { slot, n };
  }

  /** Frame n was not ingested (e.g. back-pressure). */
  drop(n) {
    const slot = n % this.T;
    if (this._slotN[slot] === n) this._slotN[slot] = -1;
  }

  get medianDtMs() {
    if (!this._dts.length) return NaN;
    const s = this._dts.slice().sort((a, b) => a - b);
    return s[s.length >> 1];
  }

  get validCount() {
    let c = 0;
    for (let j = 0; j < this.T; j++) {
      const n = this._lastN - this.T + 1 + j;
      if (n >= 0 && this._slotN[n % this.T] === n) c++;
    }
    return c;
  }

  /** Time-ordered {order, mask, used, newestN}; the oldest `guard` indices masked; shuffle = null test. */
  snapshot(opts = {}) {
    const guard = pick(opts.guard, 0);
    const order = new Uint32Array(this.T);
    const mask = new Float32Array(this.T);
    let used = 0;
    for (let j = 0; j < this.T; j++) {
      const n = this._lastN - this.T + 1 + j;
      const slot = ((n % this.T) + this.T) % this.T;
      const ok = n >= 0 && j >= guard && this._slotN[slot] === n;
      order[j] = slot;
      mask[j] = ok ? 1 : 0;
      if (ok) used++;
    }
    if (opts.shuffle) {
      const s = shuffleOrder({ order, mask }, opts.rand || Math.random);
      return { order: s.order, mask: s.mask, used, newestN: this._lastN };
    }
    return { order, mask, used, newestN: this._lastN };
  }
}
