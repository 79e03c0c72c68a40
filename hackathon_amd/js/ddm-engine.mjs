// TypeScript-facing facade of the MI355X DDM engine (ES module, ES2019).
//
// Keeps the method names, argument meaning, return values and thrown errors of
// the reference's WebGPU engine class (motility-ddm src/ddm-engine.ts:62-426)
// so callers switch by changing the import.  Every call goes through the N-API
// addon (../napi/ddm_napi.node) to libddm.so: HIP kernels on one MI355X.  The
// calls that wait on the GPU or on large copies run off the event loop (the
// addon's async work) and are chained per engine, so they take effect in the
// order they were made; ingestLuma stays synchronous, as in the reference.
// Types are declared in ddm-engine.d.ts.
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
    this._tail = Promise.resolve();   // the engine's queue of asynchronous calls
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

  // run `call` (returning the addon's promise) after every call queued before it
  _queue(call) {
    const p = this._tail.then(call);
    this._tail = p.then(() => undefined, () => undefined);
    return p;
  }

  get pendingSubmits() { return 0; }
  get saturated() { return false; }
  async waitForCapacity() {}

  /** ingestLuma (src/ddm-engine.ts:308-324): one N*N Float32Array frame into slot % T, synchronously. */
  ingestLuma(luma, slot) {
    if (this._destroyed) { this.throttled++; return false; }
    if (luma.length !== this.N * this.N) throw new Error(`ingestLuma expects ${this.N * this.N} values`);
    return addon.ingestSync(this._h, luma, 1, slot % this.T);
  }

  /** Frame stack in: nframes frames (Uint8Array | Uint16Array | Float32Array) into slots firstSlot + i. */
  async ingestFrames(frames, nframes, firstSlot = 0) {
    if (this._destroyed) { this.throttled++; return false; }
    return this._queue(() => addon.ingest(this._h, frames, nframes, firstSlot));
  }

  /** computeTemporal (src/ddm-engine.ts:331-365): resolves once D(q, tau) is ready on the GPU. */
  async computeTemporal(order, mask, opts = {}) {
    if (this._temporalBusy) throw new Error('computeTemporal is already running');
    this._temporalBusy = true;
    try {
      await this._queue(() => addon.computeTemporal(this._h, order, mask));
    } finally {
      this._temporalBusy = false;
    }
  }

  /**
   * readRadial (src/ddm-engine.ts:391-395): the Radial record of src/ddm_ref.js:336-348, with the
   * phase-drift sums as in the reference (opts.drift = false skips them: they take a per-mode pass).
   */
  async readRadial(opts = {}) {
    const r = await this._queue(() => addon.readRadial(this._h, pick(opts.drift, true)));
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
  async readTemporalFull() { return this._queue(() => addon.readTemporalFull(this._h)); }

  /** readMean (src/ddm-engine.ts:410-414): mean luma of the last frame ingested (u8/255, u16/65535). */
  async readMean() { return this._queue(() => addon.readMean(this._h)); }

  /** readHistory (src/ddm-engine.ts:397-400): hist[(mode*T + slot)*2 + {0,1}]. */
  async readHistory() { return this._queue(() => addon.readHistory(this._h)); }

  async idle() { return this._queue(() => addon.idle(this._h)); }

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

/**
 * shuffleOrder (src/ddm_ref.js:510-519), the frame-order null test: the library's Fisher-Yates
 * permutation of the (slot, mask) pairs, fed with rand()'s draws in the reference's order.
 */
export function shuffleOrder(o, rand) {
  const order = Uint32Array.from(o.order);
  const mask = Float32Array.from(o.mask);
  const draws = new Float64Array(Math.max(order.length - 1, 0));
  for (let k = 0; k < draws.length; k++) draws[k] = rand();
  addon.shuffleOrder(order, mask, draws);
  return { order, mask };
}

/** linearOrder (src/ddm_ref.js:488-502): time index j in slot j, valid where `valid` is truthy. */
export function linearOrder(valid) {
  return {
    order: Uint32Array.from(valid, (_, j) => j),
    mask: Float32Array.from(valid, (v) => (v ? 1 : 0)),
  };
}

const SOURCES = ['captureTime', 'watchdog', 'mediaTime'];

/**
 * The live frame -> history-slot clock: the library's frame clock (ddm_clock_*,
 * the reference's FrameClock contract, src/camera.ts:306-425; semantics with
 * ddm_clock_create in include/ddm.h).  Frame n = round((t - t0) / dt) goes to
 * slot n mod T; snapshot() masks dropped, late, duplicated and overwritten
 * frames.  T is the engine's ring length (the reference fixes 256).
 */
export class FrameClock {
  constructor(fps, T = DDM.T) {
    this._c = addon.clockCreate(fps, T);
    this.T = T;
    this.dtMs = 1000 / fps;
  }

  get frames() { return addon.clockState(this._c).frames; }
  get missed() { return addon.clockState(this._c).missed; }
  get duplicates() { return addon.clockState(this._c).duplicates; }
  get resets() { return addon.clockState(this._c).resets; }
  get newestN() { return addon.clockState(this._c).newestN; }
  get validCount() { return addon.clockState(this._c).valid; }
  get medianDtMs() { return addon.clockState(this._c).medianDtMs; }
  get timeSource() {
    const s = addon.clockState(this._c).timeSource;
    return s >= 0 ? SOURCES[s] : null;
  }

  reset() { addon.clockReset(this._c); }

  /** {slot, n} for a frame captured at tMs, or null for a duplicate / late frame. */
  place(tMs, source = 'captureTime') {
    const src = SOURCES.indexOf(source);
    if (src < 0) throw new Error(`FrameClock.place: unknown time source ${source}`);
    const r = addon.clockPlace(this._c, tMs, src);
    return r === null ? null : { slot: r[0], n: r[1] };
  }

  /** Frame n was not ingested (e.g. back-pressure). */
  drop(n) { addon.clockDrop(this._c, n); }

  /** Time-ordered {order, mask, used, newestN}; the oldest `guard` indices masked; shuffle = null test. */
  snapshot(opts = {}) {
    const s = addon.clockSnapshot(this._c, pick(opts.guard, 0), this.T);
    const newestN = this.newestN;
    if (opts.shuffle) {
      const sh = shuffleOrder(s, opts.rand || Math.random);
      return { order: sh.order, mask: sh.mask, used: s.used, newestN };
    }
    return { order: s.order, mask: s.mask, used: s.used, newestN };
  }
}
