// cpu_path.mjs -- the CPU baseline: a JavaScript restatement of the reference's
// CPU analysis path, History.ingest -> analyzeRadial (src/ddm_ref.js:177-205,
// 229-288, 353-437; src/fft.js:59-93, 149-183), run under Node on one thread
// or sharded over worker_threads (frames for the ingest, Fourier modes for the
// temporal stage).
//
// TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg and tests/ run it; the
// product library never loads it.  It follows the same algorithm as the
// reference so that its time is the reference's CPU cost: radix-2 Stockham
// FFTs in float64 (here the decimation-in-frequency form), spectrum of the
// mean-removed, Blackman-Harris-windowed frame (complex row FFTs, then column
// FFTs of the half plane, x 1/N), per-mode masked Wiener-Khinchin with the
// mask and |x|^2 packed as z = m + i p into one transform, and per-bin
// accumulation of D, P and the eight phase-drift sums.  tests/test_cpu_path.py
// holds its output to the C oracle (oracle/ddm_oracle.c) on the same input.
//
// Usage: node oracle/cpu_path.mjs <dir>
//   <dir>/spec.json   {"N", "T", "L", "NB", "dtype": "u8"|"u16", "threads"}
//   <dir>/frames.bin  T frames of N*N pixels (luma = pixel / 255 or / 65535), frame t to slot t
//   <dir>/binOffsets.bin (uint32, NB+1), <dir>/binModes.bin (int32, 4 per entry: mode, kx, ky, 0)
// Writes <dir>/D.bin (float64 NB*L), P.bin (NB), drift.bin (NB*L*8) and prints
// one JSON line of wall times.  With threads > 1 the worker pool is started
// before the timed region, as a long-running service would keep it.
import fs from 'fs';
import os from 'os';
import path from 'path';
import { Worker, isMainThread, parentPort, workerData } from 'worker_threads';

const now = () => Number(process.hrtime.bigint()) / 1e9;

// ------------------------------------------------------------------ FFT
const twCache = new Map();
function twiddles(n) {   // e^{-2 pi i k / n}, k < n/2, interleaved
  let t = twCache.get(n);
  if (!t) {
    t = new Float64Array(n);
    for (let k = 0; k < n / 2; k++) {
      const a = (-2 * Math.PI * k) / n;
      t[2 * k] = Math.cos(a);
      t[2 * k + 1] = Math.sin(a);
    }
    twCache.set(n, t);
  }
  return t;
}

// Unnormalised complex FFT of x (interleaved, n a power of two) in place, w a
// work array of the same length.  Stockham decimation in frequency: the stage
// of length len splits each stride-s sub-sequence into its sum and twiddled
// difference halves, output in natural order.  inverse conjugates the twiddles.
function fft(x, w, n, inverse) {
  const tw = twiddles(n);
  const sg = inverse ? -1 : 1;
  let src = x, dst = w;
  for (let len = n, s = 1; len > 1; len >>= 1, s <<= 1) {
    const h = len >> 1;
    const s2 = 2 * s;
    for (let p = 0; p < h; p++) {
      const wr = tw[2 * p * s], wi = sg * tw[2 * p * s + 1];
      const a0 = s2 * p, b0 = s2 * (p + h), o0 = 2 * s2 * p, o1 = o0 + s2;
      for (let q = 0; q < s2; q += 2) {
        const ar = src[a0 + q], ai = src[a0 + q + 1], br = src[b0 + q], bi = src[b0 + q + 1];
        dst[o0 + q] = ar + br;
        dst[o0 + q + 1] = ai + bi;
        const dr = ar - br, di = ai - bi;
        dst[o1 + q] = dr * wr - di * wi;
        dst[o1 + q + 1] = dr * wi + di * wr;
      }
    }
    const t = src;
    src = dst;
    dst = t;
  }
  if (src !== x) x.set(src.subarray(0, 2 * n));
}

function blackmanHarris(n) {   // 4-term, periodic
  const w = new Float64Array(n);
  for (let i = 0; i < n; i++) {
    const a = (2 * Math.PI * i) / n;
    w[i] = 0.35875 - 0.48829 * Math.cos(a) + 0.14128 * Math.cos(2 * a) - 0.01168 * Math.cos(3 * a);
  }
  return w;
}

// ------------------------------------------------------------------ stages
function makeState(sp, sh) {
  const N = sp.N, T = sp.T, L = sp.L, NH = (N >> 1) + 1;
  let n = 1;
  while (n < 2 * T) n <<= 1;
  const frames = sp.dtype === 'u16' ? new Uint16Array(sh.frames) : new Uint8Array(sh.frames);
  const binOffsets = new Uint32Array(sh.binOffsets), binModes = new Int32Array(sh.binModes);
  const binOf = new Int32Array(binModes.length / 4);
  for (let b = 0; b < sp.NB; b++) for (let i = binOffsets[b]; i < binOffsets[b + 1]; i++) binOf[i] = b;
  return {
    N, T, L, NH, n, NB: sp.NB, frames, scale: sp.dtype === 'u16' ? 1 / 65535 : 1 / 255, binModes, binOf,
    hist: new Float64Array(sh.hist),
    win: blackmanHarris(N), row: new Float64Array(2 * N), work: new Float64Array(2 * Math.max(N, n)),
    rows: new Float64Array(2 * N * NH), z: new Float64Array(2 * n), y: new Float64Array(2 * n),
    S: new Float64Array(2 * n), D: new Float64Array(L), C: new Float64Array(2 * L),
  };
}

// spectrum of frame f into history slot f (History.ingest)
function ingestFrame(st, f) {
  const { N, T, NH, frames, win, row, work, rows, hist } = st;
  const off = f * N * N, sc = st.scale;   // luma = pixel / full scale, as lumaOf
  let mean = 0;
  for (let i = 0; i < N * N; i++) mean += frames[off + i] * sc;
  mean /= N * N;
  for (let yy = 0; yy < N; yy++) {
    const wy = win[yy], ro = off + yy * N;
    for (let x = 0; x < N; x++) {
      row[2 * x] = (frames[ro + x] * sc - mean) * win[x] * wy;
      row[2 * x + 1] = 0;
    }
    fft(row, work, N, false);
    rows.set(row.subarray(0, 2 * NH), 2 * yy * NH);
  }
  const s = 1 / N;
  for (let kx = 0; kx < NH; kx++) {
    for (let yy = 0; yy < N; yy++) {
      row[2 * yy] = rows[2 * (yy * NH + kx)];
      row[2 * yy + 1] = rows[2 * (yy * NH + kx) + 1];
    }
    fft(row, work, N, false);
    for (let ky = 0; ky < N; ky++) {
      const o = ((ky * NH + kx) * T + f) * 2;
      hist[o] = row[2 * ky] * s;
      hist[o + 1] = row[2 * ky + 1] * s;
    }
  }
}

// masked WK of one mode (identity order, every frame valid); fills st.D, st.C, returns P
function temporalMode(st, mode) {
  const { T, L, n, z, y, S, work, hist, D, C } = st;
  z.fill(0);
  y.fill(0);
  const base = mode * T * 2;
  for (let j = 0; j < T; j++) {
    const xr = hist[base + 2 * j], xi = hist[base + 2 * j + 1];
    z[2 * j] = 1;
    z[2 * j + 1] = xr * xr + xi * xi;
    y[2 * j] = xr;
    y[2 * j + 1] = xi;
  }
  fft(z, work, n, false);
  // Z = FFT(m + i p): M = (Z + conj Z_-k)/2, P = (Z - conj Z_-k)/(2i); S = conj(M) P + i |M|^2
  for (let k = 0; k < n; k++) {
    const kk = k === 0 ? 0 : n - k;
    const zr = z[2 * k], zi = z[2 * k + 1], cr = z[2 * kk], ci = -z[2 * kk + 1];
    const mr = 0.5 * (zr + cr), mi = 0.5 * (zi + ci);
    const pr = 0.5 * (zi - ci), pi = -0.5 * (zr - cr);
    S[2 * k] = mr * pr + mi * pi;
    S[2 * k + 1] = mr * pi - mi * pr + (mr * mr + mi * mi);
  }
  fft(S, work, n, true);
  fft(y, work, n, false);
  for (let k = 0; k < n; k++) {
    const yr = y[2 * k], yi = y[2 * k + 1];
    y[2 * k] = yr * yr + yi * yi;
    y[2 * k + 1] = 0;
  }
  fft(y, work, n, true);
  const inv = 1 / n;
  for (let tau = 0; tau < L; tau++) {
    const A = S[2 * tau] * inv, B = S[2 * (tau === 0 ? 0 : n - tau)] * inv;
    const np = Math.round(S[2 * tau + 1] * inv);
    if (np > 0) {
      const gr = y[2 * tau] * inv, gi = y[2 * tau + 1] * inv;
      D[tau] = (A + B - 2 * gr) / np;
      C[2 * tau] = gr / np;
      C[2 * tau + 1] = gi / np;
    } else {
      D[tau] = 0;
      C[2 * tau] = 0;
      C[2 * tau + 1] = 0;
    }
  }
  const m0 = S[1] * inv;
  return m0 > 0.5 ? (S[0] * inv) / m0 : 0;
}

// temporal stage over bin entries [i0, i1): per-bin sums into out {D, P, drift}
function temporalRange(st, i0, i1, out) {
  const { L, binModes, binOf, D, C } = st;
  for (let i = i0; i < i1; i++) {
    const b = binOf[i], kx = binModes[4 * i + 1], ky = binModes[4 * i + 2];
    out.P[b] += temporalMode(st, binModes[4 * i]);
    const od = b * L;
    for (let tau = 0; tau < L; tau++) out.D[od + tau] += D[tau];
    for (let tau = 0; tau + 1 < L; tau++) {
      const ar = C[2 * tau], ai = C[2 * tau + 1], br = C[2 * tau + 2], bi = C[2 * tau + 3];
      const zr = br * ar + bi * ai, zi = bi * ar - br * ai;   // C(tau+1) conj C(tau)
      const w = Math.hypot(zr, zi), ph = Math.atan2(zi, zr);
      const o = (od + tau) * 8;
      out.drift[o] += w;
      out.drift[o + 1] += w * kx * kx;
      out.drift[o + 2] += w * kx * ky;
      out.drift[o + 3] += w * ky * ky;
      out.drift[o + 4] += w * kx * ph;
      out.drift[o + 5] += w * ky * ph;
      out.drift[o + 6] += w * ph * ph;
      out.drift[o + 7] += w > 0 ? 1 : 0;
    }
  }
}

function partial(sh, k, NB, L) {
  const per = NB + NB * L * 9;
  const base = k * per * 8;
  return {
    P: new Float64Array(sh.partials, base, NB),
    D: new Float64Array(sh.partials, base + NB * 8, NB * L),
    drift: new Float64Array(sh.partials, base + (NB + NB * L) * 8, NB * L * 8),
  };
}

// ------------------------------------------------------------------ worker
if (!isMainThread) {
  const { sp, sh, k } = workerData;
  const st = makeState(sp, sh);
  const out = partial(sh, k, sp.NB, sp.L);
  parentPort.on('message', (m) => {
    if (m.op === 'ingest') for (let f = m.a; f < m.b; f++) ingestFrame(st, f);
    else temporalRange(st, m.a, m.b, out);
    parentPort.postMessage(0);
  });
  parentPort.postMessage(0);   // ready
}

// ------------------------------------------------------------------ main
async function main() {
  const dir = process.argv[2];
  const sp = JSON.parse(fs.readFileSync(path.join(dir, 'spec.json'), 'utf8'));
  const { N, T, L, NB } = sp;
  const NH = (N >> 1) + 1;
  const threads = Math.max(1, sp.threads | 0);
  const share = (file) => {
    const b = fs.readFileSync(path.join(dir, file));
    const s = new SharedArrayBuffer(b.length);
    new Uint8Array(s).set(b);
    return s;
  };
  const sh = {
    frames: share('frames.bin'), binOffsets: share('binOffsets.bin'), binModes: share('binModes.bin'),
    hist: new SharedArrayBuffer(N * NH * T * 16),
    partials: new SharedArrayBuffer(threads * (NB + NB * L * 9) * 8),
  };
  const nent = new Int32Array(sh.binModes).length / 4;
  let t0, t1, t2;
  if (threads === 1) {
    const st = makeState(sp, sh);
    t0 = now();
    for (let f = 0; f < T; f++) ingestFrame(st, f);
    t1 = now();
    temporalRange(st, 0, nent, partial(sh, 0, NB, L));
    t2 = now();
  } else {
    const me = new URL(import.meta.url).pathname;
    const pool = [];
    await Promise.all(Array.from({ length: threads }, (_, k) => new Promise((res, rej) => {
      const w = new Worker(me, { workerData: { sp, sh, k } });
      w.once('message', res);
      w.on('error', rej);
      pool.push(w);
    })));
    const run = (op, count) => Promise.all(pool.map((w, k) => new Promise((res) => {
      w.once('message', res);
      w.postMessage({ op, a: Math.floor((count * k) / threads), b: Math.floor((count * (k + 1)) / threads) });
    })));
    t0 = now();
    await run('ingest', T);
    t1 = now();
    await run('temporal', nent);
    t2 = now();
    pool.forEach((w) => w.terminate());
  }
  // reduce the partial sums and divide by the bin's mode count (finalizeRadial)
  const bo = new Uint32Array(sh.binOffsets);
  const R = partial(sh, 0, NB, L);
  for (let k = 1; k < threads; k++) {
    const p = partial(sh, k, NB, L);
    for (let i = 0; i < NB; i++) R.P[i] += p.P[i];
    for (let i = 0; i < NB * L; i++) R.D[i] += p.D[i];
    for (let i = 0; i < NB * L * 8; i++) R.drift[i] += p.drift[i];
  }
  for (let b = 0; b < NB; b++) {
    const c = bo[b + 1] - bo[b];
    if (!c) continue;
    R.P[b] /= c;
    for (let tau = 0; tau < L; tau++) R.D[b * L + tau] /= c;
  }
  const t3 = now();
  const w = (name, a) => fs.writeFileSync(path.join(dir, name), Buffer.from(a.buffer, a.byteOffset, a.byteLength));
  w('D.bin', R.D);
  w('P.bin', R.P);
  w('drift.bin', R.drift);
  console.log(JSON.stringify({
    threads, ingest_s: t1 - t0, temporal_s: t2 - t1 + (t3 - t2), total_s: t3 - t0, modes: nent,
    frame_pixels_per_s: (T * N * N) / (t3 - t0), node: process.versions.node, cpu: os.cpus()[0].model,
    cpus: os.cpus().length,
  }));
}

if (isMainThread) main().catch((e) => { console.error(e); process.exit(1); });
