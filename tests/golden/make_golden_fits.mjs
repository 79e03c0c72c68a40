// Reference outputs for the parts of the oracle that only the reference's own
// code can pin: the fits on noisy records, the exact isf2dFast table values and
// the movie simulator.  Run in the build container only (node >= 12, the
// reference checked out at /root/reference), driven by make_golden_fits.py,
// which writes the input records (the oracle pipeline's Radial records of six
// validation-matrix scenarios) into <dir>:
//     node tests/golden/make_golden_fits.mjs <dir>
//
// src/motility_fit.js and test/sim.mjs use `??` and `?.`, which Node 12 does
// not parse: their text is read, those operators are rewritten to plain
// conditionals in memory, and the module body is evaluated from that string.
// Nothing of the reference is written anywhere; only the numbers below are.
import fs from 'fs';
import path from 'path';

const REF = '/root/reference/motility-ddm';

function load(file, names) {
  let src = fs.readFileSync(path.join(REF, file), 'utf8');
  const lit = `('[^']*'|-?[0-9.]+(?:e-?[0-9]+)?|true|false|[A-Za-z_$][\\w$.]*)`;
  src = src.replace(/([A-Za-z_$][\w$.]*)\?\.([A-Za-z_$][\w$]*)/g, '($1 == null ? undefined : $1.$2)');
  src = src.replace(new RegExp(`(\\([^()]*\\)|[A-Za-z_$][\\w$.]*)\\s*\\?\\?\\s*${lit}`, 'g'), '__nc($1, $2)');
  src = src.replace(/^export (function|const|class) /gm, '$1 ');
  if (/\?\?|\?\.[A-Za-z_$]/.test(src)) throw new Error(`unlowered ?? or ?. left in ${file}`);
  const body = 'const __nc = (a, b) => (a === null || a === undefined ? b : a);\n' + src +
    `\nreturn { ${names.join(', ')} };`;
  return new Function(body)();   // eslint-disable-line no-new-func
}

const fit = load('src/motility_fit.js',
  ['prepare', 'fitBrownian', 'fitSwimmer', 'estimateDrift', 'analyzeMotility', 'isf2dFast', 'isf2d']);
const sim = load('test/sim.mjs', ['simulate']);

const dir = process.argv[2];
const spec = JSON.parse(fs.readFileSync(path.join(dir, 'records.json'), 'utf8'));
const out = { node: process.versions.node, records: [], isf2dFast: [], sims: [] };

for (const r of spec.records) {
  const radial = {
    N: r.N, NB: r.NB, L: r.L, kCenter: Float64Array.from(r.kCenter), kNodes: Float64Array.from(r.kNodes),
    nModes: Uint32Array.from(r.nModes), D: Float64Array.from(r.D), drift: Float64Array.from(r.drift),
  };
  const opt = { dxUm: r.dxUm, fps: r.fps, geometry: r.geometry, exposureLocked: true };
  const pr = fit.prepare(radial, opt);
  const brown = fit.fitBrownian(pr, 0);
  const swim = fit.fitSwimmer(pr, r.geometry, 0, brown.Dd);
  const drift = fit.estimateDrift(radial, pr, opt, swim.alpha > 0.01 ? swim : null);
  const m = fit.analyzeMotility(radial, opt);
  const pick = (o, ks) => Object.fromEntries(ks.map((k) => [k, o[k]]));
  out.records.push({
    name: r.name, geometry: r.geometry, bins: pr.bins.map((b) => b.b),
    fitBrownian: pick(brown, ['Dd', 'rss']),
    fitSwimmer: pick(swim, ['vbar', 'Z', 'Dd', 'alpha', 'rss']),
    estimateDrift: pick(drift, ['Vx', 'Vy', 'mag', 'sigma', 'significant', 'pairs', 'samples']),
    analyzeMotility: Object.assign(
      pick(m, ['verdict', 'physicsVerdict', 'TSI', 'alpha', 'vbar', 'Z', 'Ddiff', 'DdiffBrownian', 'dBIC', 'mu',
        'nData', 'nEff']),
      { drift: m.drift ? pick(m.drift, ['Vx', 'Vy', 'mag', 'significant']) : null,
        gates: m.gates.map((g) => ({ id: g.id, pass: g.pass, value: g.value })),
        swimmer: m.fit ? pick(m.fit.swimmer, ['vbar', 'Z', 'Dd', 'Vmag', 'alpha', 'rss']) : null,
        brownian: m.fit ? pick(m.fit.brownian, ['Dd', 'Vmag', 'rss']) : null }),
  });
}

// the tabulated in-plane kernel against which the fits run, and the exact form it tabulates
for (const Z of [0.5, 0.8, 1.3, 2.1, 3.4, 5.5, 8.9, 14.4, 23.3, 37.7, 61.0, 99.0]) {
  for (const x of [0, 1e-5, 3e-4, 0.01, 0.07, 0.2, 0.45, 0.9, 1.6, 2.7, 4.2, 6.6, 10.5, 17, 29, 55, 120]) {
    out.isf2dFast.push([x, Z, fit.isf2dFast(x, Z), fit.isf2d(x, Z)]);
  }
}

// small movies from the simulator: every population, drift, dropped frames
for (const p of spec.sims) {
  const s = sim.simulate(p);
  const bytes = Buffer.concat(s.frames.map((f) => Buffer.from(f.buffer, f.byteOffset, f.byteLength)));
  out.sims.push({ params: p, frames: bytes.toString('base64'), valid: s.valid.map((v) => (v ? 1 : 0)),
    meanSpeed: s.truth.meanSpeed });
}

fs.writeFileSync(path.join(path.dirname(new URL(import.meta.url).pathname), 'fits_reference.json'),
  JSON.stringify(out) + '\n');
console.log(`reference outputs written: ${out.records.length} records, ${out.isf2dFast.length} isf2dFast samples, ` +
  `${out.sims.length} movies`);
