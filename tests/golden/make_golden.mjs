// Writes golden vectors for the parts of the oracle that nothing else pins.
//
// Run in the build container only (node >= 12, reference checked out at
// /root/reference):   node tests/golden/make_golden.mjs
//
//  * bins_*.json   makeBins() kCenter / kNodes / nModes / binOffsets, produced by
//                  importing motility-ddm/src/ddm_ref.js.  kNodes (the window-
//                  leakage quantile nodes) has no closed form and no reference
//                  test, so importing the twin is its only pin.
//  * v8_hypot.json this Node's Math.hypot (V8 builtin, not reference code) on
//                  the Pythagorean integer pairs where it rounds below the exact
//                  integer, plus bit patterns on half-integer pairs.
import { writeFileSync } from 'fs';
import { makeBins } from '/root/reference/motility-ddm/src/ddm_ref.js';

const out = (name, obj) => writeFileSync(new URL(name, import.meta.url), JSON.stringify(obj) + '\n');


const geoms = [
  [256, 32, 4, 100],   // reference defaults (DDM)
  [128, 64, 0, 64],    // 64 frames of 128^2, 64 unit-width bins
  [256, 128, 0, 128],  // unit width at N=256 (12 V8-hypot edge flips)
  [512, 256, 0, 256],  // 512^2, 256 unit-width bins
  [64, 16, 2, 30],     // non-unit width, kmin > 0
  [16, 8, 0, 8],
];
for (const [N, NB, kmin, kmax] of geoms) {
  const b = makeBins(N, NB, kmin, kmax);
  out(`bins_${N}_${NB}_${kmin}_${kmax}.json`, {
    N, NB, kmin, kmax,
    nModes: Array.from(b.nModes),
    binOffsets: Array.from(b.binOffsets),
    kCenter: Array.from(b.kCenter),
    kNodes: Array.from(b.kNodes),
  });
}

const bits = (x) => {
  const f = new Float64Array([x]);
  const u = new Uint32Array(f.buffer);
  return u[1].toString(16).padStart(8, '0') + u[0].toString(16).padStart(8, '0');
};
const low = [];
let pyth = 0;
for (let a = 0; a <= 1024; a++) {
  for (let b = 0; b <= 1024; b++) {
    const s = a * a + b * b;
    const r = Math.round(Math.sqrt(s));
    if (r * r !== s) continue;
    pyth++;
    if (Math.hypot(a, b) < r) low.push([a, b]);
  }
}
const half = [];
for (let a = -40; a <= 40; a++) {
  for (let b = -40; b <= 40; b += 3) {
    const x = a * 0.5 + 7, y = b * 0.5 - 3;
    half.push([x, y, bits(Math.hypot(x, y))]);
  }
}
out('v8_hypot.json', { node: process.versions.node, v8: process.versions.v8, pythagorean: pyth, below: low, half });
console.log('golden vectors written');
