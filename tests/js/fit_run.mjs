// Frame stack in -> D(q, tau) -> per-q fits out, through the TypeScript-facing
// API only (DdmEngine + analyzeMotilityLM).  Writes the fitted parameters as
// raw f64 next to the frames for the Python checker.
//   node tests/js/fit_run.mjs <dir> <N> <T> <L> <NB> <kmin> <kmax> <fps>
import fs from 'fs';
import path from 'path';
import { DdmEngine, linearOrder } from '../../hackathon_amd/js/ddm-engine.mjs';
import { analyzeMotilityLM, lagGrid, prepare, MODEL } from '../../hackathon_amd/js/motility-fit.mjs';

async function run(dir, N, T, L, NB, kmin, kmax, fps) {
  const raw = fs.readFileSync(path.join(dir, 'frames.bin'));
  const frames = new Uint8Array(raw.buffer, raw.byteOffset, raw.byteLength);
  const e = await DdmEngine.create({ N, T, L, NB, kmin, kmax });
  await e.ingestFrames(frames, T, 0);
  const { order, mask } = linearOrder(new Array(T).fill(true));
  await e.computeTemporal(order, mask);
  const radial = await e.readRadial();
  e.destroy();
  const res = analyzeMotilityLM(radial, { dxUm: 1, fps, model: MODEL.BROWNIAN });
  const th = new Float64Array(res.fits.length * 3);
  res.fits.forEach((f, i) => { th[3 * i] = f.A; th[3 * i + 1] = f.B; th[3 * i + 2] = f.D; });
  fs.writeFileSync(path.join(dir, 'fit.f64'), Buffer.from(th.buffer));
  // the same fits with a fixed drift speed in every bin's model
  const resV = analyzeMotilityLM(radial, { dxUm: 1, fps, model: MODEL.BROWNIAN, vmag: 2.5 });
  fs.writeFileSync(path.join(dir, 'fitV.f64'), Buffer.from(Float64Array.from(resV.fits.map((f) => f.D)).buffer));
  fs.writeFileSync(path.join(dir, 'radD.f64'), Buffer.from(radial.D.buffer));
  return { bins: res.fits.map((f) => f.b), chi2: res.fits.map((f) => f.chi2), chi2V: resV.fits.map((f) => f.chi2),
    lags: lagGrid(L),
    nPrepared: prepare(radial, 1, fps).bins.length, kernelMs: res.kernelMs };
}

const a = process.argv.slice(2);
run(a[0], +a[1], +a[2], +a[3], +a[4], +a[5], +a[6], +a[7])
  .then((r) => console.log('RESULT ' + JSON.stringify(r)))
  .catch((e) => { console.log('RESULT ' + JSON.stringify({ error: String(e && e.stack || e) })); process.exit(1); });
