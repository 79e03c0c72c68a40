// Drives the TypeScript-facing facade the way the reference's harness drives
// DdmEngine (test/gpu/harness.ts:75-113): create, ingest a frame stack,
// computeTemporal, readRadial / readHistory.  Frames come from a raw file;
// results are written as raw f64/f32 files next to it for the Python checker.
//   node tests/js/engine_run.mjs <dir> <N> <T> <L> <NB> <kmin> <kmax> <dtype:u8|u16|f32>
// With "errors" as the only argument it exercises the error paths only.
import fs from 'fs';
import path from 'path';
import { DdmEngine, linearOrder } from '../../hackathon_amd/js/ddm-engine.mjs';

async function errors() {
  const out = {};
  try { await DdmEngine.create({ N: 100, T: 8, L: 4, NB: 4, kmin: 0, kmax: 8 }); } catch (e) { out.create = e.message; }
  return out;
}

async function run(dir, N, T, L, NB, kmin, kmax, dtype) {
  const raw = fs.readFileSync(path.join(dir, 'frames.bin'));
  const Ctor = dtype === 'u8' ? Uint8Array : dtype === 'u16' ? Uint16Array : Float32Array;
  const frames = new Ctor(raw.buffer, raw.byteOffset, raw.byteLength / Ctor.BYTES_PER_ELEMENT);
  const e = await DdmEngine.create({ N, T, L, NB, kmin, kmax });
  const ok = await e.ingestFrames(frames, T, 0);
  const valid = new Array(T).fill(true);
  const { order, mask } = linearOrder(valid);
  await e.computeTemporal(order, mask);
  const r = await e.readRadial();
  const h = await e.readHistory();
  const mean = await e.readMean();
  let err = '';
  try { await e.computeTemporal(new Uint32Array(T - 1), new Float32Array(T - 1)); } catch (x) { err = x.message; }
  e.destroy();
  fs.writeFileSync(path.join(dir, 'D.f64'), Buffer.from(r.D.buffer));
  fs.writeFileSync(path.join(dir, 'P.f64'), Buffer.from(r.P.buffer));
  fs.writeFileSync(path.join(dir, 'hist.f32'), Buffer.from(h.buffer));
  return { ok, lenErr: err, mean, kCenter: Array.from(r.kCenter), nModes: Array.from(r.nModes), L: r.L, NB: r.NB };
}

const a = process.argv.slice(2);
(a[0] === 'errors' ? errors() : run(a[0], +a[1], +a[2], +a[3], +a[4], +a[5], +a[6], a[7]))
  .then((r) => console.log('RESULT ' + JSON.stringify(r)))
  .catch((e) => { console.log('RESULT ' + JSON.stringify({ error: String(e && e.stack || e) })); process.exit(1); });
