// The event loop keeps running while the engine works: a timer counts its
// ticks during ingestFrames of a 1024-frame 512x512 u16 stack and during
// twenty computeTemporal calls through to D(q, tau) (all lags, 256 bins).  A second
// computeTemporal while one is running is refused, as in the reference
// (src/ddm-engine.ts:331-352).
//   node tests/js/async_run.mjs
import { DdmEngine, linearOrder } from '../../hackathon_amd/js/ddm-engine.mjs';

async function run() {
  const N = 512, T = 1024;
  const frames = new Uint16Array(T * N * N);
  for (let i = 0; i < frames.length; i++) frames[i] = (i * 2654435761) >>> 20;
  const e = await DdmEngine.create({ N, T, L: T, NB: 256, kmin: 0, kmax: 256 });
  let ticks = 0;
  const timer = setInterval(() => { ticks++; }, 1);
  const t0 = Date.now();
  await e.ingestFrames(frames, T, 0);
  const ingestTicks = ticks, ingestMs = Date.now() - t0;
  const { order, mask } = linearOrder(new Array(T).fill(true));
  const t1 = Date.now();
  const first = e.computeTemporal(order, mask);
  let reentry = '';
  try { await e.computeTemporal(order, mask); } catch (x) { reentry = x.message; }
  await first;
  for (let k = 0; k < 19; k++) await e.computeTemporal(order, mask);   // twenty analyses in all
  const temporalTicks = ticks - ingestTicks, temporalMs = Date.now() - t1;
  const r = await e.readRadial({ drift: false });
  clearInterval(timer);
  e.destroy();
  let finite = true;
  for (const v of r.D) finite = finite && Number.isFinite(v);
  return { ingestTicks, ingestMs, temporalTicks, temporalMs, reentry, finite, d0: r.D[10 * T], d1: r.D[10 * T + 1] };
}

run().then((r) => console.log('RESULT ' + JSON.stringify(r)))
  .catch((e) => { console.log('RESULT ' + JSON.stringify({ error: String(e && e.stack || e) })); process.exit(1); });
