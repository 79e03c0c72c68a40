"""The TypeScript-facing facade (Node + N-API addon) over libddm.so."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "tests", "js", "engine_run.mjs")
ADDON = os.path.join(ROOT, "hackathon_amd", "napi", "ddm_napi.node")


def _node(*args, timeout=120):
    r = subprocess.run(["node", RUN, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, r.stdout + r.stderr
    return json.loads(line[-1][7:])


def test_facade_error_path_without_gpu_work():
    if not os.path.exists(ADDON):
        pytest.fail("N-API addon not built (python -m hackathon_amd.napi.build_addon)")
    out = _node("errors")
    assert "power of two" in out["create"]


@pytest.mark.gpu
def test_facade_parity(gpu, tmp_path):
    import oracle as O
    from hackathon_amd import synth
    fr = synth.config1()
    T = L = 64
    (tmp_path / "frames.bin").write_bytes(fr.tobytes())
    out = _node(tmp_path, 128, T, L, 64, 0, 64, "u8")
    assert out.get("ok") is True, out
    assert "length 64" in out["lenErr"]
    D = np.fromfile(tmp_path / "D.f64").reshape(64, L)
    P = np.fromfile(tmp_path / "P.f64")
    h = np.fromfile(tmp_path / "hist.f32", dtype=np.float32).view(np.complex64).reshape(-1, T)
    ref_h = O.history(fr, T)
    assert np.max(np.abs(h - ref_h)) / np.sqrt(np.mean(np.abs(ref_h) ** 2)) < 1e-5
    bins = O.make_bins(128, 64, 0, 64, with_nodes=False)
    refD, refP = O.analyze_radial(ref_h, np.arange(T), np.ones(T, np.float32), bins, L)
    ok = bins["nModes"] > 0
    scale = np.max(np.abs(refD[ok]), axis=1, keepdims=True)
    assert np.max(np.abs(D[ok] - refD[ok]) / scale) < 1e-5
    assert np.max(np.abs(P[ok] - refP[ok]) / refP[ok]) < 1e-5
    assert out["nModes"] == [int(x) for x in bins["nModes"]]
    assert out["mean"] == pytest.approx(float(np.mean(fr[T - 1].astype(np.float64))) / 255.0, rel=1e-15)


def test_js_lag_grid_matches_python():
    """lagGrid of the TS-facing module against the Python host restatement (no GPU work)."""
    code = ("import('" + os.path.join(ROOT, "hackathon_amd", "js", "motility-fit.mjs") +
            "').then(m => console.log('RESULT ' + JSON.stringify([2, 11, 64, 128, 1024, 4096].map(L => m.lagGrid(L)))))")
    if not os.path.exists(ADDON):
        pytest.fail("N-API addon not built (python -m hackathon_amd.napi.build_addon)")
    r = subprocess.run(["node", "--input-type=module", "-e", code], capture_output=True, text=True, timeout=60)
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, r.stdout + r.stderr
    from hackathon_amd.fit import lag_grid
    for L, lags in zip([2, 11, 64, 128, 1024, 4096], json.loads(line[0][7:])):
        assert lags == list(lag_grid(L)), L


@pytest.mark.gpu
def test_facade_frames_to_fits(gpu, tmp_path):
    """Frame stack in, fitted per-q parameters out, through the TS-facing API only; identical to the
    Python path on the same radial."""
    from hackathon_amd import synth
    from hackathon_amd.engine import Radial
    from hackathon_amd import fit as F
    fr = synth.config1()
    T = L = 64
    (tmp_path / "frames.bin").write_bytes(fr.tobytes())
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "fit_run.mjs"), str(tmp_path), "128", str(T),
                        str(L), "64", "0", "64", "30"], capture_output=True, text=True, timeout=180)
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, r.stdout + r.stderr
    out = json.loads(line[-1][7:])
    assert "error" not in out, out
    th = np.fromfile(tmp_path / "fit.f64").reshape(-1, 3)
    D = np.fromfile(tmp_path / "radD.f64")
    from hackathon_amd.engine import DdmEngine
    with DdmEngine.create(N=128, T=T, L=L, NB=64, kmin=0, kmax=64) as e:   # the same geometry's bin tables
        bins = e.bins
    rad = Radial(128, 64, L, bins.kCenter, bins.kNodes, bins.nModes, D, np.zeros(64), np.zeros(1))
    pr, res = F.analyzeMotilityLM(rad, 1.0, 30.0, model=F.BROWNIAN)
    assert [b.b for b in pr.bins] == out["bins"] and len(out["bins"]) == out["nPrepared"] > 0
    assert np.allclose(res.chi2, out["chi2"], rtol=1e-12, atol=0)
    assert np.allclose(res.theta[:, 0], th[:, 0], rtol=1e-12) and np.allclose(res.D, th[:, 2], rtol=1e-12)
    # a fixed drift speed (opt.vmag) reaches the library as the Python path passes it
    _, resV = F.analyzeMotilityLM(rad, 1.0, 30.0, model=F.BROWNIAN, vmag=2.5)
    assert np.allclose(resV.chi2, out["chi2V"], rtol=1e-12, atol=0)
    assert np.allclose(resV.D, np.fromfile(tmp_path / "fitV.f64"), rtol=1e-12)
    assert not np.allclose(resV.chi2, res.chi2, rtol=1e-6)


def test_js_shuffle_order_matches_oracle():
    """shuffleOrder + mulberry32 of the TS-facing facade against the oracle's restatement (no GPU work)."""
    import oracle as O
    if not os.path.exists(ADDON):
        pytest.fail("N-API addon not built (python -m hackathon_amd.napi.build_addon)")
    code = ("import('" + os.path.join(ROOT, "hackathon_amd", "js", "ddm-engine.mjs") + "').then(m => {"
            " const T = 50; const o = m.linearOrder(Array.from({length: T}, (_, i) => i % 7 !== 3));"
            " const s = m.shuffleOrder(o, m.mulberry32(4242));"
            " console.log('RESULT ' + JSON.stringify({order: Array.from(s.order), mask: Array.from(s.mask)})); })")
    r = subprocess.run(["node", "--input-type=module", "-e", code], capture_output=True, text=True, timeout=60)
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, r.stdout + r.stderr
    out = json.loads(line[0][7:])
    T = 50
    order, mask = O.shuffle_order(np.arange(T), np.array([i % 7 != 3 for i in range(T)], np.float32),
                                  O.mulberry32(4242))
    assert out["order"] == order.tolist() and out["mask"] == mask.tolist()


def test_js_estimate_drift_matches_python(tmp_path):
    """estimateDrift of the TS-facing module against the Python host version on a synthetic radial
    record (drift V plus a coherent swimmer term), open and ISF-gated (no GPU work)."""
    import math
    from tests.test_drift_estimate import DX, FPS, q_of, synth_radial
    from hackathon_amd import drift as DR
    from hackathon_amd.fit import prepare
    if not os.path.exists(ADDON):
        pytest.fail("N-API addon not built (python -m hackathon_amd.napi.build_addon)")
    rng = np.random.default_rng(3)

    def gen(kx, ky, t):
        qx, qy = q_of(kx), q_of(ky)
        g = np.exp(-0.05 * np.hypot(qx, qy) * 40 * t / FPS)
        ph = -(qx * 3.0 - qy * 4.0) * t / FPS
        noise = 0.02 * (rng.standard_normal(np.broadcast(ph, g).shape))
        return np.exp(1j * ph) + g * np.exp(1j * 6 * ph) + noise

    r = synth_radial(gen)
    rec = dict(N=r.N, NB=r.NB, L=r.L, kCenter=r.kCenter.tolist(), kNodes=r.kNodes.tolist(),
               nModes=[int(x) for x in r.nModes], D=r.D.tolist(), P=r.P.tolist(), drift=r.drift.tolist())
    (tmp_path / "radial.json").write_text(json.dumps(rec))
    mod = os.path.join(ROOT, "hackathon_amd", "js", "motility-fit.mjs")
    code = ("Promise.all([import('" + mod + "'), import('fs')]).then(([m, fs]) => {"
            " const r = JSON.parse(fs.readFileSync('" + str(tmp_path / "radial.json") + "', 'utf8'));"
            " for (const k of ['kCenter', 'kNodes', 'D', 'P', 'drift']) r[k] = Float64Array.from(r[k]);"
            " r.nModes = Uint32Array.from(r.nModes);"
            f" const pr = m.prepare(r, {DX}, {FPS});"
            f" const out = [m.estimateDrift(r, pr, {{dxUm: {DX}}}),"
            f"  m.estimateDrift(r, pr, {{dxUm: {DX}}}, {{vbar: 40, Z: 4}}),"
            f"  m.estimateDrift(r, pr, {{dxUm: {DX}, geometry: '2d'}}, {{vbar: 40, Z: 4}})];"
            " console.log('RESULT ' + JSON.stringify(out)); })")
    res = subprocess.run(["node", "--input-type=module", "-e", code], capture_output=True, text=True, timeout=60)
    line = [x for x in res.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, res.stdout + res.stderr
    out = json.loads(line[0][7:])
    pr = prepare(r, DX, FPS)
    want = [DR.estimate_drift(r, pr, DX), DR.estimate_drift(r, pr, DX, swimmer=(40, 4)),
            DR.estimate_drift(r, pr, DX, geometry="2d", swimmer=(40, 4))]
    for o, w in zip(out, want):
        assert o["pairs"] == w.pairs and o["samples"] == w.samples and o["significant"] == w.significant
        for k in ("Vx", "Vy", "mag", "sigma", "angleDeg"):
            assert math.isclose(o[k], getattr(w, k), rel_tol=1e-12, abs_tol=1e-12), (k, o[k], getattr(w, k))
    assert out[1]["pairs"] < out[0]["pairs"]


def test_js_analyze_motility_matches_python(tmp_path):
    """analyzeMotility of the TS-facing module (N-API -> library host code) against the Python wrapper on the
    oracle pipeline's record of a motile scenario, 3d and 2d, with and without the exposure lock (no GPU work)."""
    from tests.test_verdict import oracle_radial
    from hackathon_amd.verdict import analyze_motility
    if not os.path.exists(ADDON):
        pytest.fail("N-API addon not built (python -m hackathon_amd.napi.build_addon)")
    r = oracle_radial("motile_3d")
    rec = dict(N=r.N, NB=r.NB, L=r.L, kCenter=list(map(float, r.kCenter)), kNodes=list(map(float, r.kNodes)),
               nModes=[int(x) for x in r.nModes], D=list(map(float, r.D)), drift=list(map(float, r.drift)))
    (tmp_path / "radial.json").write_text(json.dumps(rec))
    mod = os.path.join(ROOT, "hackathon_amd", "js", "motility-fit.mjs")
    code = ("Promise.all([import('" + mod + "'), import('fs')]).then(([m, fs]) => {"
            " const r = JSON.parse(fs.readFileSync('" + str(tmp_path / "radial.json") + "', 'utf8'));"
            " const out = [m.analyzeMotility(r, {dxUm: 1, fps: 30}), m.analyzeMotility(r, {dxUm: 1, fps: 30, geometry: '2d'}),"
            "  m.analyzeMotility(r, {dxUm: 1, fps: 30, exposureLocked: false})];"
            " console.log('RESULT ' + JSON.stringify(out)); })")
    res = subprocess.run(["node", "--input-type=module", "-e", code], capture_output=True, text=True, timeout=120)
    line = [x for x in res.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, res.stdout + res.stderr
    out = json.loads(line[0][7:])
    want = [analyze_motility(r, 1.0, 30.0), analyze_motility(r, 1.0, 30.0, "2d"),
            analyze_motility(r, 1.0, 30.0, exposureLocked=False)]
    for o, w in zip(out, want):
        assert o["verdict"] == w["verdict"] and o["physicsVerdict"] == w["physicsVerdict"]
        assert [(g["id"], g["pass"]) for g in o["gates"]] == [(g["id"], g["pass"]) for g in w["gates"]]
        for k in ("TSI", "alpha", "vbar", "Z", "Ddiff", "DdiffBrownian", "dBIC", "nEff"):
            assert o[k] == w[k], k
        assert o["drift"]["mag"] == w["drift"]["mag"] and o["bins"] == w["bins"] and o["lags"] == w["lags"]
        assert o["fit"]["swimmer"]["amps"] == w["fit"]["swimmer"]["amps"]
    assert out[0]["verdict"] == "MOTILE" and out[2]["verdict"] == "UNLOCKED_EXPOSURE"



@pytest.mark.gpu
def test_js_engine_calls_leave_the_event_loop_running(gpu):
    """ingestFrames and computeTemporal run off the event loop (the addon's async work): a 1 ms timer keeps
    ticking while a 1024 x 512^2 u16 stack is ingested and analysed twenty times; a second computeTemporal while one runs
    is refused with the reference's message (src/ddm-engine.ts:331-352)."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "async_run.mjs")], capture_output=True, text=True,
                       timeout=240)
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, r.stdout + r.stderr
    out = json.loads(line[-1][7:])
    print(out)
    assert "error" not in out, out
    assert out["ingestTicks"] >= 3 and out["temporalTicks"] >= 3, out
    assert out["reentry"] == "computeTemporal is already running"
    assert out["finite"] and out["d0"] == 0 and out["d1"] > 0
