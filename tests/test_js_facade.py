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
