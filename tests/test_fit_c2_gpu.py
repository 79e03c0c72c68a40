"""The per-q Schulz-model LM at the benchmark size (1024 frames of 512x512 u16, 256 bins, every usable bin)
against the FP64 oracle LM on the same prepared data.

D(q, tau) comes from the engine; prepare (src/motility_fit.js:383-401) picks the bins and the lag grid; both
LMs start from the reference's start grid (the 3 best of src/motility_fit.js:577-611's starts) within the same
bounds.  The per-q landscape is multimodal with slow valleys, so agreement is held bin by bin as rates:
  - both LMs run up to 400 steps per start, the library's default (at a 100-step cap a few slow valleys left
    both mid-descent, where the chi^2 reached depends on the trajectory);
  - from the oracle's own three starts, the GPU LM reaches the oracle LM's chi2 (1e-3) in all but 1 % of bins
    (the LM arithmetic: FP64 parameters, Jacobian, objective and normal equations);
  - from its own start grid, the GPU reaches the oracle's chi2 within 1e-4 in 90 % of bins and is worse than
    1e-3 in at most 2 % (near-equal grid chi2 can rank a different third start from FP32 inputs);
  - where both reach the same chi2, 80 % agree in every parameter to 1e-3 (the rest sit on flat directions:
    alpha near 0 leaves vbar and Z undetermined)."""
import numpy as np
import pytest

import oracle as O
from hackathon_amd import fit as F
from hackathon_amd import synth
from hackathon_amd.engine import DdmEngine
from hackathon_amd.synth import GpuSynth

pytestmark = pytest.mark.gpu


def test_c2_per_bin_schulz_lm_against_fp64_oracle(gpu):
    c2 = dict(synth.C2)
    T, N, NB = c2["T"], c2["N"], 256
    g = GpuSynth()
    fr = g.video(g.params(T, N, c2["n_spots"], c2["sigma"], c2["amp"], c2["background"], D=c2["D"],
                          alpha=c2["alpha"], vbar_px_frame=c2["vbar_px_frame"], Z=c2["Z"], dtype="uint16", seed=2))
    with DdmEngine.create(N=N, T=T, L=T, NB=NB, kmin=0, kmax=N // 2) as e:
        e.ingestDevice(fr.data_ptr(), np.uint16, T, 0)
        e.computeTemporal(np.arange(T, dtype=np.uint32), np.ones(T, np.float32))
        r = e.readRadial()
    pr, res = F.analyzeMotilityLM(r, 1.0, 100.0, model=F.SWIMMER_3D)
    nb = len(pr.bins)
    print(f"{nb} usable bins, GPU LM {res.kernel_ms:.2f} ms (kernel)", flush=True)
    assert nb > 200
    rec = dict(N=r.N, NB=r.NB, L=r.L, D=r.D, nModes=r.nModes, kCenter=r.kCenter, kNodes=r.kNodes)
    opr = O.prepare(rec, 1.0, 100.0)
    assert [b["b"] for b in opr["bins"]] == [b.b for b in pr.bins]
    b4 = O.bounds(opr)[:4]
    # the GPU LM from the oracle's own three starts, against the oracle LM from the same starts
    starts = np.stack([O.lm_starts(opr["tau"], bn["d"], bn["w"], bn["qn"], b4)[0][:3] for bn in opr["bins"]], 1)
    probs = F.problems_from(pr)
    f = F.Fitter()
    from_starts = np.min([f.fit(*probs, model=F.SWIMMER_3D, seeds=starts[s]).chi2 for s in range(3)], axis=0)
    same_start_worse = 0
    for i, bn in enumerate(opr["bins"]):
        oc = min(O.lm_fit(opr["tau"], bn["d"], bn["w"], bn["qn"], b4, starts[s, i], maxit=400)[1] for s in range(3))
        same_start_worse += from_starts[i] > oc * (1 + 1e-3)
    print(f"from the oracle's starts, GPU worse than 1e-3: {same_start_worse}/{nb}", flush=True)
    assert same_start_worse <= 0.01 * nb
    worse, same, agree, rows = [], 0, 0, []
    for i, bn in enumerate(opr["bins"]):
        th, chi, _ = O.lm_plain(opr["tau"], bn["d"], bn["w"], bn["qn"], b4, True, "3d", maxit=400)
        rel = (res.chi2[i] - chi) / max(chi, 1e-300)
        p_gpu = np.array([res.alpha[i], res.vbar[i], res.Z[i], res.D[i]])
        p_orc = np.array([th[2], np.exp(th[3]), np.exp(th[4]), np.exp(th[5])])
        rel_p = np.max(np.abs(p_gpu - p_orc) / np.maximum(np.abs(p_orc), 1e-3))
        rows.append((bn["b"], chi, rel, rel_p))
        if rel > 1e-3:
            worse.append(rows[-1])
        if abs(rel) < 1e-4:
            same += 1
            agree += rel_p < 1e-3
    lower = sum(1 for x in rows if x[2] < -1e-4)
    print(f"chi2 within 1e-4 of the oracle: {same}/{nb}; of those, parameters within 1e-3: {agree}; GPU lower "
          f"chi2: {lower}; GPU worse than 1e-3: {len(worse)}", flush=True)
    for x in rows[:: max(1, nb // 16)] + worse[:10]:
        print("bin %d chi2 %.4e rel_chi %+.2e rel_p %.2e" % x, flush=True)
    assert len(worse) <= 0.02 * nb
    assert same >= 0.9 * nb
    assert agree >= 0.8 * same
