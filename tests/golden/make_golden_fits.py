"""Writes tests/golden/fits_reference.json: the reference's own fit outputs on noisy records, its isf2dFast
table values and movies from its simulator (the pins nothing but the reference's code provides).

Run in the build container only, with the reference checked out at /root/reference:
    python tests/golden/make_golden_fits.py
The input records are the FP64 oracle pipeline's Radial records (simulator -> history -> masked WK radial
record with drift sums) of six validation-matrix scenarios, the same records tests/test_golden_fits.py
rebuilds; make_golden_fits.mjs runs the reference's prepare / fitBrownian / fitSwimmer / estimateDrift /
analyzeMotility on them under Node and stores only the outputs.
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

RECORDS = ["motile_3d", "motile_2d", "killed_brownian", "flow_0deg_10ums", "motile_in_flow", "dropped_frames_20pct"]
SIMS = [
    dict(N=32, T=6, fps=30, dxUm=1, seed=5, Dum2s=0.4, blobSigmaPx=1.5, amp=0.15, background=0.25, noise=0.01,
         swimmers=dict(n=6, vbar=20, sigma=8, geometry="3d"), passive=dict(n=5), fixed=dict(n=3),
         drift=dict(speed=10, angleDeg=30), dropExact=2),
    dict(N=24, T=5, fps=50, dxUm=0.8, seed=9, Dum2s=1.1, blobSigmaPx=1.2, amp=0.2, background=0.3, noise=0.02,
         swimmers=dict(n=4, vbar=25, sigma=10, geometry="2d"), passive=dict(n=6), dropFrac=0.3),
]


def main():
    from tests.scenarios import SCENARIOS
    from tests.test_verdict import oracle_radial
    recs = []
    for name in RECORDS:
        sc = next(s for s in SCENARIOS if s["name"] == name)
        r = oracle_radial(name)
        recs.append(dict(name=name, geometry=sc["geometry"], dxUm=sc["sim"]["dxUm"], fps=sc["sim"]["fps"], N=int(r.N),
                         NB=int(r.NB), L=int(r.L), kCenter=[float(x) for x in r.kCenter],
                         kNodes=[float(x) for x in r.kNodes], nModes=[int(x) for x in r.nModes],
                         D=[float(x) for x in r.D], drift=[float(x) for x in r.drift]))
        print("record", name, flush=True)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "records.json"), "w") as f:
            json.dump(dict(records=recs, sims=SIMS), f)
        subprocess.run(["node", os.path.join(HERE, "make_golden_fits.mjs"), d], check=True)


if __name__ == "__main__":
    main()
