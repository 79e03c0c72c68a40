"""The reference's 11-scenario validation matrix (motility-ddm test/sim.mjs:207-305) as data:
simulator parameters, analysis geometry, the required verdict and the quantitative checks."""

BASE = dict(N=256, T=256, fps=30, dxUm=1, Dum2s=0.4, blobSigmaPx=1.5, amp=0.15, background=0.25, noise=0.01)
SWIM3D = dict(n=150, vbar=20, sigma=8, geometry="3d")


def _sim(**kw):
    d = dict(BASE)
    d.update(kw)
    return d


SCENARIOS = [
    dict(name="motile_3d", sim=_sim(seed=11, swimmers=SWIM3D, passive=dict(n=150)), geometry="3d", expect="MOTILE",
         checks=dict(vbar=(20, 0.2), alpha=(0.5, 0.2))),
    dict(name="motile_2d", sim=_sim(seed=12, swimmers=dict(n=150, vbar=25, sigma=10, geometry="2d"),
                                    passive=dict(n=100)), geometry="2d", expect="MOTILE",
         checks=dict(vbar=(25, 0.2), alpha=(0.6, 0.2))),
    dict(name="sparse_swimmers", sim=_sim(seed=13, swimmers=dict(n=25, vbar=20, sigma=8, geometry="3d"),
                                          passive=dict(n=300)), geometry="3d", expect="MOTILE",
         checks=dict(alpha=(25 / 325, 0.07))),
    dict(name="killed_brownian", sim=_sim(seed=14, passive=dict(n=300)), geometry="3d", expect="NOT_MOTILE", checks={}),
    dict(name="static_debris", sim=_sim(seed=15, fixed=dict(n=300)), geometry="3d", expect="NO_DYNAMICS", checks={}),
    dict(name="shuffle_null", sim=_sim(seed=11, swimmers=SWIM3D, passive=dict(n=150)), geometry="3d", shuffle=True,
         expect="NO_DYNAMICS", checks={}),
    dict(name="flow_0deg_10ums", sim=_sim(seed=17, passive=dict(n=250), drift=dict(speed=10, angleDeg=0)),
         geometry="3d", expect="DIRECTED_TRANSPORT", checks=dict(drift=(10, 0, 0.15, 10))),
    dict(name="flow_135deg_5ums", sim=_sim(seed=18, passive=dict(n=250), drift=dict(speed=5, angleDeg=135)),
         geometry="3d", expect="DIRECTED_TRANSPORT", checks=dict(drift=(5, 135, 0.15, 10))),
    dict(name="flow_60deg_1ums", sim=_sim(seed=19, passive=dict(n=250), drift=dict(speed=1, angleDeg=60)),
         geometry="3d", expect="NOT_MOTILE", checks={}),
    dict(name="motile_in_flow", sim=_sim(seed=20, swimmers=SWIM3D, passive=dict(n=150),
                                         drift=dict(speed=8, angleDeg=210)),
         geometry="3d", expect="DIRECTED_TRANSPORT", checks=dict(drift=(8, 210, 0.2, 15))),
    dict(name="dropped_frames_20pct", sim=_sim(seed=21, swimmers=SWIM3D, passive=dict(n=150), dropFrac=0.2),
         geometry="3d", expect="MOTILE", checks=dict(vbar=(20, 0.2), alpha=(0.5, 0.2))),
]


def check(sc, res):
    """checkScenario (test/t_validation_matrix.mjs:51-73): the failures, empty when the scenario passes."""
    fails = []
    if res["verdict"] != sc["expect"]:
        fails.append(f"verdict {res['verdict']}, expected {sc['expect']}")
    c = sc["checks"]
    if "vbar" in c:
        v, tol = c["vbar"]
        if not (abs(res.get("vbar", float("nan")) - v) / v <= tol):
            fails.append(f"vbar {res.get('vbar')}, expected {v} +/- {tol * 100}%")
    if "alpha" in c:
        a, tol = c["alpha"]
        if not (abs(res.get("alpha", float("nan")) - a) <= tol):
            fails.append(f"alpha {res.get('alpha')}, expected {a} +/- {tol}")
    if "drift" in c:
        speed, ang, rtol, dtol = c["drift"]
        d = res["drift"]
        da = abs(d["angleDeg"] - ang) % 360
        da = 360 - da if da > 180 else da
        if not (abs(d["mag"] - speed) / speed <= rtol):
            fails.append(f"|V| {d['mag']}, expected {speed} +/- {rtol * 100}%")
        if not (da <= dtol):
            fails.append(f"V angle {d['angleDeg']}, expected {ang} +/- {dtol}")
    return fails
