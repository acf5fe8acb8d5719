"""The adjoint of the implicit step (gc_adjoint_step) on the device.

The reference has no gradient, so the oracle here is the derivative of the
oracle's own step: central finite differences (relative step 1e-4) of the
FP64 restatement's rollout loss, with the solver tolerance at 1e-10 so the
difference quotients are not solver noise. The rollout is the small drape
after 14 steps (in contact with the neck: the gradient passes through the
contact filter, the release pass, Coulomb friction and the push-out), with
per-panel stretch, shear and bend factors and the friction coefficient as
parameters.
"""
import numpy as np
import pytest

import oracle as O
from garmentcraft_amd import szenen
from garmentcraft_amd.adjungiert import PanelMaterial, Rollout, rollout_funktion, verlust, verlust_gradient
from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter

pytestmark = pytest.mark.gpu

H = 1.0 / 120.0
VOR = 14        # oracle steps before the rollout (the drape touches the neck)
SCHRITTE = 8    # rollout length
TOL = 1e-10


def param(tol):
    return dict(schritte_je_bild=1, cg_max=20000, cg_toleranz=tol)


@pytest.fixture(scope="module")
def aufbau():
    sdf = szenen.koerper_sdf(aufloesung=96, spacing=0.008)
    cfg = szenen.DrapeKonfig(spalten=2, zeilen=2, zelle_mm=160.0, hoehe_m=0.335, jitter_mm=10.0, versatz_mm=(0.0, 0.0))
    s, _ = szenen.drape(cfg, sdf=sdf)
    o = O.OrakelSim(s, O.param(**param(TOL)))
    for _ in range(VOR):
        o.schritt(H)
    x, v = o.lage()
    dv = o.dv_vorher()
    mat = PanelMaterial(s)
    rng = np.random.default_rng(7)
    th0 = mat.theta0()
    th_ziel = th0 * np.concatenate([rng.uniform(0.7, 1.3, mat.n_param - 1), [1.0]])
    th_ziel[-1] = 0.4
    return {"s": s, "x": x, "v": v, "dv": dv, "mat": mat, "th0": th0, "th_ziel": th_ziel}


def orakel_rollout(a, theta):
    s = a["mat"].szene(theta)
    o = O.OrakelSim(s, O.param(**param(TOL)))
    o.setze_lage(a["x"], a["v"])
    o.setze_dv_vorher(a["dv"])
    for _ in range(SCHRITTE):
        o.schritt(H)
    return o.lage()[0]


def geraet_rollout(a, praezision, tol=TOL, alle=3):
    g = GeraetSzene(a["s"], ImplizitParameter(**param(tol)), praezision=praezision)
    return g, Rollout(g, a["s"], a["mat"], H, SCHRITTE, alle=alle, x0=a["x"], v0=a["v"], dv_vorher=a["dv"])


@pytest.fixture(scope="module")
def fd(aufbau):
    """Central differences of the oracle's loss at theta0 (target from theta_ziel)."""
    a = aufbau
    ziel = orakel_rollout(a, a["th_ziel"])
    th0 = a["th0"]
    x0 = orakel_rollout(a, th0)
    grad = np.zeros_like(th0)
    for k in range(th0.size):
        hk = 1e-4 * abs(th0[k])
        tp, tm = th0.copy(), th0.copy()
        tp[k] += hk
        tm[k] -= hk
        grad[k] = (verlust(orakel_rollout(a, tp), ziel) - verlust(orakel_rollout(a, tm), ziel)) / (2 * hk)
    return {"ziel": ziel, "x0": x0, "grad": grad}


def test_adjoint_gleich_finite_differenzen_des_orakels(aufbau, fd):
    a = aufbau
    g, r = geraet_rollout(a, "fp64")
    x = r.vorwaerts(a["th0"])
    print("forward |x_gpu - x_oracle| =", np.max(np.abs(x - fd["x0"])))
    assert np.max(np.abs(x - fd["x0"])) <= 1e-8
    st = g.statistik()
    assert st["kontakte"] > 0 and st["haftend"] > 0, st
    d = r.rueckwaerts(verlust_gradient(x, fd["ziel"]))
    print("adjoint stats", r.statistik)
    assert r.statistik["abweichungen"] == 0
    ref = fd["grad"]
    fehler = np.linalg.norm(d - ref) / np.linalg.norm(ref)
    for k in range(ref.size):
        print(f"theta[{k}]: adjoint {d[k]: .6e}  fd {ref[k]: .6e}")
    print("relative error", fehler)
    assert fehler <= 1e-3
    gross = np.abs(ref) > 1e-2 * np.max(np.abs(ref))
    assert np.all(np.abs(d - ref)[gross] <= 1e-2 * np.abs(ref)[gross])


def test_fp32_adjoint_nahe_fp64(aufbau, fd):
    """FP32 bulk (FP64 reductions) against the all-FP64 adjoint, both at the
    drape's tolerance 1e-6."""
    a = aufbau
    out = {}
    for pr in ("fp64", "fp32"):
        g, r = geraet_rollout(a, pr, tol=1e-6)
        x = r.vorwaerts(a["th0"])
        out[pr] = (x, r.rueckwaerts(verlust_gradient(x, fd["ziel"])))
    dx = np.max(np.abs(out["fp32"][0] - out["fp64"][0]))
    fehler = np.linalg.norm(out["fp32"][1] - out["fp64"][1]) / np.linalg.norm(out["fp64"][1])
    print("fp32 vs fp64: positions", dx, "gradient", fehler)
    assert dx <= 1e-4
    assert fehler <= 2e-3
    # against the oracle's differences at tolerance 1e-10 the trajectory at
    # 1e-6 itself differs (the loss gradient scales with x - x*)
    fehler_fd = np.linalg.norm(out["fp32"][1] - fd["grad"]) / np.linalg.norm(fd["grad"])
    fehler_fd64 = np.linalg.norm(out["fp64"][1] - fd["grad"]) / np.linalg.norm(fd["grad"])
    print("vs oracle fd at 1e-10: fp32", fehler_fd, "fp64", fehler_fd64)
    assert fehler_fd <= 1e-2


def test_checkpoint_abstand_aendert_nichts(aufbau, fd):
    a = aufbau
    gr = []
    for alle in (1, 3, SCHRITTE):
        g, r = geraet_rollout(a, "fp32", tol=1e-6, alle=alle)
        x = r.vorwaerts(a["th0"])
        gr.append(r.rueckwaerts(verlust_gradient(x, fd["ziel"])))
    assert np.array_equal(gr[0], gr[1]) and np.array_equal(gr[0], gr[2])


@pytest.mark.parametrize("praezision", ["fp64", "fp32"])
def test_wiedergabe_wie_neu_geloest(aufbau, fd, praezision):
    """The reverse pass with the segment re-run's solutions replayed
    (gc_step_aufnehmen, gc_adjoint_step_aus) gives the gradient and the start
    state's adjoint of the pass that solves every step again, to the bit."""
    a = aufbau
    out = []
    for wiedergabe in (True, False):
        g, r = geraet_rollout(a, praezision, tol=1e-8, alle=SCHRITTE)
        r.wiedergabe = wiedergabe
        x = r.vorwaerts(a["th0"])
        d = r.rueckwaerts(verlust_gradient(x, fd["ziel"]))
        out.append((d, *r.anfangsgradient(), r.statistik["cg_iterationen"], r.statistik["abweichungen"]))
        g.close()
    print(praezision, "adjoint CG iterations", out[0][3], out[1][3], "deviations", out[0][4], out[1][4])
    for u, w in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(u, w)
    assert out[0][4] == 0 and out[1][4] == 0


def test_torch_autograd_wie_direkt(aufbau, fd):
    import torch
    a = aufbau
    g, r = geraet_rollout(a, "fp32", tol=1e-6)
    theta = torch.tensor(a["th0"], dtype=torch.float64, requires_grad=True)
    x = rollout_funktion().apply(theta, r)
    ziel = torch.from_numpy(fd["ziel"])
    loss = ((x - ziel) ** 2).sum(dim=1).mean()
    loss.backward()
    g2, r2 = geraet_rollout(a, "fp32", tol=1e-6)
    x2 = r2.vorwaerts(a["th0"])
    d2 = r2.rueckwaerts(verlust_gradient(x2, fd["ziel"]))
    assert np.array_equal(x.detach().numpy(), x2)
    assert np.allclose(theta.grad.numpy(), d2, rtol=1e-12, atol=0.0)


def test_steifigkeit_setzen_wie_neu_gebaut(aufbau):
    """gc_set_steifigkeit with theta's values steps like a context built from
    the scene that carries them."""
    a = aufbau
    th = a["th_ziel"]
    s2 = a["mat"].szene(th)
    out = []
    for variante in ("gesetzt", "gebaut"):
        if variante == "gesetzt":
            g = GeraetSzene(a["s"], ImplizitParameter(**param(1e-8)), praezision="fp64")
            md, _, rb = a["mat"].werte(th)
            # the stiffness a context derives from the copy's compliance
            g.set_steifigkeit(membran_ek_es_g=md, biegung_k=1.0 / s2.biegung_nachgiebigkeit, reibung=rb)
        else:
            g = GeraetSzene(s2, ImplizitParameter(**param(1e-8)), praezision="fp64")
        g.set_state(a["x"], a["v"])
        g.set_dv_vorher(a["dv"])
        g.schritte(H, 3)
        out.append(g.lage()[0])
    assert np.max(np.abs(out[0] - out[1])) <= 1e-12


def test_adjoint_verweigert_selbstkollision(aufbau):
    from garmentcraft_amd._lib import GcFehler
    a = aufbau
    p = ImplizitParameter(**param(1e-6))
    p.selbstkollision = True
    g = GeraetSzene(a["s"], p, praezision="fp32")
    g.adjoint_setzen(None, None)
    with pytest.raises(GcFehler, match="self-collision"):
        g.adjoint_schritt(H)
