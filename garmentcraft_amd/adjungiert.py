"""Gradients of an implicit-step rollout with respect to material parameters.

The reference has no gradient. This module runs ``ImplizitCg::schritt``
(implizit.rs:899-1038) forward on the device and back through
``gc_adjoint_step`` (the discrete adjoint of the step, include/gcsim.h), so that
per-panel stiffnesses and the friction coefficient can be fitted to a target
drape.

* :class:`PanelMaterial` maps the parameter vector
  ``theta = (stretch[P], shear[P], bend[P], friction)`` onto the scene's
  per-element values: a panel's stretch factor scales E_k and E_s of its
  membranes (membran_daten fields 9 and 10), its shear factor G (field 11),
  its bend factor the stiffness of its bending constraints (a constraint
  belongs to the panel of its first wing point, biegung_punkte[:, 2]); the
  friction coefficient is every point's. The per-element gradients of the
  adjoint are summed back onto theta by the chain rule.
* :class:`Rollout` runs n steps from a fixed start state, saving the state
  every ``alle`` steps in device memory; its reverse pass re-runs each
  segment from its checkpoint, keeping each step's solutions, and steps the
  adjoint back through it with those solutions.
* :class:`RolloutFunktion` wraps a rollout as a ``torch.autograd.Function``
  (theta -> final positions) for checks with torch.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .geraet import GeraetSzene
from .szene import Szene

START_SLOT = -1
TEMP_SLOT = 1 << 20


class PanelMaterial:
    """theta = (stretch factor per panel, shear factor per panel, bend factor
    per panel, friction coefficient) on a scene's elements."""

    def __init__(self, s: Szene):
        self.s = s
        self.panele = int(s.teil_von_punkt.max()) + 1
        self.panel_m = s.teil_von_punkt[s.membran_punkte[:, 0]].astype(np.int64)
        self.panel_b = s.teil_von_punkt[s.biegung_punkte[:, 2]].astype(np.int64)
        self.md0 = np.array(s.membran_daten[:, 9:12], dtype=np.float64)
        self.bk0 = 1.0 / np.asarray(s.biegung_nachgiebigkeit, dtype=np.float64)

    @property
    def n_param(self) -> int:
        return 3 * self.panele + 1

    def theta0(self, reibung: Optional[float] = None) -> np.ndarray:
        th = np.ones(self.n_param)
        th[-1] = float(self.s.reibung[0]) if reibung is None else reibung
        return th

    def werte(self, theta):
        """(membran E_k E_s G per element, bending k per constraint, friction per point)."""
        th = np.asarray(theta, dtype=np.float64)
        P = self.panele
        st, sh, be, mu = th[:P], th[P:2 * P], th[2 * P:3 * P], th[3 * P]
        md = self.md0 * np.column_stack([st[self.panel_m], st[self.panel_m], sh[self.panel_m]])
        bk = self.bk0 * be[self.panel_b]
        reib = np.full(self.s.n, mu)
        return md, bk, reib

    def szene(self, theta) -> Szene:
        """A copy of the scene with theta's values (for the oracle)."""
        md, bk, reib = self.werte(theta)
        s = self.s.kopie()
        s.membran_daten = np.array(self.s.membran_daten, dtype=np.float64)
        s.membran_daten[:, 9:12] = md
        s.biegung_nachgiebigkeit = 1.0 / bk
        s.reibung = reib
        return s

    def zurueck(self, grad: dict) -> np.ndarray:
        """dL/dtheta from the per-element gradients of gc_adjoint_gradient."""
        P = self.panele
        dm, db, dr = grad["membran"], grad["biegung"], grad["reibung"]
        out = np.zeros(self.n_param)
        out[:P] = np.bincount(self.panel_m, weights=self.md0[:, 0] * dm[:, 0] + self.md0[:, 1] * dm[:, 1], minlength=P)
        out[P:2 * P] = np.bincount(self.panel_m, weights=self.md0[:, 2] * dm[:, 2], minlength=P)
        out[2 * P:3 * P] = np.bincount(self.panel_b, weights=self.bk0 * db, minlength=P)
        out[3 * P] = float(np.sum(dr))
        return out


class Rollout:
    """n steps of length h from the scene's state on a device context, with a
    checkpoint every ``alle`` steps (slots of gc_zustand_sichern) and the
    adjoint pass back through them. With ``wiedergabe`` (the default) the
    reverse pass's segment re-run records each step's solutions and the
    adjoint step replays them (gc_step_aufnehmen, gc_adjoint_step_aus), the
    same gradient to the bit with one forward solve less per step."""

    def __init__(self, g: GeraetSzene, s: Szene, material: PanelMaterial, h: float, schritte: int, alle: int = 10,
                 x0=None, v0=None, dv_vorher=None, wiedergabe: bool = True):
        if schritte < 1 or alle < 1:
            raise ValueError("schritte and alle must be >= 1")
        self.g, self.s, self.material, self.h, self.schritte, self.alle = g, s, material, h, schritte, alle
        self.wiedergabe = wiedergabe
        g.set_state(s.x if x0 is None else x0, s.v if v0 is None else v0)
        g.set_dv_vorher(dv_vorher)
        g.sichern(START_SLOT)
        self.theta = None

    def _setzen(self, theta) -> None:
        md, bk, reib = self.material.werte(theta)
        self.g.set_steifigkeit(membran_ek_es_g=md, biegung_k=bk, reibung=reib)
        self.theta = np.array(theta, dtype=np.float64)

    def vorwaerts(self, theta, halten: Optional[str] = None) -> np.ndarray:
        """Final positions (n, 3) after the rollout with theta's material.

        halten="aufnehmen" records every step's discrete decisions (contacts,
        releases, stick or slip, path entries, push-outs), halten="abspielen"
        replays the recorded ones instead of deciding: finite differences of
        the branch the adjoint differentiates."""
        self._setzen(theta)
        g = self.g
        g.laden(START_SLOT)
        modus = {None: 0, "aufnehmen": 1, "abspielen": 2}[halten]
        try:
            for k in range(self.schritte):
                if k % self.alle == 0:
                    g.sichern(k // self.alle)
                if modus:
                    g.entscheidungen(modus, k)
                g.schritte(self.h, 1)
        finally:
            if modus:
                g.entscheidungen(0)
        return g.lage()[0]

    def rueckwaerts(self, gx, gv=None) -> np.ndarray:
        """dL/dtheta for dL/dx (and dL/dv) of the final state, after vorwaerts."""
        if self.theta is None:
            raise RuntimeError("rueckwaerts needs a vorwaerts first")
        g = self.g
        g.adjoint_gradient(self.s, reset=True)
        g.adjoint_statistik(reset=True)
        g.adjoint_setzen(gx, gv)
        segmente = (self.schritte + self.alle - 1) // self.alle
        for seg in reversed(range(segmente)):
            n0, n1 = seg * self.alle, min((seg + 1) * self.alle, self.schritte)
            g.laden(seg)
            # the segment re-run keeps each step's solutions with the state
            # before it, so that the adjoint step does not solve it again (the
            # segment's last step is taken only by its adjoint step)
            for k in range(n0, n1):
                if k + 1 < n1 and self.wiedergabe:
                    g.schritt_aufnehmen(self.h, TEMP_SLOT + k - n0)
                else:
                    g.sichern(TEMP_SLOT + k - n0)
                    if k + 1 < n1:
                        g.schritte(self.h, 1)
            for k in reversed(range(n0, n1)):
                if k + 1 < n1 and self.wiedergabe:
                    g.adjoint_schritt_aus(self.h, TEMP_SLOT + k - n0)
                else:
                    g.laden(TEMP_SLOT + k - n0)
                    g.adjoint_schritt(self.h)
        self.statistik = g.adjoint_statistik()
        return self.material.zurueck(g.adjoint_gradient(self.s, reset=True))

    def anfangsgradient(self):
        """dL/dx0, dL/dv0 of the start state after rueckwaerts."""
        return self.g.adjoint_lesen()


def verlust(x, ziel) -> float:
    """Mean squared point distance, sum_i |x_i - x*_i|^2 / n."""
    d = np.asarray(x) - np.asarray(ziel)
    return float(np.sum(d * d) / d.shape[0])


def verlust_gradient(x, ziel) -> np.ndarray:
    return 2.0 * (np.asarray(x) - np.asarray(ziel)) / np.asarray(x).shape[0]


def _torch_funktion():
    import torch

    class RolloutFunktion(torch.autograd.Function):
        """theta (float64 CPU tensor) -> final positions (n, 3); backward runs
        the adjoint rollout."""

        @staticmethod
        def forward(ctx, theta, rollout):
            x = rollout.vorwaerts(theta.detach().cpu().numpy().astype(np.float64))
            ctx.rollout = rollout
            return torch.from_numpy(x)

        @staticmethod
        def backward(ctx, gx):
            d = ctx.rollout.rueckwaerts(gx.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(d), None

    return RolloutFunktion


def rollout_funktion():
    """The torch.autograd.Function class (torch imported on first use)."""
    global RolloutFunktion
    if "RolloutFunktion" not in globals():
        RolloutFunktion = _torch_funktion()
    return RolloutFunktion
