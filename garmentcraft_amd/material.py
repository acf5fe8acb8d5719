"""Fabric constants in SI units (material.rs:21-124).

The field names follow the reference's ``Material`` struct: ``dehnung_n_m`` is
the weft stretch stiffness, ``kette_n_m`` the warp one, ``schub_n_m`` the shear
one (all N/m); ``biegung_nm`` / ``biegung_kette_nm`` / ``biegung_schraege_nm``
are the weft / warp / bias bending stiffnesses (N*m).
"""
from __future__ import annotations

from dataclasses import dataclass, replace


@dataclass(frozen=True)
class Material:
    """material.rs:21-68."""
    name: str
    dehnung_n_m: float
    kette_n_m: float
    schub_n_m: float
    biegung_nm: float
    biegung_kette_nm: float
    biegung_schraege_nm: float
    flaechengewicht_kg_m2: float
    dicke_m: float
    reibung: float

    def mit_dicke(self, dicke: float) -> "Material":
        """material.rs:111-124: stretch and mass scale with the thickness,
        bending with its cube."""
        f = dicke / self.dicke_m
        return replace(
            self,
            dehnung_n_m=self.dehnung_n_m * f,
            kette_n_m=self.kette_n_m * f,
            schub_n_m=self.schub_n_m * f,
            biegung_nm=self.biegung_nm * f ** 3,
            biegung_kette_nm=self.biegung_kette_nm * f ** 3,
            biegung_schraege_nm=self.biegung_schraege_nm * f ** 3,
            flaechengewicht_kg_m2=self.flaechengewicht_kg_m2 * f,
            dicke_m=dicke,
        )


# material.rs:73-84
RADICAL_RUBBER_04 = Material(
    name="Latex_0.4mm_Radical_Rubber",
    dehnung_n_m=348.2741562,
    kette_n_m=348.2817188,
    schub_n_m=174.1590469,
    biegung_nm=2482.2424e-9,
    biegung_kette_nm=2486.4924e-9,
    biegung_schraege_nm=2486.4924e-9,
    flaechengewicht_kg_m2=0.3489899,
    dicke_m=0.4e-3,
    reibung=0.03,
)

# material.rs:89-102
RETIFIST_NR_040 = Material(
    name="NR-sheet-0.40 (Retifist)",
    dehnung_n_m=33.333332,
    kette_n_m=33.333332,
    schub_n_m=6.6666665,
    biegung_nm=60.361e-9,
    biegung_kette_nm=60.361e-9,
    biegung_schraege_nm=60.361e-9,
    flaechengewicht_kg_m2=0.380,
    dicke_m=0.4e-3,
    reibung=0.03,
)


def latex(e_mpa: float, dicke_m: float, dichte_kg_m3: float, biegung_nm: float, reibung: float = 0.03) -> Material:
    """Isotropic sheet from Young's modulus with nu = 0: membrane E*t in both
    directions and G*t = E*t/2 (so that nu = E/(2G) - 1 = 0, kontinuum.rs:29-31)."""
    et = e_mpa * 1e6 * dicke_m
    return Material(
        name=f"latex E={e_mpa} MPa t={dicke_m * 1e3} mm",
        dehnung_n_m=et, kette_n_m=et, schub_n_m=0.5 * et,
        biegung_nm=biegung_nm, biegung_kette_nm=biegung_nm, biegung_schraege_nm=biegung_nm,
        flaechengewicht_kg_m2=dichte_kg_m3 * dicke_m, dicke_m=dicke_m, reibung=reibung,
    )


def gewebe(kette_mpa: float, schuss_mpa: float, schub_mpa: float, dicke_m: float, flaechengewicht: float,
           biegung_kette: float, biegung_schuss: float, biegung_schraege: float, reibung: float) -> Material:
    """Woven orthotropic fabric from warp/weft/shear moduli (MPa) and a thickness."""
    return Material(
        name=f"woven {kette_mpa}/{schuss_mpa}/{schub_mpa} MPa t={dicke_m * 1e3} mm",
        dehnung_n_m=schuss_mpa * 1e6 * dicke_m, kette_n_m=kette_mpa * 1e6 * dicke_m,
        schub_n_m=schub_mpa * 1e6 * dicke_m,
        biegung_nm=biegung_schuss, biegung_kette_nm=biegung_kette, biegung_schraege_nm=biegung_schraege,
        flaechengewicht_kg_m2=flaechengewicht, dicke_m=dicke_m, reibung=reibung,
    )
