"""Shared transport cases for the oracle and GPU parity tests."""

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut

# the reference's synthetic SRD 64-format Mott tables (tests/data/mott_srd64_synthetic)
MOTT_FIXTURE_DIR = Path(__file__).resolve().parent / "golden" / "mott_srd64_synthetic"
# the reference's ELSEPA vendor panel (Hg, 1 keV; tests/data/xsgen/elsepa/test-run-output)
ELSEPA_PANEL = Path(__file__).resolve().parent / "golden" / "elsepa_dcs_1p000e03.dat"


def sr_cross_section_cm2(Z, E_keV):
    """Total cross section of the analytic screened-Rutherford model."""
    a = 3.4e-3 * Z**0.67 / E_keV
    rel = (E_keV + 511.0) / (E_keV + 1024.0)
    return 5.21e-21 * Z * Z * 4.0 * np.pi / E_keV**2 / (a * (1.0 + a)) * rel**2


def elsepa_test_tables(layers, energies_keV=None, vendor=("W",)):
    """ELSEPA-format tables on the vendor panel's angular grid: the elements in
    ``vendor`` carry the vendor DCS shape (a real, non-screened-Rutherford law)
    at every energy, the others the screened-Rutherford law's own DCS; all with
    the screened-Rutherford total cross section."""
    from pyrite_amd.tables.elsepa import read_dcs
    from pyrite_amd.transport.elements import TRANSPORT_ELEMENTS

    panel = read_dcs(ELSEPA_PANEL)
    E = np.geomspace(4.0, 45.0, 30) if energies_keV is None else np.asarray(energies_keV, dtype=float)
    out = []
    for _, _, comp in layers:
        row = []
        for el, _ in comp:
            Z = float(TRANSPORT_ELEMENTS[el]["Z"])
            sigma = sr_cross_section_cm2(Z, E)
            if el in vendor:
                dcs = np.tile(panel.dcs_cm2_sr, (E.size, 1))
            else:
                a = 3.4e-3 * Z**0.67 / E
                dcs = (sigma / (4.0 * np.pi) * a * (1.0 + a))[:, None] / (panel.mu[None, :] + a[:, None]) ** 2
            row.append({"energy_eV": E * 1e3, "total_elastic_cm2": sigma, "mu": panel.mu, "dcs_cm2_sr": dcs})
        out.append(row)
    return out


@dataclass
class Case:
    name: str
    E0: float
    Ne: int
    layers: list
    seed: int
    E_cut: float = 5.0
    straggle: bool = False
    energy_model: int = 1
    max_dE_frac: float = 0.0
    beam_dir: tuple = (0.0, 0.0, 1.0)
    finite: tuple | None = None
    elastic: str = "sr"
    atomic: str | None = None  # atomic_electron_deflection
    elsepa_energies: tuple | None = None  # (lo keV, hi keV, nodes) of the ELSEPA test tables
    elsepa_vendor: tuple = ("W",)  # elements on the vendor DCS shape

    def tables(self):
        extra = {}
        if self.elastic == "elsepa":
            E = None if self.elsepa_energies is None else np.geomspace(*self.elsepa_energies)
            extra = {"elastic_tables": elsepa_test_tables(self.layers, E, self.elsepa_vendor)}
        lt = build_layer_tables(self.layers, self.elastic, MOTT_FIXTURE_DIR, self.atomic, **extra)
        lut = build_transport_lut(self.E_cut, self.E0, lt)
        return lt, lut

    def state(self):
        bd = np.asarray(self.beam_dir, dtype=float)
        bd = bd / np.linalg.norm(bd)
        pos = np.zeros((self.Ne, 3))
        dirs = np.tile(bd, (self.Ne, 1))
        E = np.full(self.Ne, self.E0)
        E_cut = np.full(self.Ne, self.E_cut)
        return pos, dirs, E, E_cut


SI_30KEV = Case("si30", 30.0, 400, [(0.0, 5.0e4, [("Si", 0.04996)])], seed=11)
SI_30KEV_STRAGGLE = Case("si30s", 30.0, 400, [(0.0, 5.0e4, [("Si", 0.04996)])], seed=11, straggle=True)
HOPG_5MEV = Case(
    "hopg5", 5000.0, 300, [(0.0, 1.0e6, [("C", 0.1136)])], seed=1, straggle=True,
    beam_dir=(-np.sin(np.deg2rad(0.5)), 0.0, np.cos(np.deg2rad(0.5))),
)
MULTILAYER = Case(
    "stack", 40.0, 300,
    [(0.0, 2.0e3, [("W", 0.0632)]), (2.0e3, 3.0e4, [("Si", 0.025), ("C", 0.05)])],
    seed=5, straggle=True, max_dE_frac=0.05,
)
FRAMED = Case(
    "prism", 60.0, 300, [(0.0, 4.0e4, [("Ge", 0.0442)])], seed=3,
    beam_dir=(0.3, 0.1, 0.9), finite=(2.0e4, 1.5e4), energy_model=0,
)
MULTILAYER_MOTT = Case(
    "stack_mott", 40.0, 300,
    [(0.0, 2.0e3, [("W", 0.0632)]), (2.0e3, 3.0e4, [("Si", 0.025), ("C", 0.05)])],
    seed=6, straggle=True, max_dE_frac=0.05, elastic="mott",
)
SI_30KEV_ATOMIC = Case("si30_aed", 30.0, 400, [(0.0, 5.0e4, [("Si", 0.04996)])], seed=11, straggle=True,
                       atomic="kawrakow")
MULTILAYER_ELSEPA = Case(
    "stack_elsepa", 40.0, 300,
    [(0.0, 2.0e3, [("W", 0.0632)]), (2.0e3, 3.0e4, [("Si", 0.025), ("C", 0.05)])],
    seed=7, straggle=True, max_dE_frac=0.05, elastic="elsepa",
)
CASES = [SI_30KEV, SI_30KEV_STRAGGLE, HOPG_5MEV, MULTILAYER, FRAMED, MULTILAYER_MOTT, SI_30KEV_ATOMIC,
         MULTILAYER_ELSEPA]
