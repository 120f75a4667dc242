"""Mechanism data model: parses the Cantera-style YAML files under ``mech/``.

The reference compiles one mechanism per build (``Chemistry_Model`` selects
``PelePhysics/Mechanisms/<model>/mechanism.{H,cpp}``, CMake/BuildPelePhysicsLib.cmake:113-116);
those generated files are absent from the reference tree, so the mechanism is carried here as
data and turned into device code by :mod:`pelec_amd.mechgen`.

Units are CGS as in PeleC (Exec/RegTests/zeroD/prob_parm.H:9 uses p = 1013250 dyn/cm^2):
A in (cm^3/mol)^(order-1)/s, activation energies converted from cal/mol to K (Ea/R).
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
import re
from typing import Dict, List, Optional

import yaml

#: universal gas constant, erg/(mol K)
R_ERG = 8.31446261815324e7
#: gas constant in cal/(mol K) (R / 4.184 J/cal)
R_CAL = 8.31446261815324 / 4.184
#: reference pressure of the NASA-7 fits, dyn/cm^2 (1 atm)
P_ATM = 1013250.0

MECH_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mech")


@dataclasses.dataclass
class Species:
    name: str
    composition: Dict[str, float]
    W: float
    Tmin: float
    Tmid: float
    Tmax: float
    low: List[float]
    high: List[float]


@dataclasses.dataclass
class Reaction:
    equation: str
    reactants: Dict[int, float]      # species index -> stoichiometric coefficient
    products: Dict[int, float]
    reversible: bool
    kind: str                        # "elementary" | "three-body" | "falloff"
    A: float                         # (high-pressure) pre-exponential factor
    b: float
    Ea_R: float                      # activation temperature, K
    low: Optional[List[float]] = None    # falloff: [A0, b0, Ea0/R]
    troe: Optional[List[float]] = None   # [a, T***, T*, (T**)]
    efficiencies: Optional[List[float]] = None  # per species, default 1.0
    duplicate: bool = False


@dataclasses.dataclass
class Mechanism:
    name: str
    elements: List[str]
    element_weights: List[float]
    species: List[Species]
    reactions: List[Reaction]
    source_sha256: str

    @property
    def nsp(self) -> int:
        return len(self.species)

    @property
    def nrxn(self) -> int:
        return len(self.reactions)

    def species_index(self, name: str) -> int:
        for i, s in enumerate(self.species):
            if s.name == name:
                return i
        raise KeyError(name)

    @property
    def weights(self) -> List[float]:
        return [s.W for s in self.species]

    def element_matrix(self):
        """a[e][k]: atoms of element e in species k."""
        return [[s.composition.get(e, 0.0) for s in self.species] for e in self.elements]


_TERM = re.compile(r"^\s*(\d+(?:\.\d+)?)?\s*(.+?)\s*$")


def _parse_side(side: str, names: Dict[str, int]):
    """Parse one side of an equation into ({species: coeff}, collider) where collider is
    None, 'M' (three-body) or '(+M)' (falloff)."""
    collider = None
    s = side.strip()
    if s.endswith("(+M)"):
        collider = "(+M)"
        s = s[: -len("(+M)")].strip()
    terms = [t for t in re.split(r"\s+\+\s+", s) if t.strip()]
    out: Dict[int, float] = {}
    for t in terms:
        t = t.strip()
        if t == "M":
            collider = "M"
            continue
        m = _TERM.match(t)
        coeff = float(m.group(1)) if m.group(1) else 1.0
        name = m.group(2)
        if name not in names:
            raise ValueError(f"unknown species '{name}' in '{side}'")
        k = names[name]
        out[k] = out.get(k, 0.0) + coeff
    return out, collider


def load(name_or_path: str) -> Mechanism:
    path = name_or_path
    if not os.path.exists(path):
        path = os.path.join(MECH_DIR, f"{name_or_path}.yaml")
    with open(path, "rb") as f:
        raw = f.read()
    doc = yaml.safe_load(raw)
    units = doc.get("units", {})
    if units.get("activation-energy", "cal/mol") != "cal/mol":
        raise ValueError("only cal/mol activation energies are supported")
    ew = doc["elements"]
    elements = list(ew.keys())
    species: List[Species] = []
    for sp in doc["species"]:
        comp = {k: float(v) for k, v in sp["composition"].items()}
        W = sum(ew[e] * n for e, n in comp.items())
        T = sp["T"]
        species.append(Species(sp["name"], comp, W, float(T[0]), float(T[1]), float(T[2]),
                               [float(x) for x in sp["low"]], [float(x) for x in sp["high"]]))
    names = {s.name: i for i, s in enumerate(species)}
    reactions: List[Reaction] = []
    for r in doc["reactions"]:
        eq = r["equation"]
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
            rev = True
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            rev = False
        else:
            raise ValueError(eq)
        reac, c1 = _parse_side(lhs, names)
        prod, c2 = _parse_side(rhs, names)
        if c1 != c2:
            raise ValueError(f"collider mismatch in {eq}")
        kind = r.get("type", "elementary")
        expect = {None: "elementary", "M": "three-body", "(+M)": "falloff"}[c1]
        if kind != expect:
            raise ValueError(f"type {kind} does not match equation {eq}")
        eff = None
        if kind in ("three-body", "falloff"):
            eff = [1.0] * len(species)
            for sname, v in (r.get("efficiencies") or {}).items():
                eff[names[sname]] = float(v)
        if kind == "falloff":
            hi = r["high"]
            lo = r["low"]
            A, b, Ea = float(hi[0]), float(hi[1]), float(hi[2])
            low = [float(lo[0]), float(lo[1]), float(lo[2]) / R_CAL]
            troe = [float(x) for x in r["troe"]] if "troe" in r else None
        else:
            A, b, Ea = (float(x) for x in r["rate"])
            low = None
            troe = None
        reactions.append(Reaction(eq, reac, prod, rev, kind, A, b, Ea / R_CAL, low, troe, eff,
                                  bool(r.get("duplicate", False))))
    return Mechanism(doc["name"], elements, [float(ew[e]) for e in elements], species, reactions,
                     hashlib.sha256(raw).hexdigest())


def available() -> List[str]:
    return sorted(f[:-5] for f in os.listdir(MECH_DIR) if f.endswith(".yaml"))
