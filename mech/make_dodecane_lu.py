"""Writes mech/dodecane_lu.yaml: an n-dodecane/air mechanism with the 53 species of PeleC's
dodecane_lu runs, in the order of Exec/RegTests/PMF/PMF_NC12H26_1bar_300K_DodecaneLu.dat:1, and
229 reactions.

The dodecane_lu data files themselves (PelePhysics/Mechanisms/dodecane_lu) are not in the
reference tree, and no file there holds its rate parameters or NASA fits. This mechanism is
therefore assembled here, openly, from two parts:

* C0-C2 and its thermo: every GRI-Mech 3.0 reaction (mech/gri30.yaml) whose species all belong to
  the 53 (127 reactions), with GRI-Mech 3.0's NASA-7 fits for those species. Third-body
  efficiencies of species outside the set are dropped.
* C3-C12: 102 reactions written here from the usual alkane rate rules (H abstraction by H, O, OH,
  HO2, CH3 and O2 by site type; C-C fission; beta scission of alkyl radicals; alkyl + O2
  addition, RO2 -> QOOH isomerisation, second O2 addition and ketohydroperoxide chain branching
  for the low-temperature path; lumped alkene consumption), with parameters of the magnitude
  those rules give. They are this build's choice, not dodecane_lu's values.
* Thermo of the C3+ species is built by group increments from GRI-Mech 3.0 fits: NASA-7
  coefficients are linear in cp/R, h/RT and s/R, so a species is written as a signed sum of GRI
  species with its composition (a CH2 increment C3H8 - C2H6, an H-abstraction increment C2H5 -
  C2H6, ...), and its a6 coefficients are shifted so that h(298.15 K) equals the standard
  enthalpy of formation stated below (kcal/mol, from tabulated values and the same increments).
  Both ranges share T_mid = 1000 K, so the fits stay continuous.

What the build needs from this mechanism is the shape of the configuration: n = 54 unknowns per
cell, 229 reactions, stiff low-temperature branching, and its species order. The kinetics are
parity-unpinned against the real dodecane_lu; the oracle and the GPU consume the same YAML.

Run: python mech/make_dodecane_lu.py  (rewrites mech/dodecane_lu.yaml)
"""
from __future__ import annotations

import os
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from pelec_amd import mechanism as M  # noqa: E402

SPECIES = ("NC12H26 H O OH HO2 H2 H2O H2O2 O2 CH2 CH2S CH3 CH4 HCO CH2O CH3O CO CO2 C2H2 C2H3 "
           "C2H4 C2H5 C2H6 CH2CHO aC3H5 C3H6 nC3H7 C2H3CHO C4H7 C4H81 pC4H9 C5H9 C5H10 PXC5H11 "
           "C6H12 PXC6H13 C7H14 PXC7H15 C8H16 PXC8H17 C9H18 PXC9H19 C10H20 PXC10H21 PXC12H25 "
           "SXC12H25 S3XC12H25 C12H24 C12H25O2 C12OOH O2C12H24OOH OC12H23OOH N2").split()
GRI_NAME = {"CH2S": "CH2(S)"}   # dodecane_lu name -> GRI-Mech 3.0 name
R_CAL = M.R_CAL
T_REF = 298.15

# species built by increments: (composition, {GRI species: coefficient}, dHf298 kcal/mol)
D_CH2 = {"C3H8": 1.0, "C2H6": -1.0}          # one CH2 group
D_ABS = {"C2H5": 1.0, "C2H6": -1.0}          # R-H -> R.


def _add(*terms):
    out = {}
    for coef, d in terms:
        for k, v in d.items():
            out[k] = out.get(k, 0.0) + coef * v
    return {k: v for k, v in out.items() if abs(v) > 1e-12}


def _alkene(n):      # 1-CnH2n = C2H4 + (n-2) CH2
    return _add((1, {"C2H4": 1.0}), (n - 2, D_CH2))


def _alkyl(n):       # n-CnH2n+1 = C2H5 + (n-2) CH2
    return _add((1, {"C2H5": 1.0}), (n - 2, D_CH2))


BUILT = {
    "NC12H26": ({"C": 12, "H": 26}, _add((1, {"C2H6": 1.0}), (10, D_CH2)), -69.9),
    "aC3H5": ({"C": 3, "H": 5}, _add((1, _alkene(3)), (1, D_ABS)), 40.9),
    "C3H6": ({"C": 3, "H": 6}, _alkene(3), 4.9),
    "nC3H7": ({"C": 3, "H": 7}, _alkyl(3), 24.0),
    "C2H3CHO": ({"C": 3, "H": 4, "O": 1}, {"CH3CHO": 1.0, "C2H3": 1.0, "CH3": -1.0}, -17.0),
    "C4H7": ({"C": 4, "H": 7}, _add((1, _alkene(4)), (1, D_ABS)), 32.3),
    "C4H81": ({"C": 4, "H": 8}, _alkene(4), 0.0),
    "pC4H9": ({"C": 4, "H": 9}, _alkyl(4), 19.2),
    "C5H9": ({"C": 5, "H": 9}, _add((1, _alkene(5)), (1, D_ABS)), 27.4),
}
for _n in range(5, 11):
    BUILT[f"C{_n}H{2 * _n}"] = ({"C": _n, "H": 2 * _n}, _alkene(_n), -0.1 - 4.93 * (_n - 4))
    BUILT[f"PXC{_n}H{2 * _n + 1}"] = ({"C": _n, "H": 2 * _n + 1}, _alkyl(_n),
                                      19.2 - 4.93 * (_n - 4))
_RH = BUILT["NC12H26"][1]
BUILT.update({
    # primary (101 kcal/mol) and secondary (98.5) C-H bonds of n-dodecane, H at 52.1
    "PXC12H25": ({"C": 12, "H": 25}, _add((1, _RH), (1, D_ABS)), -69.9 + 101.0 - 52.1),
    "SXC12H25": ({"C": 12, "H": 25}, _add((1, _RH), (1, D_ABS)), -69.9 + 98.5 - 52.1),
    "S3XC12H25": ({"C": 12, "H": 25}, _add((1, _RH), (1, D_ABS)), -69.9 + 98.5 - 52.1),
    "C12H24": ({"C": 12, "H": 24}, _alkene(12), -0.1 - 4.93 * 8),
    # RO2: R + O2 well depth 35.5 kcal/mol; QOOH 15 above RO2; O2QOOH 35 below QOOH + O2;
    # ketohydroperoxide from the group values (R-H -> R-OOH -13, CH2 -> C=O -26.5)
    "C12H25O2": ({"C": 12, "H": 25, "O": 2}, _add((1, _RH), (1, {"HO2": 1.0, "H2": -1.0})), -59.0),
    "C12OOH": ({"C": 12, "H": 25, "O": 2}, _add((1, _RH), (1, {"HO2": 1.0, "H2": -1.0})), -44.0),
    "O2C12H24OOH": ({"C": 12, "H": 25, "O": 4},
                    _add((1, _RH), (2, {"HO2": 1.0}), (-1, {"H2": 1.0}), (-1, {"H": 1.0})), -79.0),
    "OC12H23OOH": ({"C": 12, "H": 24, "O": 3},
                   _add((1, _RH), (2, {"HO2": 1.0}), (-1, {"H2": 1.0}), (-1, {"H": 1.0}),
                        (-1, {"OH": 1.0})), -109.4),
})

# C3-C12 reactions: (equation, rate [A cm-mol-s, b, Ea cal/mol]) or a dict for falloff
EFF = {"H2": 2.0, "H2O": 6.0, "CH4": 2.0, "CO": 1.5, "CO2": 2.0, "C2H6": 3.0}
C3PLUS = [
    # C3: propene, allyl, n-propyl, acrolein
    {"equation": "aC3H5 + H (+M) <=> C3H6 (+M)", "high": [2.0e14, 0.0, 0.0],
     "low": [1.33e60, -12.0, 5967.8], "troe": [0.02, 1096.6, 1096.6, 6859.5]},
    {"equation": "C2H3 + CH3 (+M) <=> C3H6 (+M)", "high": [2.5e13, 0.0, 0.0],
     "low": [4.27e58, -11.94, 9769.8], "troe": [0.175, 1340.6, 60000.0, 10139.8]},
    ("C3H6 + H <=> aC3H5 + H2", [1.7e5, 2.5, 2490.0]),
    ("C3H6 + O <=> aC3H5 + OH", [5.24e11, 0.7, 5884.0]),
    ("C3H6 + OH <=> aC3H5 + H2O", [3.1e6, 2.0, -298.0]),
    ("C3H6 + HO2 <=> aC3H5 + H2O2", [2.7e4, 2.5, 12340.0]),
    ("C3H6 + CH3 <=> aC3H5 + CH4", [2.2, 3.5, 5675.0]),
    ("C3H6 + H <=> C2H4 + CH3", [8.0e21, -2.39, 11180.0]),
    ("C3H6 + O <=> C2H5 + HCO", [1.58e7, 1.76, -1216.0]),
    ("C3H6 + O <=> CH2CHO + CH3", [1.2e8, 1.65, 327.0]),
    ("aC3H5 + O2 <=> C2H3CHO + OH", [2.47e13, -0.44, 23020.0]),
    ("aC3H5 + HO2 <=> C3H6 + O2", [2.66e12, 0.0, 0.0]),
    ("aC3H5 + O <=> C2H3CHO + H", [6.0e13, 0.0, 0.0]),
    ("C2H3CHO + H <=> C2H4 + HCO", [1.08e11, 0.454, 5820.0]),
    ("C2H3CHO + O => C2H3 + OH + CO", [3.0e13, 0.0, 3540.0]),
    ("C2H3CHO + OH => C2H3 + H2O + CO", [3.43e9, 1.18, -447.0]),
    ("nC3H7 <=> CH3 + C2H4", [9.6e13, 0.0, 31000.0]),
    ("nC3H7 <=> H + C3H6", [2.67e15, -0.64, 36820.0]),
    ("nC3H7 + O2 <=> C3H6 + HO2", [3.0e11, 0.0, 3000.0]),
    ("CH2CHO <=> CH3 + CO", [7.8e41, -9.147, 46900.0]),
    # C4: 1-butene, 1-methylallyl, 1-butyl
    {"equation": "aC3H5 + CH3 (+M) <=> C4H81 (+M)", "high": [1.0e14, -0.32, -262.3],
     "low": [3.91e60, -12.81, 6250.0], "troe": [0.104, 1606.0, 60000.0, 6118.4]},
    ("C4H81 + H <=> C4H7 + H2", [6.5e5, 2.54, 6756.0]),
    ("C4H81 + OH <=> C4H7 + H2O", [7.0e2, 2.66, 527.0]),
    ("C4H81 + O <=> C4H7 + OH", [2.6e11, 0.7, 3256.0]),
    ("C4H81 + H <=> C2H4 + C2H5", [1.6e22, -2.39, 11180.0]),
    ("C4H81 + H <=> C3H6 + CH3", [3.2e22, -2.39, 11180.0]),
    ("C4H7 <=> C2H4 + C2H3", [1.0e11, 0.0, 37000.0]),
    ("C4H7 + O2 => C2H3CHO + CH3O", [1.0e9, 0.0, 0.0]),
    ("pC4H9 <=> C2H5 + C2H4", [1.06e13, 0.0, 28000.0]),
    ("pC4H9 <=> C4H81 + H", [1.26e13, 0.0, 38000.0]),
    ("pC4H9 + O2 <=> C4H81 + HO2", [8.4e11, 0.0, 3875.0]),
    # C5: pentenyl
    ("C5H10 + H <=> C5H9 + H2", [6.5e5, 2.54, 6756.0]),
    ("C5H10 + OH <=> C5H9 + H2O", [7.0e2, 2.66, 527.0]),
    ("C5H9 <=> C2H4 + aC3H5", [2.5e13, 0.0, 30000.0]),
]
# C5-C10 cascade: per carbon number, alkyl beta scission, alkene allylic C-C fission, H addition
# with beta scission, OH addition with decomposition
_ALKYL = {2: "C2H5", 3: "nC3H7", 4: "pC4H9"}
for _n in range(5, 11):
    _ALKYL[_n] = f"PXC{_n}H{2 * _n + 1}"
for _n in range(5, 11):
    alk, ene = _ALKYL[_n], f"C{_n}H{2 * _n}"
    C3PLUS += [
        (f"{alk} <=> C2H4 + {_ALKYL[_n - 2]}", [1.0e13, 0.0, 28300.0]),
        (f"{ene} <=> aC3H5 + {_ALKYL[_n - 3]}", [2.5e16, 0.0, 71000.0]),
        (f"{ene} + H <=> C2H4 + {_ALKYL[_n - 2]}", [8.0e21, -2.39, 11180.0]),
    ]
    if _n >= 6:
        C3PLUS.append((f"{ene} + OH => CH2O + {_ALKYL[_n - 1]}", [1.0e12, 0.0, -1000.0]))
# n-dodecane
C3PLUS += [
    ("NC12H26 <=> 2 PXC6H13", [4.0e16, 0.0, 81000.0]),
    ("NC12H26 <=> PXC5H11 + PXC7H15", [8.0e16, 0.0, 81000.0]),
    ("NC12H26 <=> pC4H9 + PXC8H17", [8.0e16, 0.0, 81000.0]),
    # H abstraction: 6 primary H (PX), 4 secondary H at C2 (SX), 16 secondary H at C3-C6 (S3X)
    ("NC12H26 + H <=> PXC12H25 + H2", [1.3e6, 2.54, 6756.0]),
    ("NC12H26 + H <=> SXC12H25 + H2", [2.6e6, 2.4, 4471.0]),
    ("NC12H26 + H <=> S3XC12H25 + H2", [1.04e7, 2.4, 4471.0]),
    ("NC12H26 + O <=> PXC12H25 + OH", [1.9e6, 2.68, 3716.0]),
    ("NC12H26 + O <=> SXC12H25 + OH", [9.54e4, 2.71, 2106.0]),
    ("NC12H26 + O <=> S3XC12H25 + OH", [3.8e5, 2.71, 2106.0]),
    ("NC12H26 + OH <=> PXC12H25 + H2O", [2.7e7, 1.8, 1000.0]),
    ("NC12H26 + OH <=> SXC12H25 + H2O", [9.4e7, 1.61, -35.0]),
    ("NC12H26 + OH <=> S3XC12H25 + H2O", [3.76e8, 1.61, -35.0]),
    ("NC12H26 + HO2 <=> PXC12H25 + H2O2", [4.08e1, 3.59, 17160.0]),
    ("NC12H26 + HO2 <=> SXC12H25 + H2O2", [1.26e2, 3.37, 13720.0]),
    ("NC12H26 + HO2 <=> S3XC12H25 + H2O2", [5.04e2, 3.37, 13720.0]),
    ("NC12H26 + CH3 <=> PXC12H25 + CH4", [1.81, 3.65, 7153.0]),
    ("NC12H26 + CH3 <=> SXC12H25 + CH4", [6.0, 3.46, 5480.0]),
    ("NC12H26 + CH3 <=> S3XC12H25 + CH4", [2.4e1, 3.46, 5480.0]),
    ("NC12H26 + O2 <=> PXC12H25 + HO2", [6.0e13, 0.0, 52800.0]),
    ("NC12H26 + O2 <=> SXC12H25 + HO2", [4.0e13, 0.0, 50150.0]),
    ("NC12H26 + O2 <=> S3XC12H25 + HO2", [1.6e14, 0.0, 50150.0]),
    # dodecyl isomerisation and beta scission
    ("PXC12H25 <=> S3XC12H25", [3.67e12, -0.6, 15300.0]),
    ("SXC12H25 <=> S3XC12H25", [1.3e12, 0.0, 18000.0]),
    ("PXC12H25 <=> C2H4 + PXC10H21", [1.0e13, 0.0, 28300.0]),
    ("SXC12H25 <=> C3H6 + PXC9H19", [1.6e13, 0.0, 28300.0]),
    ("SXC12H25 <=> C12H24 + H", [3.0e13, 0.0, 37000.0]),
    ("S3XC12H25 <=> C4H81 + PXC8H17", [2.5e12, 0.0, 28300.0]),
    ("S3XC12H25 <=> C5H10 + PXC7H15", [2.5e12, 0.0, 28300.0]),
    ("S3XC12H25 <=> C6H12 + PXC6H13", [2.5e12, 0.0, 28300.0]),
    ("S3XC12H25 <=> C7H14 + PXC5H11", [2.5e12, 0.0, 28300.0]),
    ("S3XC12H25 <=> C8H16 + pC4H9", [2.5e12, 0.0, 28300.0]),
    ("S3XC12H25 <=> C9H18 + nC3H7", [2.5e12, 0.0, 28300.0]),
    ("S3XC12H25 <=> C10H20 + C2H5", [2.5e12, 0.0, 28300.0]),
    ("C12H24 <=> aC3H5 + PXC9H19", [2.5e16, 0.0, 71000.0]),
    ("C12H24 + OH => CH2O + C2H4 + PXC9H19", [1.0e12, 0.0, -1000.0]),
    # low-temperature path
    ("PXC12H25 + O2 <=> C12H25O2", [5.0e13, 0.0, 0.0]),
    ("SXC12H25 + O2 <=> C12H25O2", [5.0e13, 0.0, 0.0]),
    ("S3XC12H25 + O2 <=> C12H25O2", [5.0e13, 0.0, 0.0]),
    ("C12H25O2 <=> C12OOH", [2.5e10, 0.0, 20450.0]),
    ("C12H25O2 <=> C12H24 + HO2", [5.0e12, 0.0, 29000.0]),
    ("C12OOH <=> C12H24 + HO2", [1.0e12, 0.0, 18000.0]),
    ("C12OOH => C4H81 + C8H16 + HO2", [3.0e13, 0.0, 28000.0]),
    ("C12OOH + O2 <=> O2C12H24OOH", [4.6e10, 0.0, 0.0]),
    ("O2C12H24OOH => OC12H23OOH + OH", [2.0e11, 0.0, 17000.0]),
    ("OC12H23OOH => OH + CH2CHO + CH2O + C9H18", [1.0e16, 0.0, 39000.0]),
]


def _fmt(x):
    return float(f"{x:.10g}")


def build():
    gri = M.load("gri30")
    gidx = {s.name: i for i, s in enumerate(gri.species)}
    with open(os.path.join(HERE, "gri30.yaml")) as f:
        gdoc = yaml.safe_load(f)
    gsp = {s["name"]: s for s in gdoc["species"]}
    species = []
    for name in SPECIES:
        gname = GRI_NAME.get(name, name)
        if name not in BUILT:
            s = dict(gsp[gname])
            s["name"] = name
            species.append(s)
            continue
        comp, combo, dhf = BUILT[name]
        low = [0.0] * 7
        high = [0.0] * 7
        for g, c in combo.items():
            sg = gri.species[gidx[g]]
            for i in range(7):
                low[i] += c * sg.low[i]
                high[i] += c * sg.high[i]
        # shift a6 so that h(298.15 K) is the target enthalpy of formation
        a = low
        hRT = (a[0] + T_REF * (a[1] / 2 + T_REF * (a[2] / 3 + T_REF * (a[3] / 4 + T_REF * a[4] / 5)))
               + a[5] / T_REF)
        shift = dhf * 1000.0 / R_CAL - hRT * T_REF
        low[5] += shift
        high[5] += shift
        species.append({"name": name, "composition": comp, "T": [200.0, 1000.0, 3500.0],
                        "low": [_fmt(x) for x in low], "high": [_fmt(x) for x in high]})
    keep = set(SPECIES)
    to_local = {v: k for k, v in GRI_NAME.items()}
    reactions = []
    for r in gdoc["reactions"]:
        eq = r["equation"]
        toks = [t for t in eq.replace("(+M)", " ").replace("<=>", " ").replace("=>", " ")
                .replace("+", " ").split()]
        names = [t for t in toks if not t.replace(".", "").isdigit() and t != "M"]
        local = [to_local.get(t, t) for t in names]
        if not all(t in keep for t in local):
            continue
        r = dict(r)
        for g, l in to_local.items():
            r["equation"] = r["equation"].replace(g, l)
        if "efficiencies" in r:
            r["efficiencies"] = {to_local.get(k, k): v for k, v in r["efficiencies"].items()
                                 if to_local.get(k, k) in keep}
        reactions.append(r)
    n_gri = len(reactions)
    for item in C3PLUS:
        if isinstance(item, dict):
            d = {"equation": item["equation"], "type": "falloff", "high": item["high"],
                 "low": item["low"], "troe": item["troe"], "efficiencies": dict(EFF)}
        else:
            d = {"equation": item[0], "rate": item[1]}
        reactions.append(d)
    return species, reactions, n_gri


def main():
    species, reactions, n_gri = build()
    out = os.path.join(HERE, "dodecane_lu.yaml")
    with open(out, "w") as f:
        f.write("# n-dodecane/air, 53 species in the order of PeleC's dodecane_lu runs\n"
                "# (Exec/RegTests/PMF/PMF_NC12H26_1bar_300K_DodecaneLu.dat:1), 229 reactions.\n"
                f"# Written by mech/make_dodecane_lu.py: {n_gri} GRI-Mech 3.0 reactions (C0-C2) and\n"
                f"# {len(reactions) - n_gri} C3-C12 reactions from alkane rate rules; C3+ thermo by group\n"
                "# increments of GRI-Mech 3.0 fits. Units: A in cm-mol-s, Ea in cal/mol.\n")
        f.write("name: dodecane_lu\n")
        f.write("units: {length: cm, quantity: mol, activation-energy: cal/mol}\n")
        f.write("elements: {O: 15.9994, H: 1.00797, C: 12.01115, N: 14.0067}\n")
        f.write("species:\n")
        for s in species:
            f.write(f"- name: {s['name']}\n")
            f.write(f"  composition: {yaml.safe_dump(s['composition'], default_flow_style=True).strip()}\n")
            f.write(f"  T: {list(map(float, s['T']))}\n")
            f.write(f"  low: {[float(x) for x in s['low']]}\n")
            f.write(f"  high: {[float(x) for x in s['high']]}\n")
        f.write("reactions:\n")
        for r in reactions:
            f.write("- " + yaml.safe_dump(r, default_flow_style=True, sort_keys=False, width=1000).strip() + "\n")
    print(f"{out}: {len(species)} species, {len(reactions)} reactions ({n_gri} from GRI-Mech 3.0)")


if __name__ == "__main__":
    main()
