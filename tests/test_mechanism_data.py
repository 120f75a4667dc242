"""Mechanism data checks (CPU): what the data files must satisfy before any kernel uses them.

The reference ships no mechanism data (PelePhysics is an empty submodule), so the authored files are
pinned here by properties a transcription error would break: element balance of every reaction,
C0/C1/C2 continuity of the NASA-7 fits at T_mid, and the species orders the reference's PMF files fix
(Exec/RegTests/PMF/*.dat:1).
"""
import numpy as np
import pytest

from pelec_amd import mechanism as M
from pelec_amd import mechgen

MECHS = ["LiDryer", "drm19", "gri30", "dodecane_lu"]

PMF_ORDER = {
    # Exec/RegTests/PMF/LiDryer_H2_p1_phi0_4000tu0300.dat:1
    "LiDryer": "H2 O2 H2O H O OH HO2 H2O2 N2".split(),
    # Exec/RegTests/PMF/PMF_CH4_1bar_300K_DRM_MixAvg.dat:1
    "drm19": "H2 H O O2 OH H2O HO2 CH2 CH2(S) CH3 CH4 CO CO2 HCO CH2O CH3O C2H4 C2H5 C2H6 N2 AR".split(),
}
PMF_ORDER["gri30"] = (  # GRI-Mech 3.0's published species order (no reference file holds it)
    "H2 H O O2 OH H2O HO2 H2O2 C CH CH2 CH2(S) CH3 CH4 CO CO2 HCO CH2O CH2OH CH3O CH3OH C2H C2H2 "
    "C2H3 C2H4 C2H5 C2H6 HCCO CH2CO HCCOH N NH NH2 NH3 NNH NO NO2 N2O HNO CN HCN H2CN HCNN HCNO "
    "HOCN HNCO NCO N2 AR C3H7 C3H8 CH2CHO CH3CHO").split()
# Exec/RegTests/PMF/PMF_NC12H26_1bar_300K_DodecaneLu.dat:1
PMF_ORDER["dodecane_lu"] = (
    "NC12H26 H O OH HO2 H2 H2O H2O2 O2 CH2 CH2S CH3 CH4 HCO CH2O CH3O CO CO2 C2H2 C2H3 C2H4 C2H5 "
    "C2H6 CH2CHO aC3H5 C3H6 nC3H7 C2H3CHO C4H7 C4H81 pC4H9 C5H9 C5H10 PXC5H11 C6H12 PXC6H13 C7H14 "
    "PXC7H15 C8H16 PXC8H17 C9H18 PXC9H19 C10H20 PXC10H21 PXC12H25 SXC12H25 S3XC12H25 C12H24 "
    "C12H25O2 C12OOH O2C12H24OOH OC12H23OOH N2").split()
SIZES = {"LiDryer": (9, 21), "drm19": (21, 84), "gri30": (53, 325), "dodecane_lu": (53, 229)}
#: continuity of the published fits at T_mid, in cp/R, h/RT and s/R. GRI-Mech 3.0's C3 fits
#: (C3H7 L 9/84, C3H8 L 4/85) jump by up to 9e-5 in cp/R; a mistyped digit breaks this by far more.
#: dodecane_lu's C3+ fits are sums of GRI-Mech 3.0 fits with up to ten C3H8 - C2H6 increments,
#: so they inherit ten times C3H8's jump
CONTINUITY = {"LiDryer": 5e-5, "drm19": 5e-5, "gri30": 1e-4, "dodecane_lu": 1.5e-3}


@pytest.mark.parametrize("name", MECHS)
def test_species_order_and_size(name):
    m = M.load(name)
    assert [s.name for s in m.species] == PMF_ORDER[name]
    assert (m.nsp, m.nrxn) == SIZES[name]


def test_gri30_structure():
    # GRI-Mech 3.0: 29 falloff (Troe or Lindemann) and 12 third-body reactions with efficiencies,
    # 16 irreversible reactions, 3 duplicate pairs; drm19 is a subset of its species and shares
    # their NASA-7 fits
    m = M.load("gri30")
    kinds = [r.kind for r in m.reactions]
    assert kinds.count("falloff") == 29 and kinds.count("three-body") == 12
    assert sum(not r.reversible for r in m.reactions) == 16
    assert sum(r.duplicate for r in m.reactions) == 6
    d = M.load("drm19")
    gs = {s.name: s for s in m.species}
    for s in d.species:
        g = gs[s.name]
        assert (g.low, g.high, g.Tmid, g.composition) == (s.low, s.high, s.Tmid, s.composition)


# standard enthalpies of formation at 298.15 K, kJ/mol (JANAF tables, 4th ed.; the C3+
# hydrocarbons from the NIST Chemistry WebBook gas-phase values)
DHF298 = {"H": 218.0, "O": 249.2, "N": 472.7, "C": 716.7, "H2O": -241.8, "CO": -110.5,
          "CO2": -393.5, "CH4": -74.6, "NO": 90.3, "NO2": 33.1, "N2O": 82.05, "NH3": -45.9,
          "C2H2": 228.2, "C2H4": 52.5, "C2H6": -84.0, "CH3OH": -201.0, "H2O2": -136.1,
          "C3H8": -104.7, "CH3CHO": -166.2}
DHF298_C3PLUS = {"NC12H26": -290.9, "C3H6": 20.0, "C4H81": -0.1, "C6H12": -42.0,
                 "C8H16": -82.9, "C10H20": -123.3, "C12H24": -165.4, "aC3H5": 170.7}


def test_dodecane_lu_structure_and_enthalpies():
    # dodecane_lu (mech/make_dodecane_lu.py): the GRI-Mech 3.0 species it shares keep GRI's fits;
    # the C0-C2 reactions are GRI-Mech 3.0's; the built C3+ fits give the tabulated heats of
    # formation (the a6 shift fixes h(298.15 K); the increments fix the rest of the curve)
    m = M.load("dodecane_lu")
    g = {s.name: s for s in M.load("gri30").species}
    shared = 0
    for s in m.species:
        gs = g.get({"CH2S": "CH2(S)"}.get(s.name, s.name))
        if gs is not None:
            assert (gs.low, gs.high) == (s.low, s.high), s.name
            shared += 1
    assert shared == 24
    for s in m.species:
        if s.name in DHF298_C3PLUS:
            h = _nasa(s.low, 298.15)[1] * 8.31446261815324e-3 * 298.15
            assert abs(h - DHF298_C3PLUS[s.name]) < 2.5, (s.name, h)
    # heat capacity of n-dodecane vapour: 280 J/(mol K) tabulated at 298 K. The C3H8 - C2H6
    # increment is 21.1 J/(mol K) per CH2 at 298 K against Benson's 22.9 for the CH2 group, so ten
    # of them leave the built fit about 6 % low there; at 1000 K the increment matches the group
    # value and the deficit is gone
    cp = _nasa(m.species[m.species_index("NC12H26")].low, 298.15)[0] * 8.31446261815324
    assert 0.90 < cp / 280.0 < 0.97, cp
    # every species is reachable: each appears in some reaction
    used = set()
    for r in m.reactions:
        used |= set(r.reactants) | set(r.products)
    assert used == set(range(m.nsp))


def test_gri30_formation_enthalpies():
    # h(298.15 K) of the low-range fit against tabulated heats of formation: pins the
    # integration constants a6 that continuity alone leaves free (one shift for both ranges)
    for s in M.load("gri30").species:
        if s.name in DHF298:
            h = _nasa(s.low, 298.15)[1] * 8.31446261815324e-3 * 298.15
            assert abs(h - DHF298[s.name]) < 2.0, (s.name, h)


@pytest.mark.parametrize("name", MECHS)
def test_element_balance(name):
    m = M.load(name)
    A = np.array(m.element_matrix())
    for r in m.reactions:
        nu = np.zeros(m.nsp)
        for k, v in r.products.items():
            nu[k] += v
        for k, v in r.reactants.items():
            nu[k] -= v
        assert np.abs(A @ nu).max() < 1e-12, r.equation


def _nasa(a, T):
    cp = a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    s = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return np.array([cp, h, s])


@pytest.mark.parametrize("name", MECHS)
def test_thermo_continuity(name):
    # published fits are continuous at T_mid to ~1e-5; a mistyped digit breaks this by far more
    for s in M.load(name).species:
        d = _nasa(s.low, s.Tmid) - _nasa(s.high, s.Tmid)
        assert np.abs(d).max() < CONTINUITY[name], (s.name, d)


@pytest.mark.parametrize("name", MECHS)
def test_molecular_weights(name):
    m = M.load(name)
    W = dict(zip([s.name for s in m.species], m.weights))
    # Chemkin/PelePhysics atomic weights: H2O = 18.01534, N2 = 28.0134
    assert abs(W["H2O"] - 18.01534) < 1e-9
    assert abs(W["N2"] - 28.0134) < 1e-9


@pytest.mark.parametrize("name", MECHS)
def test_generated_header_matches_data(name, tmp_path):
    m = M.load(name)
    path = mechgen.generate(name, out_dir=str(tmp_path))
    text = open(path).read()
    assert m.source_sha256 in text
    assert f"NSP = {m.nsp};" in text and f"NRXN = {m.nrxn};" in text
    # the committed header is what the data file generates
    committed = open(mechgen.os.path.join(mechgen.GEN_DIR, f"mech_{name}.h")).read()
    assert committed == text
