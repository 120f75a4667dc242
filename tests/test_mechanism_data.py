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

MECHS = ["LiDryer", "drm19"]

PMF_ORDER = {
    # Exec/RegTests/PMF/LiDryer_H2_p1_phi0_4000tu0300.dat:1
    "LiDryer": "H2 O2 H2O H O OH HO2 H2O2 N2".split(),
    # Exec/RegTests/PMF/PMF_CH4_1bar_300K_DRM_MixAvg.dat:1
    "drm19": "H2 H O O2 OH H2O HO2 CH2 CH2(S) CH3 CH4 CO CO2 HCO CH2O CH3O C2H4 C2H5 C2H6 N2 AR".split(),
}
SIZES = {"LiDryer": (9, 21), "drm19": (21, 84)}


@pytest.mark.parametrize("name", MECHS)
def test_species_order_and_size(name):
    m = M.load(name)
    assert [s.name for s in m.species] == PMF_ORDER[name]
    assert (m.nsp, m.nrxn) == SIZES[name]


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
        assert np.abs(d).max() < 5e-5, (s.name, d)


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
