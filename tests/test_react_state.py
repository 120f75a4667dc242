"""The react_state restatement (oracle/react_state_np.py) pinned by physics it must respect, not by
its own formulas: chemistry conserves mass (sum_k I_R,k = 0) and, with no hydrodynamic forcing,
total energy (I_R,E = 0); the heat release of a short step equals rho c_v dT/dt - R T sum_k
I_R,k / W_k (constant-volume energy balance, h_k - e_k = R T / W_k); initialisation leaves the state
alone; covered cells are untouched with I_R = 0; pc_cmpTemp inverts e(T)."""
import numpy as np
import pytest

from oracle import react_state_np as RS
from oracle.kinetics_np import NpMech
from pelec_amd import inputs, mechanism as M
from pelec_amd.inputs import make_conservative_state as make_state


@pytest.fixture(scope="module")
def drm(oracle_lib):
    m = M.load("drm19")
    return oracle_lib.OracleMech(m), NpMech(m)


def test_react_state_conserves_mass_and_energy(drm):
    om, npm = drm
    W = inputs.WORKLOADS["drm19_64"]
    S, _ = make_state("drm19_64", 48, hydro=False)
    Snew, IR, FC = RS.react_state(om, npm, S, S, np.zeros_like(S), W.dt, rtol=W.rtol, atol=W.atol)
    N = npm.N
    u = np.finfo(float).eps
    # round-off of (rhoY_new - rhoY_old) / dt summed over species, and of the energy difference
    assert (np.abs(IR[:N].sum(0)) <= 64 * u * N * S[RS.URHO] / W.dt).all()       # mass
    assert (np.abs(IR[N]) <= 64 * u * np.abs(S[RS.UEDEN]) / W.dt).all()          # total energy
    assert np.abs(IR[:N]).max() > 1e6 * 64 * u * S[RS.URHO].max() / W.dt         # not vacuous
    np.testing.assert_allclose(Snew[RS.URHO], S[RS.URHO], rtol=1e-13)
    assert (FC > 0).all()


def test_heat_release_energy_balance(drm):
    om, npm = drm
    dt = 1e-8
    S, _ = make_state("drm19_64", 32, hydro=False)
    Snew, IR, _ = RS.react_state(om, npm, S, S, np.zeros_like(S), dt, rtol=1e-12, atol=1e-20)
    N = npm.N
    rho = S[RS.URHO]
    T0, T1 = S[RS.UTEMP], Snew[RS.UTEMP]
    Y = (S[7:7 + N] / rho).T
    cv = (Y * npm.cv_mass(0.5 * (T0 + T1))).sum(-1)
    rhs = rho * cv * (T1 - T0) / dt - RS.R_ERG * T1 * (IR[:N] / npm.W[:, None]).sum(0)
    hr = IR[N + 1]
    big = np.abs(hr) > 1e-3 * np.abs(hr).max()
    assert big.sum() >= 8
    assert (np.abs(hr - rhs)[big] / np.abs(hr[big])).max() < 1e-3


def test_react_init_and_covered_cells(drm):
    om, npm = drm
    W = inputs.WORKLOADS["drm19_64"]
    S, nonrs = make_state("drm19_64", 16)
    mask = np.ones(16, np.int32)
    mask[[2, 9]] = 0
    Snew, IR, FC = RS.react_state(om, npm, None, S, np.zeros_like(S), W.dt, react_init=True,
                                  mask=mask, rtol=W.rtol, atol=W.atol)
    assert np.array_equal(Snew, S)                     # initialisation: state unchanged
    assert (IR[:, [2, 9]] == 0).all() and (FC[[2, 9]] == 0).all()
    Snew, IR, _ = RS.react_state(om, npm, S, S + W.dt * nonrs, nonrs, W.dt, mask=mask,
                                 rtol=W.rtol, atol=W.atol)
    assert np.array_equal(Snew[:, [2, 9]], (S + W.dt * nonrs)[:, [2, 9]])


def test_compute_temp_round_trip(drm):
    _, npm = drm
    S, _ = make_state("drm19_64", 64)
    T0 = S[RS.UTEMP].copy()
    S[RS.UTEMP] = T0 * np.random.default_rng(3).uniform(0.7, 1.3, T0.shape)
    out = RS.compute_temp(npm, S)
    np.testing.assert_allclose(out[RS.UTEMP], T0, rtol=1e-11)
