"""Pins the CPU FP64 oracle (oracle/oracle.cpp) before anything is compared against it.

Each oracle function is checked against something it was not derived from:
  * the reference's golden Cantera ignition trajectories (Exec/RegTests/zeroD/cantera_soln,
    Exec/RegTests/EB-C3/cantera_soln) -- kinetics + thermo + integrator end to end (LiDryer);
  * scipy Radau solutions at rtol 1e-12 (tests/golden/radau_*.npz) -- the BDF integrator;
  * the independent numpy restatement (oracle/kinetics_np.py) -- rates and RHS;
  * torch autograd through a torch restatement, and central differences -- the analytic Jacobian;
  * closed forms: falloff limits (Pr -> 0 gives k0[M], Pr -> inf gives k_inf), detailed balance
    at equilibrium, T(e(T)) = T, mass/element/energy conservation, LU against numpy.
"""
import os

import numpy as np
import pytest
import torch

from oracle.kinetics_np import NpMech
from pelec_amd import inputs, mechanism as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def mechs(oracle_lib):
    out = {}
    for name in ("LiDryer", "drm19", "gri30", "dodecane_lu"):
        m = M.load(name)
        out[name] = (m, NpMech(m), oracle_lib.OracleMech(m))
    return out


def _random_state(npm, rng, T=1500.0, p=3 * 1013250.0):
    N = npm.N
    Y = rng.uniform(0.01, 1.0, N)
    Y /= Y.sum()
    rho, e = npm.mixture(p, T, Y)
    return Y, rho, e


# ---------------------------------------------------------------- golden trajectories
def _cantera_case(mechs, fname, T0, skip):
    m, npm, om = mechs["LiDryer"]
    ref = np.loadtxt(os.path.join(GOLDEN, fname), skiprows=skip)
    N = m.nsp
    Y = np.zeros(N)
    Y[m.species_index("H2")], Y[m.species_index("O2")], Y[m.species_index("N2")] = 0.06, 0.5, 0.44
    rho, e = npm.mixture(1013250.0, T0, Y)
    rows = []
    for t, Tref, yh2, yo2, yoh in ref:
        out = om.react((rho * Y)[:, None], np.array([T0]), np.array([rho * e]), t, rtol=1e-10,
                       atol=1e-16)
        assert out["status"][0] == 0
        Yo = out["rY"][:, 0] / out["rY"][:, 0].sum()
        rows.append((t, out["T"][0] - Tref, Yo[m.species_index("H2")] - yh2,
                     Yo[m.species_index("O2")] - yo2, Yo[m.species_index("OH")] - yoh, Tref))
    return np.array(rows)


def test_cantera_eb_c3_trajectory(mechs):
    # T0 = 1500 K, p = 1 atm, Y(H2, O2, N2) = (0.06, 0.5, 0.44), Exec/RegTests/EB-C3/prob_parm.H:11-14
    d = _cantera_case(mechs, "EB-C3_cantera_soln.txt", 1500.0, 0)
    assert np.abs(d[:, 1]).max() < 0.05          # K, through ignition (values printed to 1e-6 K)
    assert np.abs(d[-1, 1]) < 0.01                # equilibrium end point 3238.304869 K
    assert np.abs(d[:, 2:5]).max() < 2e-5         # mass fractions printed to 1e-6


def test_cantera_zerod_trajectory(mechs):
    # T0 = 1000 K (Exec/RegTests/zeroD/zerod-1.inp:62-66); ignition near 0.135 ms
    d = _cantera_case(mechs, "zeroD_cantera_soln.txt", 1000.0, 1)
    dTdt = np.abs(np.gradient(d[:, 5], d[:, 0]))
    steep = dTdt > 1e7   # K/s: on the ignition front compare as an ignition-time shift
    assert np.abs(d[~steep, 1]).max() < 0.05
    assert (np.abs(d[steep, 1]) / dTdt[steep]).max() < 5e-8   # s, vs ignition at ~1.35e-4 s
    assert np.abs(d[-1, 1]) < 0.01                # 3223.856190 K
    assert np.abs(d[~steep, 2:5]).max() < 2e-5


@pytest.mark.parametrize("wl", ["lidryer_32", "drm19_64"])
def test_bdf_against_radau(mechs, wl):
    g = np.load(os.path.join(GOLDEN, f"radau_{wl}.npz"))
    W = inputs.WORKLOADS[wl]
    m, npm, om = mechs[W.mech]
    assert str(g["mech_sha256"]) == m.source_sha256
    for rtol, tolT in ((1e-8, 2e-4), (1e-10, 1e-5)):
        out = om.react(g["rY0"], g["T0"], g["rE0"], float(g["dt"]), rtol=rtol, atol=1e-20)
        assert (out["status"] == 0).all()
        dT = np.abs(out["T"] - g["T1"])
        assert (dT < tolT * np.maximum(1.0, np.abs(g["T1"] - g["T0"]))).all(), (rtol, dT)
        rho = g["rY1"].sum(0)
        dY = np.abs(out["rY"] - g["rY1"]) / rho
        assert dY.max() < 100 * rtol


# ---------------------------------------------------------------- rates and RHS
@pytest.mark.parametrize("name", ["LiDryer", "drm19", "gri30", "dodecane_lu"])
def test_rhs_matches_numpy_restatement(mechs, name):
    m, npm, om = mechs[name]
    rng = np.random.default_rng(1)
    for T in (700.0, 999.0, 1001.0, 1800.0, 2600.0):
        Y, rho, e = _random_state(npm, rng, T)
        src = rng.normal(0, 1e-4, m.nsp)
        y = np.concatenate([rho * Y, [T * 1.01]])
        f_c, T_c = om.rhs(2e-6, y, rho * e, 1e7, src)
        f_n, T_n = npm.rhs(2e-6, y[None], rho * e, 1e7, src[None])
        assert abs(T_c - T_n[0]) < 1e-9 * T
        assert np.abs(f_c - f_n[0]).max() <= 1e-12 * np.abs(f_n).max()


def test_falloff_limits(oracle_lib):
    # Lindemann form (Troe factor removed): Pr -> 0 gives k0 [M], Pr -> inf gives k_inf
    import copy
    m = M.load("drm19")
    r = copy.deepcopy([x for x in m.reactions if x.kind == "falloff"][1])  # H + CH3 (+M) <=> CH4 (+M)
    r.troe = None
    m.reactions = [r]
    om = oracle_lib.OracleMech(m)
    T = 1500.0
    kinf = r.A * T ** r.b * np.exp(-r.Ea_R / T)
    k0 = r.low[0] * T ** r.low[1] * np.exp(-r.low[2] / T)
    C = np.zeros(m.nsp)
    iH, iCH3 = m.species_index("H"), m.species_index("CH3")
    for scale, expect in ((1e-14, "low"), (1e14, "high")):
        C[:] = 0.0
        C[m.species_index("N2")] = scale          # efficiency-1 collider
        C[iH], C[iCH3] = 1e-12 * scale, 1e-12 * scale
        q = om.progress(C, T)[0]                  # no CH4: reverse rate is zero
        fwd = q / (C[iH] * C[iCH3])
        ratio = fwd / (k0 * C.sum()) if expect == "low" else fwd / kinf
        assert abs(ratio - 1.0) < 1e-6, (expect, ratio)


def test_troe_factor_against_numpy(oracle_lib):
    # the Troe (Gilbert et al. 1983) blending over 12 decades of Pr, oracle vs numpy restatement
    m = M.load("drm19")
    npm = NpMech(m)
    om = oracle_lib.OracleMech(m)
    iN2 = m.species_index("N2")
    for T in (800.0, 1500.0, 2500.0):
        for scale in 10.0 ** np.arange(-12, 7, 2):
            C = np.full(m.nsp, 1e-7 * scale)
            C[iN2] = scale
            q_c = om.progress(C, T)
            q_n = npm.rates_of_progress(C[None, :], np.array([T]))[0]
            fall = [j for j, r in enumerate(m.reactions) if r.kind == "falloff"]
            np.testing.assert_allclose(q_c[fall], q_n[fall], rtol=1e-11)


@pytest.mark.parametrize("name", ["LiDryer", "drm19", "gri30", "dodecane_lu"])
def test_conservation_of_rates(mechs, name):
    m, npm, om = mechs[name]
    rng = np.random.default_rng(2)
    A = np.array(m.element_matrix())
    for T in (900.0, 1500.0, 2400.0):
        Y, rho, e = _random_state(npm, rng, T)
        C = rho * Y / npm.W
        w = om.wdot(C, T)
        scale = np.abs(w * npm.W).max()
        assert abs((w * npm.W).sum()) <= 1e-12 * scale          # mass
        assert np.abs(A @ w).max() <= 1e-12 * np.abs(w).max() * 10  # elements


def test_detailed_balance_at_equilibrium(mechs):
    # integrate the EB-C3 mixture to constant-volume equilibrium; every reaction's net rate of
    # progress must then vanish relative to its forward rate
    m, npm, om = mechs["LiDryer"]
    N = m.nsp
    Y = np.zeros(N)
    Y[0], Y[1], Y[8] = 0.06, 0.5, 0.44
    rho, e = npm.mixture(1013250.0, 1500.0, Y)
    out = om.react((rho * Y)[:, None], np.array([1500.0]), np.array([rho * e]), 2e-3, rtol=1e-12,
                   atol=1e-20)
    C = out["rY"][:, 0] / npm.W
    T = out["T"][0]
    q = om.progress(C, T)
    # forward-only magnitude from the numpy restatement
    Tc = np.array([T])
    kf = npm.A * np.exp(npm.b * np.log(Tc) - npm.EaR / Tc)
    fwd = np.abs(kf * np.prod(C[None, :] ** npm.nuf, axis=-1))
    assert (np.abs(q) <= 1e-5 * fwd + 1e-30).all(), np.abs(q) / fwd


def test_detailed_balance_at_equilibrium_gri30(oracle_lib):
    # GRI-Mech 3.0 with its 16 irreversible reactions removed (they drain CH2, CH3CHO and others
    # even at equilibrium, so the remaining network cannot balance exactly): a stoichiometric
    # CH4/air mixture at 2400 K integrated to constant-volume equilibrium has every reaction's
    # net rate of progress vanishing relative to its forward rate. Exercises the equilibrium
    # constants of all 53 species through 309 reactions.
    import copy
    m = copy.deepcopy(M.load("gri30"))
    m.reactions = [r for r in m.reactions if r.reversible]
    npm, om = NpMech(m), oracle_lib.OracleMech(m)
    N = m.nsp
    X = np.zeros(N)
    X[m.species_index("CH4")], X[m.species_index("O2")], X[m.species_index("N2")] = 0.5, 1.0, 3.76
    Y = X * npm.W / (X * npm.W).sum()
    rho, e = npm.mixture(1013250.0, 2400.0, Y)
    out = om.react((rho * Y)[:, None], np.array([2400.0]), np.array([rho * e]), 1.0, rtol=1e-12,
                   atol=1e-20)
    assert out["status"][0] == 0
    C = np.maximum(out["rY"][:, 0], 0.0) / npm.W
    T = out["T"][0]
    q = om.progress(C, T)
    Tc = np.array([T])
    kf = npm.A * np.exp(npm.b * np.log(Tc) - npm.EaR / Tc)
    fwd = np.abs(kf * np.prod(C[None, :] ** npm.nuf, axis=-1))
    big = fwd > 1e-12 * fwd.max()   # reactions with a measurable forward flux
    assert big.sum() > 100
    assert (np.abs(q[big]) <= 1e-5 * fwd[big]).all(), (np.abs(q[big]) / fwd[big]).max()


def test_detailed_balance_at_equilibrium_dodecane_lu(oracle_lib):
    # dodecane_lu with its irreversible reactions removed: a stoichiometric n-dodecane/air mixture
    # at 2400 K integrated to constant-volume equilibrium has every reaction with a measurable
    # forward flux in balance. At equilibrium the C3+ species are too scarce for their reactions
    # to carry flux, so this pins the C0-C2 network and the C3 reactions that still do; the
    # C3+ equilibrium constants are pinned against numpy by test_rhs_matches_numpy_restatement
    import copy
    m = copy.deepcopy(M.load("dodecane_lu"))
    m.reactions = [r for r in m.reactions if r.reversible]
    npm, om = NpMech(m), oracle_lib.OracleMech(m)
    N = m.nsp
    X = np.zeros(N)
    X[m.species_index("NC12H26")], X[m.species_index("O2")] = 1.0 / 18.5, 1.0
    X[m.species_index("N2")] = 3.76
    Y = X * npm.W / (X * npm.W).sum()
    rho, e = npm.mixture(1013250.0, 2400.0, Y)
    out = om.react((rho * Y)[:, None], np.array([2400.0]), np.array([rho * e]), 1.0, rtol=1e-12,
                   atol=1e-20)
    assert out["status"][0] == 0
    C = np.maximum(out["rY"][:, 0], 0.0) / npm.W
    T = out["T"][0]
    q = om.progress(C, T)
    Tc = np.array([T])
    kf = npm.A * np.exp(npm.b * np.log(Tc) - npm.EaR / Tc)
    fwd = np.abs(kf * np.prod(C[None, :] ** npm.nuf, axis=-1))
    big = fwd > 1e-12 * fwd.max()
    assert big.sum() > 40
    assert (np.abs(q[big]) <= 1e-5 * fwd[big]).all(), (np.abs(q[big]) / fwd[big]).max()


# ---------------------------------------------------------------- thermo / EOS
@pytest.mark.parametrize("name", ["LiDryer", "drm19", "gri30", "dodecane_lu"])
def test_T_from_e_round_trip(mechs, name):
    m, npm, om = mechs[name]
    rng = np.random.default_rng(3)
    for T in (350.0, 999.9, 1000.1, 1700.0, 3300.0):
        Y, rho, e = _random_state(npm, rng, T)
        Tr, it = om.T_from_e(e, Y, T * 1.3)
        assert abs(Tr - T) <= 1e-11 * T and it < 20
        # monotone: cv > 0
        assert (npm.cv_mass(T) > 0).all()


def test_thermo_against_numpy(mechs):
    m, npm, om = mechs["drm19"]
    for T in (300.0, 1000.0, 2000.0):
        cp, h, s = om.thermo(T)
        np.testing.assert_allclose(cp, npm.cp_R(np.array(T)), rtol=1e-14)
        np.testing.assert_allclose(h, npm.h_RT(np.array(T)), rtol=1e-13)
        np.testing.assert_allclose(s, npm.s_R(np.array(T)), rtol=1e-13)


# ---------------------------------------------------------------- Jacobian
def _torch_rhs(npm, y, rhoe0, rhoesrc, src, t):
    """Torch restatement of the RHS; T enters through one implicit-function Newton step."""
    N = npm.N
    W = torch.tensor(npm.W)
    rY = y[:N]
    rho = rY.sum()
    Y = rY / rho
    e = (rhoe0 + t * rhoesrc) / rho
    T0 = float(npm.T_from_e(np.array(e.item()), Y.detach().numpy(), np.array(1500.0)))
    Tt = torch.tensor(T0, dtype=torch.float64)

    def thermo(T):
        lo = T < torch.tensor(npm.Tmid)
        a = torch.where(lo[:, None], torch.tensor(npm.low), torch.tensor(npm.high))
        cp = a[:, 0] + T * (a[:, 1] + T * (a[:, 2] + T * (a[:, 3] + T * a[:, 4])))
        hRT = a[:, 0] + T * (a[:, 1] / 2 + T * (a[:, 2] / 3 + T * (a[:, 3] / 4 + T * a[:, 4] / 5))) + a[:, 5] / T
        sR = a[:, 0] * torch.log(T) + T * (a[:, 1] + T * (a[:, 2] / 2 + T * (a[:, 3] / 3 + T * a[:, 4] / 4))) + a[:, 6]
        return cp, hRT, sR

    cp, hRT, _ = thermo(Tt)
    ek = (hRT - 1) * 8.31446261815324e7 * Tt / W
    cvk = (cp - 1) * 8.31446261815324e7 / W
    T = Tt - ((Y * ek).sum() - e) / (Y * cvk).sum()   # exact first derivatives of T(e, Y)
    cp, hRT, sR = thermo(T)
    C = rY / W
    nu = torch.tensor(npm.nu)
    kf = torch.tensor(npm.A) * torch.exp(torch.tensor(npm.b) * torch.log(T) - torch.tensor(npm.EaR) / T)
    dG = (hRT - sR) @ nu.T
    Kc = torch.exp(-dG) * (1013250.0 / (8.31446261815324e7 * T)) ** nu.sum(1)
    Mc = torch.tensor(npm.eff) @ C
    kf_list = []
    for j, kind in enumerate(npm.kind):
        if kind == "three-body":
            kf_list.append(kf[j] * Mc[j])
        elif kind == "falloff":
            r = npm.mech.reactions[j]
            k0 = r.low[0] * torch.exp(r.low[1] * torch.log(T) - r.low[2] / T)
            Pr = k0 * Mc[j] / kf[j]
            F = torch.tensor(1.0, dtype=torch.float64)
            if r.troe:
                a_, T3, T1 = r.troe[:3]
                Fc = (1 - a_) * torch.exp(-T / T3) + a_ * torch.exp(-T / T1)
                if len(r.troe) > 3:
                    Fc = Fc + torch.exp(-r.troe[3] / T)
                lFc = torch.log10(Fc)
                c = -0.4 - 0.67 * lFc
                n = 0.75 - 1.27 * lFc
                f1 = (torch.log10(Pr) + c) / (n - 0.14 * (torch.log10(Pr) + c))
                F = 10.0 ** (lFc / (1 + f1 * f1))
            kf_list.append(kf[j] * Pr / (1 + Pr) * F)
        else:
            kf_list.append(kf[j])
    kfe = torch.stack(kf_list)
    kr = torch.where(torch.tensor(npm.rev), kfe / Kc, torch.zeros_like(kfe))
    q = kfe * torch.prod(C[None, :] ** torch.tensor(npm.nuf), dim=1) - kr * torch.prod(C[None, :] ** torch.tensor(npm.nur), dim=1)
    wdot = q @ nu
    f = W * wdot + src
    ek = (hRT - 1) * 8.31446261815324e7 * T / W
    cvk = (cp - 1) * 8.31446261815324e7 / W
    fT = (rhoesrc - (ek * f).sum()) / (rY * cvk).sum()
    return torch.cat([f, fT[None]])


@pytest.mark.parametrize("name", ["LiDryer", "drm19", "gri30", "dodecane_lu"])
def test_jacobian_against_autograd_and_fd(mechs, name):
    m, npm, om = mechs[name]
    rng = np.random.default_rng(4)
    for T in (1100.0, 1900.0):
        Y, rho, e = _random_state(npm, rng, T)
        src = rng.normal(0, 1e-4, m.nsp)
        y = np.concatenate([rho * Y, [T]])
        J = om.jac(3e-6, y, rho * e, 1e8, src)
        yt = torch.tensor(y, dtype=torch.float64)
        Ja = torch.autograd.functional.jacobian(
            lambda v: _torch_rhs(npm, v, rho * e, 1e8, torch.tensor(src), 3e-6), yt).numpy()
        scale = np.abs(Ja).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J - Ja) / scale).max() < 1e-9
        Jfd = np.zeros_like(J)
        for k in range(m.nsp + 1):
            h = 1e-6 * max(abs(y[k]), 1e-14)
            yp, ym = y.copy(), y.copy()
            yp[k] += h
            ym[k] -= h
            Jfd[:, k] = (om.rhs(3e-6, yp, rho * e, 1e8, src)[0] - om.rhs(3e-6, ym, rho * e, 1e8, src)[0]) / (2 * h)
        assert (np.abs(J - Jfd) / scale).max() < 1e-5


def test_lu_against_numpy(oracle_lib):
    rng = np.random.default_rng(5)
    for n in (1, 2, 10, 22, 54):
        A = rng.normal(size=(n, n)) + np.eye(n) * 0.1
        b = rng.normal(size=n)
        a = A.copy()
        piv = np.zeros(n, dtype=np.int32)
        L = oracle_lib.lib()
        assert L.orc_getrf(n, oracle_lib._d(a), oracle_lib._i(piv)) == 0
        x = b.copy()
        L.orc_getrs(n, oracle_lib._d(a), oracle_lib._i(piv), oracle_lib._d(x))
        np.testing.assert_allclose(A @ x, b, atol=1e-10 * np.abs(b).max() * np.linalg.cond(A))


# ---------------------------------------------------------------- reactor invariants
@pytest.mark.parametrize("wl", ["lidryer_32", "drm19_64"])
def test_react_invariants(mechs, wl):
    W = inputs.WORKLOADS[wl]
    m, npm, om = mechs[W.mech]
    d = inputs.make_inputs(wl, 256)
    rng = np.random.default_rng(6)
    rEsrc = rng.normal(0, 1e9, 256)
    mask = (rng.uniform(size=256) > 0.2).astype(np.int32)
    out = om.react(d["rY"], d["T"], d["rE"], W.dt, rEsrc=rEsrc, mask=mask, rtol=W.rtol, atol=W.atol)
    act = mask == 1
    assert (out["status"] == 0).all()
    # masked cells untouched, FC = 0
    assert np.array_equal(out["rY"][:, ~act], d["rY"][:, ~act])
    assert np.array_equal(out["T"][~act], d["T"][~act]) and (out["FC"][~act] == 0).all()
    # mass per cell (zero species source) and energy rho e(dt) = rho e0 + dt src (React.cpp:292)
    rho0, rho1 = d["rY"].sum(0), out["rY"].sum(0)
    assert np.abs(rho1 - rho0).max() <= 1e-13 * rho0.max()
    np.testing.assert_array_equal(out["rE"][act], d["rE"][act] + W.dt * rEsrc[act])
    # elements
    A = np.array(m.element_matrix())
    el0 = A @ (d["rY"] / npm.W[:, None])
    el1 = A @ (out["rY"] / npm.W[:, None])
    assert np.abs(el1 - el0)[:, act].max() <= 1e-12 * np.abs(el0).max()
    # returned T is consistent with the returned energy
    Y1 = out["rY"] / rho1
    e1 = (Y1.T * npm.e_mass(out["T"])).sum(-1)
    assert np.abs(e1 - out["rE"] / rho1)[act].max() <= 1e-11 * np.abs(out["rE"] / rho1).max()
    assert (out["FC"][act] > 0).all()


def test_react_bad_input_reports_status(mechs):
    m, npm, om = mechs["LiDryer"]
    d = inputs.make_inputs("lidryer_32", 4)
    d["T"][1] = -5.0
    d["rY"][:, 2] = 0.0
    out = om.react(d["rY"], d["T"], d["rE"], 1e-6, rtol=1e-10, atol=1e-10)
    assert list(out["status"]) == [0, 4, 4, 0]
    assert out["nfail"] == 2


# ------------------------------------------------------------------ ReactorRK64 restatement
# The published RK4(3)6[2R+]C tableau (Kennedy, Carpenter & Lewis 2000) in register form; the
# order conditions below are the mathematics that fixes it (a transcription slip in any of the 18
# coefficients breaks at least one of them).
RK_ALPHA = [3296351145737 / 15110423921029, 1879360555526 / 7321162733569,
            10797097731880 / 20472212111779, 754636544611 / 15563872110659,
            3260218886217 / 2618290685819, 5069185909380 / 12292927838509]
RK_BETA = [-1204558336989 / 10607789004752, -3028468927040 / 14078136890693,
           -455570672869 / 8930094212428, -17275898420483 / 15997285579755,
           -2453906524165 / 9868353053862, 0.0]
RK_ERR = [-530312978447 / 9560368366154, 473021958881 / 2984707536468,
          -947229622805 / 10456009803779, -2921473878215 / 13334914072261,
          1519535112975 / 9264196100452, 167623581683 / 3930932046784]


def _rk_order_residuals(b, A, c):
    return np.array([b.sum() - 1, b @ c - 1 / 2, b @ c**2 - 1 / 3, b @ A @ c - 1 / 6,
                     b @ c**3 - 1 / 4, b @ (c * (A @ c)) - 1 / 8, b @ A @ c**2 - 1 / 12,
                     b @ A @ A @ c - 1 / 24])


def test_rk64_tableau_order_conditions():
    s = 6
    A = np.zeros((s, s))
    for i in range(1, s):
        for j in range(i):
            A[i, j] = RK_ALPHA[j] if j == i - 1 else RK_ALPHA[j] + RK_BETA[j]
    b = np.array(RK_ALPHA) + np.array(RK_BETA)
    c = A.sum(1)
    r4 = _rk_order_residuals(b, A, c)
    assert np.abs(r4).max() < 1e-14, r4                       # fourth order
    rh = _rk_order_residuals(b - np.array(RK_ERR), A, c)
    assert np.abs(rh[:4]).max() < 1e-14, rh                   # embedded: third order ...
    assert np.abs(rh[4:]).max() > 1e-4, rh                    # ... and not fourth


def test_rk64_fourth_order_convergence(mechs):
    """Fixed substeps (guess = min = max): the error against the Radau solution falls by ~2^4
    per halving of h until it reaches the fixture's own accuracy."""
    om = mechs["LiDryer"][2]
    d = np.load(os.path.join(GOLDEN, "radau_lidryer_32.npz"))
    errs = []
    for m in (16, 32, 64):
        r = om.react_rk64(d["rY0"], d["T0"], d["rE0"], float(d["dt"]), rtol=1.0, atol=1.0,
                          nsub_guess=m, nsub_min=m, nsub_max=m)
        assert r["nfail"] == 0 and (r["stats"][:, 0] == m).all()
        errs.append(np.abs(r["T"] - d["T1"]).max())
    assert errs[0] / errs[1] > 11 and errs[1] / errs[2] > 11, errs


@pytest.mark.parametrize("wl,fx,tol_T", [("lidryer_32", "radau_lidryer_32", 1e-6),
                                         ("drm19_64", "radau_drm19_64", 1e-4)])
def test_rk64_against_radau(mechs, wl, fx, tol_T):
    W = inputs.WORKLOADS[wl]
    d = np.load(os.path.join(GOLDEN, f"{fx}.npz"))
    r = mechs[W.mech][2].react_rk64(d["rY0"], d["T0"], d["rE0"], float(d["dt"]), rtol=1e-8,
                                 atol=1e-12, nsub_max=10**6, max_steps=10**6)
    assert r["nfail"] == 0
    assert np.abs(r["T"] - d["T1"]).max() < tol_T
    rho = d["rY1"].sum(0)
    assert (np.abs(r["rY"] - d["rY1"]) / rho).max() < 1e-8
    # the energy update is exact and mass is conserved
    assert np.array_equal(r["rE"], d["rE0"])
    np.testing.assert_allclose(r["rY"].sum(0), d["rY0"].sum(0), rtol=1e-13)


def test_rk64_mask_and_bad_input(mechs):
    om = mechs["LiDryer"][2]
    d = inputs.make_inputs("lidryer_32", 8)
    T = d["T"].copy()
    T[3] = -5.0
    mask = np.ones(8, np.int32)
    mask[5] = 0
    r = om.react_rk64(d["rY"], T, d["rE"], 1e-7, mask=mask)
    assert r["status"][3] == 4 and r["nfail"] == 1
    assert r["status"][5] == 0 and r["FC"][5] == 0
    assert np.array_equal(r["rY"][:, 5], d["rY"][:, 5]) and r["T"][5] == T[5]
