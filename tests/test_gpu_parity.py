"""GPU parity: the gfx950 kernels, called through the C-ABI, against the FP64 CPU oracle.

Tolerances (acceptance gate for the reactor outputs): max |dT| <= 1e-3 K and max |dY_k| <= 1e-8
against the oracle on identical inputs; EOS helpers to 1e-12 relative; integer/bookkeeping outputs
(mask, energy update, FC) exact.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pelec_amd import inputs, mechanism as M  # noqa: E402
from pelec_amd.reactor import DIAG_NAMES, Eos, Reactor  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
GATE_T = 1e-3
GATE_Y = 1e-8


def note(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def oracles(oracle_lib):
    return {m: oracle_lib.OracleMech(M.load(m)) for m in ("LiDryer", "drm19")}


def to_dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def run_gpu(wl_name, n, form="plain", lu_fp32=False, first=0, rYsrc=None, rEsrc=None, mask=None,
            d=None, kernel_timing=False, cooperative=False):
    W = inputs.WORKLOADS[wl_name]
    if d is None:
        d = inputs.make_inputs(wl_name, n, first)
    r = Reactor(W.mech)
    r.configure(rtol=W.rtol, atol=W.atol, form=form, lu_fp32=lu_fp32,
                kernel_timing=kernel_timing, cooperative=cooperative)
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    FC = torch.zeros_like(T)
    diag = torch.zeros((8, n), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    src = None if rYsrc is None else to_dev(rYsrc)
    esrc = None if rEsrc is None else to_dev(rEsrc)
    msk = None if mask is None else to_dev(mask, torch.int32)
    note(f"react {wl_name} n={n} form={form} lu_fp32={lu_fp32} cooperative={cooperative} "
         f"mask={mask is not None}")
    tnew = r.react_flat(rY, T, rE, W.dt, rYsrc=src, rEsrc=esrc, FC=FC, mask=msk, time=0.0)
    torch.cuda.synchronize()
    assert abs(tnew - W.dt) < 1e-20
    out = {"rY": rY.cpu().numpy(), "T": T.cpu().numpy(), "rE": rE.cpu().numpy(),
           "FC": FC.cpu().numpy(), "diag": diag.cpu().numpy(), "nfail": r.last_failures()}
    if kernel_timing:
        out["kernel_ms"], out["launches"] = r.kernel_times()
    r.close()
    return d, out


#: Where tight tolerances (typical-value atol), hot cells (T near 2400 K at rtol 1e-6, a local
#: error tolerance of 2.4e-3 K) or stiff ignition chemistry make the integration's step decisions
#: sensitive to rounding, other FP64 evaluations of the same integrations already disagree with
#: the oracle beyond the gate on some cells: its FMA build by up to 1.9e-3 K on the PistonBowl
#: typical-value case, 8.5e-3 K on the GRI-Mech 3.0 AMR sample and 1.3e-2 K on dodecane_lu; the
#: oracle itself, given one drm19 AMR cell's rho*e0 moved by a few ulps, by 1.5e-2 K and 6.9e-7,
#: with step counts from 241 to 281. There the GPU must meet the gate -- or the spread of
#: oracle_alternatives() on the same cells, where that is larger -- on every cell whose step and
#: RHS-evaluation counts equal the oracle's, and stay within SPREAD_FACTOR times that spread
#: elsewhere. The spread is the maximum of a heavy-tailed sample, hence the factor.
SPREAD_FACTOR = 8.0


def oracle_alternatives(om_fma, om, rY, T, rE, dt, **kw):
    """Other FP64 evaluations of the same integrations: the FMA build of the oracle, and the
    oracle with every cell's rho*e0 one ulp above and one ulp below (inputs equal to the last
    bit). Their distance from the oracle is the rounding sensitivity of each integration."""
    alts = [om_fma.react(rY, T, rE, dt, **kw)]
    for sgn in (1.0, -1.0):
        alts.append(om.react(rY, T, rE + sgn * np.spacing(rE), dt, **kw))
    return alts


def check_with_spread(label, T, rY, ref, alts, diag=None):
    """alts: oracle_alternatives() of the same cells (or one such result)."""
    alts = alts if isinstance(alts, list) else [alts]
    rho = ref["rY"].sum(0)
    dT = np.abs(T - ref["T"])
    dY = (np.abs(rY - ref["rY"]) / rho).max(0)
    sT = max(np.abs(a["T"] - ref["T"]).max() for a in alts)
    sY = max((np.abs(a["rY"] - ref["rY"]) / rho).max() for a in alts)
    ref_fma = alts[0]
    limT, limY = max(GATE_T, SPREAD_FACTOR * sT), max(GATE_Y, SPREAD_FACTOR * sY)
    msg = (f"{label}: max|dT| {dT.max():.2e} K max|dY| {dY.max():.2e}; oracle vs its FMA build "
           f"and one-ulp inputs {sT:.2e} K {sY:.2e}; cells over the gate "
           f"{int(((dT > GATE_T) | (dY > GATE_Y)).sum())} of {dT.size} (FMA build: "
           f"{int(((np.abs(ref_fma['T'] - ref['T']) > GATE_T)).sum())})")
    if diag is not None:
        same = (diag[1] == ref["stats"][:, 0]) & (diag[2] == ref["stats"][:, 1])
        msg += f"; {same.mean():.3f} with the oracle's step and RHS counts"
        note(msg)
        if same.any():
            # equal counts do not mean equal decisions throughout (an order choice or a Newton
            # iteration can differ and another compensate): these cells are held to the gate,
            # or to the oracle-to-oracle spread itself where that is larger
            assert (dT[same].max() <= max(GATE_T, sT) and
                    dY[same].max() <= max(GATE_Y, sY)), msg
    else:
        note(msg)
    assert dT.max() <= limT and dY.max() <= limY, msg


def compare(out, ref, act=None):
    if act is None:
        act = np.ones(out["T"].shape[0], dtype=bool)
    dT = np.abs(out["T"] - ref["T"])[act]
    rho = ref["rY"].sum(0)
    dY = (np.abs(out["rY"] - ref["rY"]) / rho)[:, act]
    return dT.max(), dY.max(), np.median(dT)


# ---------------------------------------------------------------- EOS helpers
@pytest.mark.parametrize("mech", ["LiDryer", "drm19"])
def test_eos_kernels(mech, oracles):
    from oracle.kinetics_np import NpMech
    m = M.load(mech)
    npm = NpMech(m)
    rng = np.random.default_rng(11)
    n = 4096
    T = rng.uniform(300.0, 3000.0, n)
    p = rng.uniform(0.5, 50.0, n) * 1013250.0
    Y = rng.uniform(0.0, 1.0, (n, m.nsp))
    Y /= Y.sum(1, keepdims=True)
    rho_ref, e_ref = npm.mixture(p, T, Y)
    eos = Eos(mech)
    note(f"eos {mech}")
    Yd = to_dev(Y.T)
    rho, e = eos.pyt2re(to_dev(p), to_dev(T), Yd)
    np.testing.assert_allclose(rho.cpu().numpy(), rho_ref, rtol=1e-13)
    np.testing.assert_allclose(e.cpu().numpy(), e_ref, rtol=1e-12, atol=1e-12 * np.abs(e_ref).max())
    Tg = to_dev(T * rng.uniform(0.7, 1.3, n))
    eos.rey2t(rho, e, Yd, Tg)
    np.testing.assert_allclose(Tg.cpu().numpy(), T, rtol=1e-11)
    hi = eos.rty2hi(to_dev(T)).cpu().numpy()
    hi_ref = (npm.h_RT(T) * 8.31446261815324e7 * T[:, None] / npm.W).T
    np.testing.assert_allclose(hi, hi_ref, rtol=1e-12, atol=1e-12 * np.abs(hi_ref).max())
    wd = eos.rty2wdot(to_dev(rho_ref), to_dev(T), Yd).cpu().numpy()
    om = oracles[mech]
    for c in range(0, n, 97):
        w_ref = om.wdot(rho_ref[c] * Y[c] / npm.W, T[c])
        assert np.abs(wd[:, c] - w_ref).max() <= 1e-11 * np.abs(w_ref).max() + 1e-300


# ---------------------------------------------------------------- reactor vs oracle
@pytest.mark.parametrize("wl,n", [("lidryer_32", 2048), ("drm19_64", 4096)])
def test_react_plain_fp64_vs_oracle(wl, n, oracles):
    W = inputs.WORKLOADS[wl]
    d, out = run_gpu(wl, n, form="plain")
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    assert out["nfail"] == 0 and (out["diag"][0] == 0).all()
    dT, dY, medT = compare(out, ref)
    note(f"{wl} plain: max|dT|={dT:.3e} K max|dY|={dY:.3e} median|dT|={medT:.3e}")
    assert dT <= GATE_T and dY <= GATE_Y
    # same algorithm: the step sequences agree cell by cell, except where rounding flips a decision
    same = (out["diag"][1] == ref["stats"][:, 0]).mean()
    assert same >= 0.9, same
    np.testing.assert_array_equal(out["FC"], out["diag"][2].astype(float))
    np.testing.assert_array_equal(out["rE"], d["rE"])


@pytest.mark.parametrize("lu_fp32,cooperative", [(False, False), (True, False), (False, True),
                                                   (True, True)])
def test_react_batched_vs_oracle(lu_fp32, cooperative, oracles):
    wl, n = "drm19_64", 4096
    W = inputs.WORKLOADS[wl]
    d, out = run_gpu(wl, n, form="batched", lu_fp32=lu_fp32, cooperative=cooperative)
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    assert out["nfail"] == 0 and (out["diag"][0] == 0).all()
    dT, dY, medT = compare(out, ref)
    note(f"batched lu_fp32={lu_fp32} cooperative={cooperative}: max|dT|={dT:.3e} K "
         f"max|dY|={dY:.3e} median|dT|={medT:.3e} "
         f"mean nni gpu/oracle={out['diag'][5].mean():.2f}/{ref['stats'][:, 4].mean():.2f}")
    assert dT <= GATE_T and dY <= GATE_Y
    np.testing.assert_array_equal(out["FC"], out["diag"][2].astype(float))


def test_react_batched_cooperative_lidryer(oracles):
    """Cooperative kernels on the 9-species mechanism (one reaction round, two efficiency sets)."""
    wl, n = "lidryer_32", 2048
    W = inputs.WORKLOADS[wl]
    d, out = run_gpu(wl, n, form="batched", lu_fp32=True, cooperative=True)
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    assert out["nfail"] == 0 and (out["diag"][0] == 0).all()
    dT, dY, medT = compare(out, ref)
    note(f"lidryer cooperative: max|dT|={dT:.3e} K max|dY|={dY:.3e} median|dT|={medT:.3e}")
    assert dT <= GATE_T and dY <= GATE_Y


@pytest.mark.parametrize("mech,wl", [("LiDryer", "lidryer_32"), ("drm19", "drm19_64")])
@pytest.mark.parametrize("cooperative", [False, True])
def test_model_eval_vs_oracle(mech, wl, cooperative, oracles):
    """RHS and exact Jacobian at given states (prc_model_eval) against the oracle's, for the
    one-cell-per-lane model and the cooperative (one cell per 32-lane group, table-driven
    kinetics) model. States: the workload's initial cells and the oracle's end states of the same
    cells (radical pools), with and without rhoY / rho e forcing. Bounds: species rates within
    1e-9 of the cell's largest species rate, the T rate within 1e-9 of its energy scale
    max_k |e_k f_k| / (rho c_v), and Jacobian rows within 1e-8 of the row's largest entry."""
    from oracle.kinetics_np import NpMech
    W = inputs.WORKLOADS[wl]
    om = oracles[mech]
    npm = NpMech(M.load(mech))
    d = inputs.make_inputs(wl, 48)
    ref = om.react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    rY = np.concatenate([d["rY"], ref["rY"]], axis=1)
    T = np.concatenate([d["T"], ref["T"]])
    rE = np.concatenate([d["rE"], d["rE"]])
    nc = T.shape[0]
    y = np.vstack([rY, T[None]])
    rng = np.random.default_rng(5)
    rho = rY.sum(0)
    rYsrc = rng.normal(size=rY.shape) * 1e-2 * rho
    rEsrc = rng.normal(size=nc) * 1e-3 * np.abs(rE)
    eos = Eos(mech)
    worst = [0.0, 0.0, 0.0]
    for forced in (False, True):
        note(f"model eval {mech} cooperative={cooperative} forced={forced} n={nc}")
        yd, J = eos.model(0.0, to_dev(y), to_dev(rE),
                          rhoesrc=to_dev(rEsrc) if forced else None,
                          rYsrc=to_dev(rYsrc) if forced else None, cooperative=cooperative)
        torch.cuda.synchronize()
        yd, J = yd.cpu().numpy(), J.cpu().numpy()
        for c in range(nc):
            args = (0.0, y[:, c], rE[c], rEsrc[c] if forced else 0.0,
                    rYsrc[:, c] if forced else None)
            f_ref, Tc = om.rhs(*args)
            J_ref = om.jac(*args)
            sc = np.abs(f_ref[:-1]).max()
            ek = npm.e_mass(Tc)
            cv = (y[:-1, c] * npm.cv_mass(Tc)).sum()
            scT = (np.abs(ek * f_ref[:-1]).max() + (abs(rEsrc[c]) if forced else 0.0)) / cv
            ey = np.abs(yd[:-1, c] - f_ref[:-1]).max() / sc
            eT = abs(yd[-1, c] - f_ref[-1]) / scT
            rows = np.abs(J_ref).max(1, keepdims=True)
            rows[rows == 0.0] = 1.0
            eJ = (np.abs(J[c] - J_ref) / rows).max()
            worst = [max(worst[0], ey), max(worst[1], eT), max(worst[2], eJ)]
            assert ey <= 1e-9 and eT <= 1e-9 and eJ <= 1e-8, (c, forced, ey, eT, eJ)
    note(f"  worst rel: ydot species {worst[0]:.2e}, T {worst[1]:.2e}, J {worst[2]:.2e}")


@pytest.mark.parametrize("wl,n,dt,rtol,atol", [("lidryer_32", 2048, None, None, None),
                                               ("drm19_64", 512, 1e-6, 1e-8, 1e-12)])
def test_react_rk64_vs_oracle(wl, n, dt, rtol, atol, oracles):
    """ReactorRK64 drop-in against the oracle's RK64 restatement (same tableau and controller).
    An explicit adaptive method's trajectories separate by about its truncation error when a
    round-off difference flips one of its ~10^3 accept/reject decisions, so the drm19 case runs at
    rtol 1e-8, where that error (2e-5 K against the Radau fixtures) is below the gate."""
    W = inputs.WORKLOADS[wl]
    dt = W.dt if dt is None else dt
    rtol = W.rtol if rtol is None else rtol
    atol = W.atol if atol is None else atol
    d = inputs.make_inputs(wl, n)
    r = Reactor(W.mech, chem_integrator="ReactorRK64")
    r.configure(rtol=rtol, atol=atol)
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    FC = torch.zeros_like(T)
    diag = torch.zeros((8, n), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    note(f"react_rk64 {wl} n={n} dt={dt}")
    r.react_flat(rY, T, rE, dt, FC=FC)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    r.close()
    ref = oracles[W.mech].react_rk64(d["rY"], d["T"], d["rE"], dt, rtol=rtol, atol=atol,
                                     nsub_guess=10, nsub_min=5, nsub_max=100000, nthreads=8)
    assert ref["nfail"] == 0
    out = {"rY": rY.cpu().numpy(), "T": T.cpu().numpy(), "rE": rE.cpu().numpy()}
    dT, dY, _ = compare(out, ref)
    note(f"  rk64 max|dT| {dT:.3e} K  max|dY| {dY:.3e}")
    assert dT <= GATE_T and dY <= GATE_Y
    assert np.array_equal(out["rE"], ref["rE"])
    dg = diag.cpu().numpy()
    assert (dg[0] == 0).all() and (dg[1] > 0).all()
    assert np.array_equal(FC.cpu().numpy(), dg[2].astype(np.float64))


def test_kernel_timing_mode_matches_graph():
    """The eager, event-timed iteration loop runs the same kernels as the graph replay: every
    per-cell result and counter is bit-identical, and the timed kernels report positive time."""
    d, g = run_gpu("drm19_64", 4096, form="batched", lu_fp32=True)
    _, e = run_gpu("drm19_64", 4096, form="batched", lu_fp32=True, d=d, kernel_timing=True)
    for k in ("rY", "T", "rE", "FC", "diag"):
        assert np.array_equal(g[k], e[k]), k
    assert e["launches"] > 0 and e["launches"] % 3 == 0
    assert all(v > 0.0 for v in e["kernel_ms"].values()), e["kernel_ms"]


def test_mask_and_sources(oracles):
    wl, n = "drm19_64", 1024
    W = inputs.WORKLOADS[wl]
    d = inputs.make_inputs(wl, n, 5000)
    rng = np.random.default_rng(12)
    mask = (rng.uniform(size=n) > 0.2).astype(np.int32)
    rEsrc = rng.normal(0, 1e9, n)
    rYsrc = rng.normal(0, 1e-3, (M.load(W.mech).nsp, n)) * d["rY"]
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rYsrc=rYsrc, rEsrc=rEsrc, mask=mask,
                                rtol=W.rtol, atol=W.atol)
    act = mask == 1
    for form in ("plain", "batched"):
        _, out = run_gpu(wl, n, form=form, rYsrc=rYsrc, rEsrc=rEsrc, mask=mask, d=d)
        assert out["nfail"] == 0
        assert np.array_equal(out["rY"][:, ~act], d["rY"][:, ~act])
        assert np.array_equal(out["T"][~act], d["T"][~act])
        assert np.array_equal(out["rE"][~act], d["rE"][~act])
        assert (out["FC"][~act] == 0).all()
        np.testing.assert_array_equal(out["rE"][act], d["rE"][act] + W.dt * rEsrc[act])
        dT, dY, _ = compare(out, ref, act)
        note(f"mask+src {form}: max|dT|={dT:.3e} max|dY|={dY:.3e}")
        assert dT <= GATE_T and dY <= GATE_Y


def test_react_box_array4_views():
    """The Array4 boundary: STemp-like storage (React.cpp:99,173-175) with T and rhoE as component
    views, a box with non-zero lo, must give what the flat call gives for the same cells."""
    wl = "lidryer_32"
    W = inputs.WORKLOADS[wl]
    nx, ny, nz = 8, 4, 3
    n = nx * ny * nz
    d = inputs.make_inputs(wl, n, 77)
    N = d["rY"].shape[0]
    lo = (5, -2, 11)
    hi = (lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
    stemp = torch.zeros((N + 2, nz, ny, nx), dtype=torch.float64, device=DEV)
    stemp[:N] = to_dev(d["rY"]).reshape(N, nz, ny, nx)
    stemp[N] = to_dev(d["T"]).reshape(nz, ny, nx)
    stemp[N + 1] = to_dev(d["rE"]).reshape(nz, ny, nx)
    fc = torch.zeros((1, nz, ny, nx), dtype=torch.float64, device=DEV)
    mask = torch.ones((1, nz, ny, nx), dtype=torch.int32, device=DEV)
    r = Reactor(W.mech)
    r.configure(rtol=W.rtol, atol=W.atol)
    note("react_box array4")
    t = r.react_box(lo, hi, stemp[:N], None, stemp[N:N + 1], stemp[N + 1:N + 2], None, fc, mask, W.dt,
                    time=1.5)
    torch.cuda.synchronize()
    assert abs(t - (1.5 + W.dt)) < 1e-15
    _, flat = run_gpu(wl, n, d=d)
    np.testing.assert_array_equal(stemp[:N].reshape(N, n).cpu().numpy(), flat["rY"])
    np.testing.assert_array_equal(stemp[N].reshape(n).cpu().numpy(), flat["T"])
    np.testing.assert_array_equal(fc.reshape(n).cpu().numpy(), flat["FC"])
    r.close()


def test_react_boxes_batched_vs_box_by_box():
    wl = "drm19_64"
    W = inputs.WORKLOADS[wl]
    N = M.load(W.mech).nsp
    shapes = [(4, 3, 2), (7, 1, 1), (2, 5, 3), (8, 8, 2), (1, 1, 1)]
    outs = {}
    for form in ("plain", "batched"):
        boxes, tensors = [], []
        first = 0
        for ib, (nx, ny, nz) in enumerate(shapes):
            n = nx * ny * nz
            d = inputs.make_inputs(wl, n, 9000 + first)
            first += n
            lo = (ib * 100, 0, -ib)
            hi = (lo[0] + nx - 1, ny - 1, -ib + nz - 1)
            rY = to_dev(d["rY"]).reshape(N, nz, ny, nx)
            T = to_dev(d["T"]).reshape(1, nz, ny, nx)
            rE = to_dev(d["rE"]).reshape(1, nz, ny, nx)
            FC = torch.zeros_like(T)
            boxes.append({"lo": lo, "hi": hi, "rY": rY, "T": T, "rE": rE, "FC": FC})
            tensors.append((rY, T, rE, FC))
        r = Reactor(W.mech)
        r.configure(rtol=W.rtol, atol=W.atol, form=form)
        note(f"react_boxes {form}")
        r.react_boxes(boxes, W.dt)
        torch.cuda.synchronize()
        assert r.last_failures() == 0
        outs[form] = [tuple(x.cpu().numpy().ravel() for x in t) for t in tensors]
        r.close()
    for a, b in zip(outs["plain"], outs["batched"]):
        assert np.abs(a[1] - b[1]).max() <= 1e-6   # T
        assert np.abs(a[0] - b[0]).max() <= 1e-12 * np.abs(a[0]).max()


def test_null_reactor_leaves_state():
    d = inputs.make_inputs("lidryer_32", 64)
    r = Reactor("LiDryer", chem_integrator="ReactorNull")
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    note("ReactorNull")
    t = r.react_flat(rY, T, rE, 1e-6, time=2.0)
    torch.cuda.synchronize()
    assert abs(t - (2.0 + 1e-6)) < 1e-15
    assert np.array_equal(rY.cpu().numpy(), d["rY"]) and np.array_equal(T.cpu().numpy(), d["T"])
    r.close()


def test_bad_input_status():
    d = inputs.make_inputs("lidryer_32", 8)
    d["T"][2] = -1.0
    d["rY"][:, 5] = 0.0
    for form in ("plain", "batched"):
        _, out = run_gpu("lidryer_32", 8, form=form, d={k: v.copy() for k, v in d.items()})
        assert list(np.nonzero(out["diag"][0])[0]) == [2, 5]
        assert (out["diag"][0][[2, 5]] == 4).all() and out["nfail"] == 2
        assert np.array_equal(out["T"][[2, 5]], d["T"][[2, 5]])


def test_full_size_invariants():
    """drm19 at 64^3 through the batched mixed-precision form: size-independent properties."""
    wl = "drm19_64"
    W = inputs.WORKLOADS[wl]
    m = M.load(W.mech)
    n = W.ncells
    d, out = run_gpu(wl, n, form="batched", lu_fp32=True)
    assert out["nfail"] == 0 and (out["diag"][0] == 0).all()
    rho0, rho1 = d["rY"].sum(0), out["rY"].sum(0)
    assert np.abs(rho1 / rho0 - 1).max() <= 1e-10
    np.testing.assert_array_equal(out["rE"], d["rE"])
    Wt = np.array(m.weights)
    A = np.array(m.element_matrix())
    el0 = A @ (d["rY"] / Wt[:, None])
    el1 = A @ (out["rY"] / Wt[:, None])
    assert (np.abs(el1 - el0) <= 1e-9 * np.abs(el0) + 1e-20).all()
    # returned T consistent with returned energy and composition
    th = inputs.Thermo(m)
    e1 = ((out["rY"] / rho1).T * th.e_mass(out["T"])).sum(-1)
    assert np.abs(e1 - out["rE"] / rho1).max() <= 1e-10 * np.abs(out["rE"] / rho1).max()
    steps = out["diag"][1]
    note(f"full size: steps mean {steps.mean():.1f} max {steps.max()} ignited cells "
         f"{(out['T'] - d['T'] > 100).mean():.3f}")
    assert steps.max() > 20 * max(steps.min(), 1)   # uneven stiffness is really there


@pytest.mark.parametrize("form,lu_fp32", [("plain", False), ("batched", True)])
def test_react_state_box_vs_oracle(form, lu_fp32, oracle_lib):
    """PeleC::react_state on one 8x8x4 box (S_new = S_old + dt S_nr, covered cells, cost output)
    against the numpy restatement around the oracle reactor."""
    from oracle import react_state_np as RS
    from oracle.kinetics_np import NpMech
    from pelec_amd.inputs import make_conservative_state as make_state
    W = inputs.WORKLOADS["drm19_64"]
    m = M.load("drm19")
    om, npm = oracle_lib.OracleMech(m), NpMech(m)
    nx, ny, nz = 8, 8, 4
    n = nx * ny * nz
    S, nonrs = make_state("drm19_64", n, seed=11)
    Snew0 = S + W.dt * nonrs
    mask = np.ones(n, np.int32)
    mask[[5, 77, 200]] = 0
    refS, refIR, refFC = RS.react_state(om, npm, S, Snew0, nonrs, W.dt, mask=mask, rtol=W.rtol,
                                        atol=W.atol, nthreads=8)
    shp = lambda a: to_dev(a.reshape(a.shape[0], nz, ny, nx))
    dS_old, dS_new, dnr = shp(S), shp(Snew0), shp(nonrs)
    IR = torch.zeros((m.nsp + 2, nz, ny, nx), dtype=torch.float64, device=DEV)
    cost = torch.zeros((nz, ny, nx), dtype=torch.float64, device=DEV)
    dmask = to_dev(mask.reshape(nz, ny, nx), torch.int32)
    r = Reactor("drm19")
    r.configure(rtol=W.rtol, atol=W.atol, form=form, lu_fp32=lu_fp32)
    lo, hi = (3, -2, 5), (3 + nx - 1, -2 + ny - 1, 5 + nz - 1)
    note(f"react_state_box drm19 {nx}x{ny}x{nz} form={form} lu_fp32={lu_fp32}")
    tnew = r.react_state_box(lo, hi, dS_old, dS_new, dnr, IR, W.dt, mask=dmask, cost=cost)
    torch.cuda.synchronize()
    assert abs(tnew - W.dt) < 1e-20 and r.last_failures() == 0
    r.close()
    gS = dS_new.cpu().numpy().reshape(S.shape[0], n)
    gIR = IR.cpu().numpy().reshape(m.nsp + 2, n)
    N = m.nsp
    rho = refS[0]
    assert np.abs(gS[6] - refS[6]).max() <= GATE_T                                 # T
    assert (np.abs(gS[7:7 + N] - refS[7:7 + N]) / rho).max() <= GATE_Y             # rhoY
    np.testing.assert_allclose(gS[:6], refS[:6], rtol=1e-12, atol=1e-300)         # rho, rho u, E, e
    assert (np.abs(gIR[:N] - refIR[:N]) / (rho / W.dt)).max() <= GATE_Y            # rho wdot
    np.testing.assert_allclose(gIR[N], refIR[N], rtol=1e-9, atol=1e-6 * np.abs(S[4]).max() / W.dt)
    np.testing.assert_allclose(gIR[N + 1], refIR[N + 1], rtol=1e-4,
                               atol=1e-6 * np.abs(refIR[N + 1]).max())
    cov = mask == 0
    assert np.array_equal(gS[:, cov], Snew0[:, cov]) and (gIR[:, cov] == 0).all()
    gc = cost.cpu().numpy().reshape(n)
    assert (gc[cov] == 0).all() and (gc[~cov] > 0).all()
    # the cost output is the cell's RHS-evaluation count, the oracle's FC: equal wherever the two
    # integrations took the same step sequence (rounding can flip a step decision in a few cells)
    same = gc[~cov] == refFC[~cov]
    note(f"  cost == oracle FC on {same.mean():.3f} of the active cells")
    assert same.mean() >= 0.9
    assert np.abs(gc[~cov] - refFC[~cov]).max() <= 0.25 * refFC[~cov].max()


PISTONBOWL_TYP_VALS = [1.e-7, 3.e-9, 1.e-8, 3.e-5, 7.e-8, 1.e-6, 1.e-7, 5.e-9, 1.e-9, 9.e-7, 6.e-6,
                       1.e-6, 1.7e-8, 1.9e-8, 7.4e-7, 6.7e-9, 3.5e-7, 1.6e-8, 1.3e-8, 1.e-4,
                       1.8e-33, 2300.0]  # Exec/Production/PistonBowl/inputs/example.inp:34 (drm19)


@pytest.fixture(scope="module")
def oracle_fma(oracle_lib):
    cache = {}

    def get(mech):
        if mech not in cache:
            cache[mech] = oracle_lib.OracleMech(M.load(mech), variant="fma")
        return cache[mech]
    return get


def test_graph_compaction_matches_host_compaction(oracles):
    """prc.compact_every: the captured graph compacts the work queue itself (gate kernel, then the
    compaction kernels) between iterations. Compaction only moves slots, so with no tail graph the
    end states are bit-identical to host-only compaction, in graph replays and in the eager,
    event-timed iterations; both meet the oracle gates (drm19, 16384 cells, FP64 factors)."""
    wl, n = "drm19_64", 16384
    W = inputs.WORKLOADS[wl]
    d = inputs.make_inputs(wl, n)
    outs = {}
    for every, kt in ((0, False), (3, False), (3, True)):
        r = Reactor(W.mech)
        r.configure(rtol=W.rtol, atol=W.atol, form="batched", lu_fp32=False, kernel_timing=kt)
        r.set_param("prc.iters_per_graph", 15)
        r.set_param("prc.compact_pct", 90)
        r.set_param("prc.compact_every", every)
        rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
        note(f"graph compaction every {every} kernel_timing={kt} n={n}")
        r.react_flat(rY, T, rE, W.dt)
        torch.cuda.synchronize()
        assert r.last_failures() == 0
        assert r.get_param("prc.compact_every") == every
        r.close()
        outs[(every, kt)] = {"rY": rY.cpu().numpy(), "T": T.cpu().numpy()}
    for k in ((3, False), (3, True)):
        assert np.array_equal(outs[k]["T"], outs[(0, False)]["T"]), k
        assert np.array_equal(outs[k]["rY"], outs[(0, False)]["rY"]), k
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    dT, dY, _ = compare(outs[(3, False)], ref)
    note(f"  max|dT|={dT:.3e} K max|dY|={dY:.3e}")
    assert dT <= GATE_T and dY <= GATE_Y


def test_react_rk64_workload_tolerance_vs_oracle(oracles, oracle_fma):
    """ReactorRK64 on drm19 at the BDF workload's own tolerances (rtol 1e-6, atol 1e-10; dt 1e-6,
    512 cells). At this tolerance a round-off difference flips accept/reject decisions of the
    explicit controller, and trajectories separate by its truncation error: every cell is held to
    max(gate, 8 x) the spread the oracle shows against its own FMA build and one-ulp inputs
    (check_with_spread), and cells whose step and RHS counts equal the oracle's to the gate."""
    W = inputs.WORKLOADS["drm19_64"]
    n, dt, rtol, atol = 512, 1e-6, W.rtol, W.atol
    d = inputs.make_inputs("drm19_64", n)
    r = Reactor(W.mech, chem_integrator="ReactorRK64")
    r.configure(rtol=rtol, atol=atol)
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    diag = torch.zeros((8, n), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    note(f"react_rk64 drm19 n={n} dt={dt} rtol={rtol}")
    r.react_flat(rY, T, rE, dt)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    r.close()
    kw = dict(rtol=rtol, atol=atol, nsub_guess=10, nsub_min=5, nsub_max=100000, nthreads=8)
    om, omf = oracles[W.mech], oracle_fma(W.mech)
    ref = om.react_rk64(d["rY"], d["T"], d["rE"], dt, **kw)
    alts = [omf.react_rk64(d["rY"], d["T"], d["rE"], dt, **kw)]
    for sgn in (1.0, -1.0):
        alts.append(om.react_rk64(d["rY"], d["T"], d["rE"] + sgn * np.spacing(d["rE"]), dt, **kw))
    assert ref["nfail"] == 0
    dg = diag.cpu().numpy()
    check_with_spread("rk64 drm19 rtol 1e-6", T.cpu().numpy(), rY.cpu().numpy(), ref, alts, diag=dg)


@pytest.mark.parametrize("knob", ["typ_vals", "max_order4"])
@pytest.mark.parametrize("form,lu_fp32", [("plain", False), ("batched", True)])
def test_production_knobs_vs_oracle(knob, form, lu_fp32, oracles, oracle_fma):
    """The production decks' chemistry settings against the oracle with the same settings:
    typical-value absolute tolerances (set_typ_vals_ode with PistonBowl's drm19 ode.typ_vals,
    atol_i = atol * typ_i) and cvode.max_order = 4 (JetFlame/inputs.inp:31,
    ChallengeProblem/challenge.inp:36). The typical values tighten some species' tolerances to
    1e-18, which makes step decisions rounding-sensitive: the gate is check_with_spread's."""
    W = inputs.WORKLOADS["drm19_64"]
    n = 2048
    d = inputs.make_inputs(W.name, n)
    kw = {"typ": np.array(PISTONBOWL_TYP_VALS)} if knob == "typ_vals" else {"max_order": 4}
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol,
                                nthreads=8, **kw)
    ref_fma = oracle_alternatives(oracle_fma(W.mech), oracles[W.mech], d["rY"], d["T"], d["rE"],
                                  W.dt, rtol=W.rtol, atol=W.atol, nthreads=8, **kw)
    assert (ref["status"] == 0).all()
    r = Reactor(W.mech)
    r.configure(rtol=W.rtol, atol=W.atol, form=form, lu_fp32=lu_fp32,
                max_order=kw.get("max_order"))
    if knob == "typ_vals":
        r.set_typ_vals_ode(PISTONBOWL_TYP_VALS)
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    diag = torch.zeros((8, n), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    note(f"react drm19 n={n} {knob} form={form} lu_fp32={lu_fp32}")
    r.react_flat(rY, T, rE, W.dt)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    r.close()
    dg = diag.cpu().numpy()
    check_with_spread(f"  {knob} {form}", T.cpu().numpy(), rY.cpu().numpy(), ref, ref_fma, dg)
    assert (dg[1] == ref["stats"][:, 0]).mean() >= 0.9


def test_plain_fp32_lu_refinement_vs_oracle(oracles):
    """The plain form with FP32 Newton-matrix factors honours prc.lu_refine: with the default two
    FP64 refinement steps it meets the gate against the oracle."""
    W = inputs.WORKLOADS["drm19_64"]
    n = 2048
    d, out = run_gpu(W.name, n, form="plain", lu_fp32=True)
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol,
                                nthreads=8)
    assert out["nfail"] == 0
    dT, dY, _ = compare(out, ref)
    note(f"plain FP32 LU (2 refinements): max|dT|={dT:.3e} K max|dY|={dY:.3e}")
    assert dT <= GATE_T and dY <= GATE_Y


@pytest.mark.parametrize("cooperative", [False, True])
def test_batched_slot_refill_vs_oracle(cooperative, oracles):
    """A call with more cells than work-queue slots (prc.max_slots): freed slots are refilled from
    the call's cell queue with matrix blocks from the free ring. Per-cell arithmetic does not depend
    on the slot, so the result equals the one-slot-per-cell call bit for bit, and meets the gate
    against the oracle. 20 % of the cells are masked (their blocks go straight back to the ring)."""
    W = inputs.WORKLOADS["drm19_64"]
    n = 6000
    d = inputs.make_inputs(W.name, n)
    mask = (inputs.uniform(77, 0, np.arange(n)) >= 0.2).astype(np.int32)
    outs = []
    for slots in (0, 1024):
        r = Reactor(W.mech)
        r.configure(rtol=W.rtol, atol=W.atol, form="batched", lu_fp32=True,
                    cooperative=cooperative, max_slots=slots)
        rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
        diag = torch.zeros((8, n), dtype=torch.int32, device=DEV)
        r.set_diagnostics(diag)
        note(f"react drm19 n={n} batched max_slots={slots} cooperative={cooperative}")
        r.react_flat(rY, T, rE, W.dt, mask=to_dev(mask, torch.int32))
        torch.cuda.synchronize()
        assert r.last_failures() == 0
        assert r.get_param("prc.slots") == (1024 if slots else n)
        r.close()
        outs.append({"rY": rY.cpu().numpy(), "T": T.cpu().numpy(), "rE": rE.cpu().numpy(),
                     "diag": diag.cpu().numpy()})
    np.testing.assert_array_equal(outs[1]["T"], outs[0]["T"])
    np.testing.assert_array_equal(outs[1]["rY"], outs[0]["rY"])
    np.testing.assert_array_equal(outs[1]["diag"], outs[0]["diag"])
    off = mask == 0
    np.testing.assert_array_equal(outs[1]["T"][off], d["T"][off])
    act = np.flatnonzero(~off)[::4]
    ref = oracles[W.mech].react(d["rY"][:, act], d["T"][act], d["rE"][act], W.dt, rtol=W.rtol,
                                atol=W.atol, nthreads=8)
    dT, dY, _ = compare({"rY": outs[1]["rY"][:, act], "T": outs[1]["T"][act]}, ref)
    note(f"  refilled call: max|dT|={dT:.3e} K max|dY|={dY:.3e} on {act.size} cells")
    assert dT <= GATE_T and dY <= GATE_Y


# ---------------------------------------------------------------- GRI-Mech 3.0 (53 species)
@pytest.fixture(scope="module")
def gri_oracle(oracle_lib):
    return oracle_lib.OracleMech(M.load("gri30"))


def test_gri30_vs_oracle(gri_oracle, oracle_fma):
    """GRI-Mech 3.0 (53 species, 325 reactions; n = 54): plain FP64, batched FP64-LU and batched
    FP32-LU against the oracle on 512 cells of the gri30 AMR state distribution (check_with_spread).
    The Newton and Jacobian kernels run one cell per 64-lane wave, the Newton matrix is factorised
    by one 64-lane group; the plain form keeps its matrices in a device workspace."""
    W = inputs.WORKLOADS["gri30_flat"]
    n = 512
    d = inputs.make_inputs(W.name, n)
    args = (d["rY"], d["T"], d["rE"], W.dt)
    ref = gri_oracle.react(*args, rtol=W.rtol, atol=W.atol, nthreads=8)
    ref_fma = oracle_alternatives(oracle_fma("gri30"), gri_oracle, *args, rtol=W.rtol,
                                  atol=W.atol, nthreads=8)
    assert (ref["status"] == 0).all()
    for form, fp32 in (("batched", False), ("batched", True), ("plain", False)):
        d2, out = run_gpu(W.name, n, form=form, lu_fp32=fp32, d=d)
        assert out["nfail"] == 0
        note(f"  Newton iterations/cell {out['diag'][5].mean():.2f} (oracle "
             f"{ref['stats'][:, 4].mean():.2f})")
        check_with_spread(f"  gri30 {form} lu_fp32={fp32}", out["T"], out["rY"], ref, ref_fma,
                          out["diag"])
        np.testing.assert_array_equal(out["FC"], out["diag"][2].astype(float))


def test_gri30_amr_batch_vs_oracle(gri_oracle, oracle_fma):
    """The GRI-Mech 3.0 AMR-shaped batch (8 boxes, edges from {8..64}, 20 % masked) in one
    prc_react_boxes call of the batched FP32-LU form with fewer slots than cells, so the call
    refills freed slots. Masked cells come back bit-identical; a sample of unmasked cells is
    checked against the oracle (check_with_spread)."""
    b = inputs.make_amr_batch(8, "gri30")
    wl = b["workload"]
    nc = b["ncells"]
    N = b["rY"].shape[0]
    rY, T, rE = to_dev(b["rY"]), to_dev(b["T"]), to_dev(b["rE"])
    mask = to_dev(b["mask"], torch.int32)
    views = []
    for bx in b["boxes"]:
        nx, ny, nz = bx["shape"]
        o, n = bx["first"], nx * ny * nz
        views.append({"lo": bx["lo"], "hi": bx["hi"], "rY": rY[:, o:o + n].view(N, nz, ny, nx),
                      "T": T[o:o + n].view(1, nz, ny, nx), "rE": rE[o:o + n].view(1, nz, ny, nx),
                      "mask": mask[o:o + n].view(nz, ny, nx)})
    r = Reactor(wl.mech)
    r.configure(rtol=wl.rtol, atol=wl.atol, form="batched", lu_fp32=True, max_slots=131072)
    diag = torch.zeros((8, nc), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    note(f"react_boxes gri30_amr: {len(views)} boxes, {nc} cells, 131072 slots")
    r.react_boxes(views, wl.dt)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    assert r.get_param("prc.slots") == 131072 < nc
    r.close()
    Tn, Yn, En, dg = T.cpu().numpy(), rY.cpu().numpy(), rE.cpu().numpy(), diag.cpu().numpy()
    off = b["mask"] == 0
    np.testing.assert_array_equal(Tn[off], b["T"][off])
    np.testing.assert_array_equal(Yn[:, off], b["rY"][:, off])
    np.testing.assert_array_equal(En[off], b["rE"][off])
    assert (dg[2, off] == 0).all() and (dg[2, ~off] > 0).all()
    act = np.flatnonzero(~off)
    idx = act[:: max(1, act.size // 256)][:256]
    args = (b["rY"][:, idx], b["T"][idx], b["rE"][idx], wl.dt)
    ref = gri_oracle.react(*args, rtol=wl.rtol, atol=wl.atol, nthreads=8)
    ref_fma = oracle_alternatives(oracle_fma("gri30"), gri_oracle, *args, rtol=wl.rtol,
                                  atol=wl.atol, nthreads=8)
    check_with_spread(f"  gri30 amr sample {idx.size}", Tn[idx], Yn[:, idx], ref, ref_fma,
                      dg[:, idx])


# ---------------------------------------------------------------- dodecane_lu (53 species)
def test_dodecane_lu_fp32_vs_fp64_lu_vs_oracle(oracle_lib, oracle_fma):
    """dodecane_lu (53 species, 229 reactions; n = 54), the mixed-precision configuration: the
    batched form with FP32 Newton-matrix factors (equilibrated, FP64 refinement) and with FP64
    factors, and the plain FP64 form, against the oracle on 384 cells of the dodecane_64 states
    (T ~ U[700, 1800] K, p ~ U[10, 60] atm; 40 % ignite within dt). Newton iteration counts of
    the two factor precisions are reported beside the oracle's; the gate is check_with_spread's
    (the ignition chemistry makes step decisions rounding-sensitive: two FP64 oracle builds take
    different step sequences on a quarter of these cells)."""
    W = inputs.WORKLOADS["dodecane_64"]
    n = 384
    d = inputs.make_inputs(W.name, n)
    args = (d["rY"], d["T"], d["rE"], W.dt)
    om = oracle_lib.OracleMech(M.load(W.mech))
    ref = om.react(*args, rtol=W.rtol, atol=W.atol, nthreads=8)
    ref_fma = oracle_alternatives(oracle_fma(W.mech), om, *args, rtol=W.rtol, atol=W.atol,
                                  nthreads=8)
    assert (ref["status"] == 0).all()
    for form, fp32 in (("batched", True), ("batched", False), ("plain", False)):
        d2, out = run_gpu(W.name, n, form=form, lu_fp32=fp32, d=d)
        assert out["nfail"] == 0
        note(f"  Newton iterations/cell {out['diag'][5].mean():.2f} (oracle "
             f"{ref['stats'][:, 4].mean():.2f})")
        check_with_spread(f"  dodecane_lu {form} lu_fp32={fp32}", out["T"], out["rY"], ref,
                          ref_fma, out["diag"])


def test_compute_temp_vs_oracle():
    from oracle import react_state_np as RS
    from oracle.kinetics_np import NpMech
    from pelec_amd.inputs import make_conservative_state as make_state
    m = M.load("drm19")
    npm = NpMech(m)
    n = 4 * 4 * 4
    S, _ = make_state("drm19_64", n, seed=5)
    S[6] *= np.random.default_rng(1).uniform(0.8, 1.2, n)   # stale temperature guess
    mask = np.ones(n, np.int32)
    mask[9] = 0
    ref = RS.compute_temp(npm, S, mask=mask)
    dS = to_dev(S.reshape(S.shape[0], 4, 4, 4))
    note("compute_temp drm19 4x4x4")
    Eos("drm19").compute_temp((0, 0, 0), (3, 3, 3), dS, mask=to_dev(mask.reshape(4, 4, 4), torch.int32))
    torch.cuda.synchronize()
    g = dS.cpu().numpy().reshape(S.shape[0], n)
    np.testing.assert_allclose(g[6], ref[6], rtol=1e-11)
    assert g[6, 9] == S[6, 9]
    assert np.array_equal(np.delete(g, 6, 0), np.delete(S, 6, 0))


def test_tail_graph_switch_vs_oracle(oracles):
    """The tail graph (cooperative Newton kernel once at most prc.coop_tail_slots slots are in use)
    taken over mid-call, after compactions: 16384 cells, compaction below 75 % live, tail at 8192
    slots. Same oracle gates; the kernel-timing (eager) replay of the same shape agrees."""
    wl, n = "drm19_64", 16384
    W = inputs.WORKLOADS[wl]
    d = inputs.make_inputs(wl, n)
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    outs = []
    for kt in (False, True):
        r = Reactor(W.mech)
        r.configure(rtol=W.rtol, atol=W.atol, form="batched", lu_fp32=True, cooperative=2,
                    kernel_timing=kt)
        r.set_param("prc.compact_pct", 75)
        r.set_param("prc.coop_tail_slots", 8192)
        r.set_param("prc.blocks_tail", 16)
        rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
        note(f"tail graph switch n={n} kernel_timing={kt}")
        r.react_flat(rY, T, rE, W.dt)
        torch.cuda.synchronize()
        assert r.last_failures() == 0
        assert r.shape()["prc.coop_tail_slots"] == 8192
        out = {"rY": rY.cpu().numpy(), "T": T.cpu().numpy()}
        r.close()
        dT, dY, medT = compare(out, ref)
        note(f"  max|dT|={dT:.3e} K max|dY|={dY:.3e}")
        assert dT <= GATE_T and dY <= GATE_Y
        outs.append(out)
    np.testing.assert_array_equal(outs[0]["T"], outs[1]["T"])


def test_large_system_n41_vs_oracle(oracle_lib):
    """n = 41 > 32 (mech/drm19_inert40.yaml: drm19 plus 19 inert species with N2's thermo): the
    code paths of the 53-species configurations -- one 64-lane group per Newton-matrix LU, the
    Newton and Jacobian kernels with one cell per 64-lane wave, the plain kernel with its
    matrices in a device workspace and its kinetics out of line -- for plain FP64, batched FP64
    and batched FP32-LU, against the oracle on the same mechanism. 40 % of each cell's N2 is moved
    into the inert species (same W and e_k, so rho and rho*e are unchanged); they must come out
    unchanged. The 32-lane cooperative kernels (prc.cooperative) do not exist above n = 32."""
    m = M.load("drm19_inert40")
    om = oracle_lib.OracleMech(m)
    W = inputs.WORKLOADS["drm19_64"]
    n = 1024
    d = inputs.make_inputs("drm19_64", n)
    iN2 = m.species_index("N2")
    rY = np.zeros((m.nsp, n))
    rY[:21] = d["rY"]
    moved = 0.4 * rY[iN2]
    rY[iN2] -= moved
    rY[21:] = moved / 19.0
    ref = om.react(rY, d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol)
    assert (ref["status"] == 0).all()
    for form, fp32 in (("batched", False), ("batched", True), ("plain", False)):
        r = Reactor("drm19_inert40")
        r.configure(rtol=W.rtol, atol=W.atol, form=form, lu_fp32=fp32)
        if form == "batched":
            with pytest.raises(Exception):
                r.set_param("prc.cooperative", 2)  # no cooperative kernels above n = 32
        rYd, T, rE = to_dev(rY), to_dev(d["T"]), to_dev(d["rE"])
        note(f"n=41 react form={form} lu_fp32={fp32} cells={n}")
        r.react_flat(rYd, T, rE, W.dt)
        torch.cuda.synchronize()
        assert r.last_failures() == 0
        out = {"rY": rYd.cpu().numpy(), "T": T.cpu().numpy()}
        r.close()
        dT, dY, _ = compare(out, ref)
        note(f"  max|dT|={dT:.3e} K max|dY|={dY:.3e}")
        assert dT <= GATE_T and dY <= GATE_Y
        np.testing.assert_allclose(out["rY"][21:], rY[21:], rtol=1e-12)


def test_autotune_shape_and_parity(oracles):
    """prc_autotune picks a launch shape by timing the batched form on a sample; the tuned shape
    is valid and the reactor still matches the oracle with it."""
    W = inputs.WORKLOADS["drm19_64"]
    n = 4096
    d = inputs.make_inputs("drm19_64", n)
    r = Reactor(W.mech)
    r.configure(rtol=W.rtol, atol=W.atol, form="batched", lu_fp32=True)
    note("autotune drm19 4096 cells")
    shape, ms = r.autotune(to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"]), W.dt)
    note(f"  tuned shape {shape}, {ms:.2f} ms")
    assert ms > 0.0 and shape["prc.iters_per_graph"] % 3 == 0
    assert all(shape[k] >= 1 for k in ("prc.blocks_newton", "prc.blocks_jac", "prc.blocks_lu"))
    assert 0 <= shape["prc.cooperative"] <= 3 and shape["prc.coop_tail_slots"] >= 0
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    r.react_flat(rY, T, rE, W.dt)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    r.close()
    ref = oracles[W.mech].react(d["rY"], d["T"], d["rE"], W.dt, rtol=W.rtol, atol=W.atol,
                                nthreads=8)
    dT, dY, _ = compare({"rY": rY.cpu().numpy(), "T": T.cpu().numpy()}, ref)
    assert dT <= GATE_T and dY <= GATE_Y


# ---------------------------------------------------------------- singular Newton matrices
@pytest.mark.parametrize("coop,tail", [(0, 0), (3, 0), (0, 4096)])
def test_singular_newton_matrix_handling(coop, tail):
    """Every factorisation reports a singular Newton matrix (prc.test_lu_fail, set from the host).
    The factorisation kernel tags the cell and the next iteration's Newton kernel runs CVODE's
    convergence-failure handling (cvHandleNFlag: restore the prediction, h *= ETACF, a fresh
    set-up) until MXNCF = 10 failures, when the cell ends with CELL_CONVFAIL and keeps its inputs.
    Cases: the one-cell-per-lane Newton kernel, the one-cell-per-group kernel (cooperative), and
    the tail graph's group kernel taking over from the first iteration."""
    W = inputs.WORKLOADS["lidryer_32"]
    n = 512
    d = inputs.make_inputs(W.name, n)
    r = Reactor(W.mech)
    r.configure(rtol=W.rtol, atol=W.atol, form="batched", lu_fp32=True, cooperative=coop)
    if tail:
        r.set_param("prc.coop_tail_slots", tail)
    r.set_param("prc.test_lu_fail", 1)
    rY, T, rE = to_dev(d["rY"]), to_dev(d["T"]), to_dev(d["rE"])
    diag = torch.zeros((8, n), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    note(f"react lidryer_32 n={n} batched, every Newton matrix reported singular, "
         f"cooperative={coop} tail={tail}")
    r.react_flat(rY, T, rE, W.dt)
    torch.cuda.synchronize()
    dg = diag.cpu().numpy()
    assert r.last_failures() == n
    np.testing.assert_array_equal(dg[0], 3)   # CELL_CONVFAIL
    np.testing.assert_array_equal(dg[7], 10)  # ncfn: MXNCF convergence failures
    np.testing.assert_array_equal(dg[1], 0)   # no step accepted
    np.testing.assert_array_equal(dg[4], 10)  # one set-up per attempt
    np.testing.assert_array_equal(T.cpu().numpy(), d["T"])
    np.testing.assert_array_equal(rY.cpu().numpy(), d["rY"])
    np.testing.assert_array_equal(rE.cpu().numpy(), d["rE"])
    # the knob off again: the same reactor integrates normally
    r.set_param("prc.test_lu_fail", 0)
    T2 = to_dev(d["T"])
    r.react_flat(to_dev(d["rY"]), T2, to_dev(d["rE"]), W.dt)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    r.close()


# ---------------------------------------------------------------- AMR-shaped batch
def test_amr_batch_vs_oracle(oracles, oracle_fma):
    """The AMR-shaped batch (inputs.make_amr_batch: boxes with per-axis edges from {8..64}, 20 %
    of the cells masked out, T ~ U[800, 2400] K, p ~ U[1, 20] atm), here 8 boxes, integrated in one
    prc_react_boxes call of the batched FP32-LU form. Masked cells come back bit-identical. On a
    sample of unmasked cells against the oracle: check_with_spread (cells with the oracle's step
    and RHS counts at the gate; every cell, T and Y, within the oracle-to-oracle spread)."""
    b = inputs.make_amr_batch(8)
    wl = b["workload"]
    nc = b["ncells"]
    N = b["rY"].shape[0]
    rY, T, rE = to_dev(b["rY"]), to_dev(b["T"]), to_dev(b["rE"])
    mask = to_dev(b["mask"], torch.int32)
    views = []
    for bx in b["boxes"]:
        nx, ny, nz = bx["shape"]
        o, n = bx["first"], nx * ny * nz
        views.append({"lo": bx["lo"], "hi": bx["hi"], "rY": rY[:, o:o + n].view(N, nz, ny, nx),
                      "T": T[o:o + n].view(1, nz, ny, nx), "rE": rE[o:o + n].view(1, nz, ny, nx),
                      "mask": mask[o:o + n].view(nz, ny, nx)})
    r = Reactor(wl.mech)
    r.configure(rtol=wl.rtol, atol=wl.atol, form="batched", lu_fp32=True)
    diag = torch.zeros((8, nc), dtype=torch.int32, device=DEV)
    r.set_diagnostics(diag)
    note(f"react_boxes drm19_amr: {len(views)} boxes, {nc} cells")
    r.react_boxes(views, wl.dt)
    torch.cuda.synchronize()
    assert r.last_failures() == 0
    r.close()
    Tn, Yn, En, dg = T.cpu().numpy(), rY.cpu().numpy(), rE.cpu().numpy(), diag.cpu().numpy()
    off = b["mask"] == 0
    np.testing.assert_array_equal(Tn[off], b["T"][off])
    np.testing.assert_array_equal(Yn[:, off], b["rY"][:, off])
    np.testing.assert_array_equal(En[off], b["rE"][off])
    act = np.flatnonzero(~off)
    idx = act[:: max(1, act.size // 512)][:512]
    args = (b["rY"][:, idx], b["T"][idx], b["rE"][idx], wl.dt)
    ref = oracles[wl.mech].react(*args, rtol=wl.rtol, atol=wl.atol, nthreads=8)
    ref_fma = oracle_alternatives(oracle_fma(wl.mech), oracles[wl.mech], *args, rtol=wl.rtol,
                                  atol=wl.atol, nthreads=8)
    check_with_spread(f"amr sample {idx.size}", Tn[idx], Yn[:, idx], ref, ref_fma, dg[:, idx])
    assert ((dg[1, idx] == ref["stats"][:, 0]) & (dg[2, idx] == ref["stats"][:, 1])).mean() >= 0.99
