"""Writes tests/golden/radau_<workload>.npz: tight-tolerance reference solutions of the reactor ODE.

For sample cells of each workload, the constant-volume rho*e reactor (oracle/kinetics_np.py RHS,
the numpy restatement) is integrated with scipy's Radau IIA (order 5, finite-difference Jacobian,
rtol 1e-12) over the workload's dt. These solutions pin the oracle's BDF integrator independently
of CVODE's step sequence: a BDF run at tolerance tau must land within O(tau) of them.

It also writes tests/golden/pmf_burned_<mech>.txt: the variable line and the last (burned-side)
row of the reference's premixed-flame profiles (Exec/RegTests/PMF/*.dat, data files of the
reference's regression tests), the state the partial-equilibrium test checks the thermo against.

Run from the repo root:  python tests/golden/make_golden.py [--pmf-only]
"""
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.kinetics_np import NpMech  # noqa: E402
from pelec_amd import inputs, mechanism as M  # noqa: E402

CASES = {
    # workload: cell indices (drm19_64 cell 3 is ignition-seeded; cell 4, at 2012 K, ignites within dt)
    "lidryer_32": [0, 1, 2, 3, 4, 5],
    "drm19_64": [0, 1, 2, 3, 4, 13, 39, 40],
}


def solve(wl_name, cells):
    wl = inputs.WORKLOADS[wl_name]
    mech = M.load(wl.mech)
    npm = NpMech(mech)
    N = mech.nsp
    d = inputs.make_inputs(wl_name, max(cells) + 1)
    out_y = []
    for c in cells:
        y0 = np.concatenate([d["rY"][:, c], [d["T"][c]]])
        rE = d["rE"][c]

        def f(t, y):
            return npm.rhs(t, y[None, :], rE, 0.0, np.zeros((1, N)))[0][0]

        sol = solve_ivp(f, (0.0, wl.dt), y0, method="Radau", rtol=1e-12, atol=1e-20)
        assert sol.status == 0, sol.message
        y = sol.y[:, -1]
        rho = y[:N].sum()
        T = npm.T_from_e(np.array(rE / rho), y[:N] / rho, np.array(y[N]))
        out_y.append(np.concatenate([y[:N], [float(T)]]))
        print(wl_name, c, "T0", d["T"][c], "T1", float(T), flush=True)
    return {
        "cells": np.array(cells), "rY0": d["rY"][:, cells], "T0": d["T"][cells], "rE0": d["rE"][cells],
        "rY1": np.array(out_y)[:, :N].T.copy(), "T1": np.array(out_y)[:, N].copy(), "dt": wl.dt,
        "mech_sha256": mech.source_sha256,
    }


PMF_FILES = {
    "drm19": "Exec/RegTests/PMF/PMF_CH4_1bar_300K_DRM_MixAvg.dat",
    "LiDryer": "Exec/RegTests/PMF/LiDryer_H2_p1_phi0_4000tu0300.dat",
    "dodecane_lu": "Exec/RegTests/PMF/PMF_NC12H26_1bar_300K_DodecaneLu.dat",
}


def pmf_rows(here, ref="/root/reference"):
    for mech, rel in PMF_FILES.items():
        lines = [ln for ln in open(os.path.join(ref, rel)).read().splitlines() if ln.strip()]
        with open(os.path.join(here, f"pmf_burned_{mech}.txt"), "w") as fh:
            fh.write(f"# {rel}: variable line and last row (x [cm], T [K], u [cm/s], rho [g/cm^3], "
                     "species mole fractions)\n")
            fh.write(lines[0].strip() + "\n")
            fh.write(lines[-1].strip() + "\n")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    pmf_rows(here)
    if "--pmf-only" in sys.argv:
        sys.exit(0)
    for wl_name, cells in CASES.items():
        res = solve(wl_name, cells)
        np.savez(os.path.join(here, f"radau_{wl_name}.npz"), **res)
