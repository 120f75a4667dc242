"""Oracle (test infrastructure only): numpy restatement of the constant-volume rho*e reactor RHS.

This is a checker, never a product path. It restates, vectorised over cells, what the reference
path computes per cell (the RHS of ReactorCvode/ReactorRK64 for reactor type 1, listed at
CMake/BuildPelePhysicsLib.cmake:118-142; the bodies live in the absent PelePhysics submodule):

  rho = sum_k rhoY_k,  Y = rhoY/rho,  e = (rhoe0 + t*rhoe_src)/rho          (React.cpp:212-219,292)
  T   = T(e, Y) by Newton on sum_k Y_k e_k(T) = e                            (EOS REY2T, Utilities.H:81-82)
  d(rhoY_k)/dt = W_k wdot_k + rhoY_src_k
  dT/dt        = (rhoe_src - sum_k e_k d(rhoY_k)/dt) / (rho cv)

with wdot from mass-action kinetics (Arrhenius, third-body, Lindemann/Troe falloff, reverse rates
from NASA-7 equilibrium constants). Energy uses e = sum Y_k e_k including formation energy
(Docs/sphinx/Algorithms.rst:519-525, Equations.rst:36-37).

Pinned by: tests/test_oracle_golden.py (Cantera trajectories of Exec/RegTests/zeroD/cantera_soln and
Exec/RegTests/EB-C3/cantera_soln, constant-volume equilibrium end points), detailed balance at
equilibrium, mass/element conservation, and Jacobian checks against torch autograd / finite
differences.
"""
from __future__ import annotations

import numpy as np

R_ERG = 8.31446261815324e7
P_ATM = 1013250.0


class NpMech:
    def __init__(self, mech):
        self.mech = mech
        N = mech.nsp
        R = mech.nrxn
        self.N, self.R = N, R
        self.W = np.array(mech.weights)
        self.low = np.array([s.low for s in mech.species])
        self.high = np.array([s.high for s in mech.species])
        self.Tmid = np.array([s.Tmid for s in mech.species])
        self.nuf = np.zeros((R, N))
        self.nur = np.zeros((R, N))
        for j, r in enumerate(mech.reactions):
            for k, v in r.reactants.items():
                self.nuf[j, k] += v
            for k, v in r.products.items():
                self.nur[j, k] += v
        self.nu = self.nur - self.nuf
        self.A = np.array([r.A for r in mech.reactions])
        self.b = np.array([r.b for r in mech.reactions])
        self.EaR = np.array([r.Ea_R for r in mech.reactions])
        self.rev = np.array([r.reversible for r in mech.reactions])
        self.kind = [r.kind for r in mech.reactions]
        self.eff = np.array([r.efficiencies if r.efficiencies else [0.0] * N
                             for r in mech.reactions])

    # ---- thermo (NASA-7) -------------------------------------------------------------
    def _coef(self, T):
        T = np.asarray(T, dtype=float)[..., None]
        return np.where((T < self.Tmid)[..., None], self.low, self.high), T

    def cp_R(self, T):
        a, t = self._coef(T)   # t: (..., 1) broadcasts against the species axis
        return a[..., 0] + t * (a[..., 1] + t * (a[..., 2] + t * (a[..., 3] + t * a[..., 4])))

    def h_RT(self, T):
        a, t = self._coef(T)
        return (a[..., 0] + t * (a[..., 1] / 2 + t * (a[..., 2] / 3 + t * (a[..., 3] / 4 + t * a[..., 4] / 5)))
                + a[..., 5] / t)

    def s_R(self, T):
        a, t = self._coef(T)
        return (a[..., 0] * np.log(t) + t * (a[..., 1] + t * (a[..., 2] / 2 + t * (a[..., 3] / 3 + t * a[..., 4] / 4)))
                + a[..., 6])

    def e_mass(self, T):
        T = np.asarray(T, dtype=float)
        return (self.h_RT(T) - 1.0) * R_ERG * T[..., None] / self.W

    def cv_mass(self, T):
        return (self.cp_R(T) - 1.0) * R_ERG / self.W

    def T_from_e(self, e, Y, T0, tol=1e-12, maxit=50):
        T = np.array(T0, dtype=float, copy=True)
        for _ in range(maxit):
            f = (Y * self.e_mass(T)).sum(-1) - e
            cv = (Y * self.cv_mass(T)).sum(-1)
            dT = -f / cv
            T = T + dT
            if np.all(np.abs(dT) <= tol * np.abs(T)):
                break
        return T

    def mixture(self, p, T, Y):
        Wbar = 1.0 / (Y / self.W).sum(-1)
        rho = p * Wbar / (R_ERG * T)
        e = (Y * self.e_mass(T)).sum(-1)
        return rho, e

    # ---- kinetics ----------------------------------------------------------------------
    def rates_of_progress(self, C, T):
        T = np.asarray(T, dtype=float)
        Tc = T[..., None]
        lnT = np.log(Tc)
        kf = self.A * np.exp(self.b * lnT - self.EaR / Tc)
        g_RT = self.h_RT(T) - self.s_R(T)
        dG = g_RT @ self.nu.T
        dn = self.nu.sum(1)
        Kc = np.exp(-dG) * (P_ATM / (R_ERG * Tc)) ** dn
        M = C @ self.eff.T
        for j, kind in enumerate(self.kind):
            if kind == "three-body":
                kf[..., j] *= M[..., j]
            elif kind == "falloff":
                r = self.mech.reactions[j]
                A0, b0, E0 = r.low
                k0 = A0 * np.exp(b0 * lnT[..., 0] - E0 / T)
                kinf = kf[..., j]
                Pr = k0 * M[..., j] / kinf
                F = np.ones_like(Pr)
                if r.troe is not None:
                    a, T3, T1 = r.troe[:3]
                    Fc = (1 - a) * np.exp(-T / T3) + a * np.exp(-T / T1)
                    if len(r.troe) > 3:
                        Fc = Fc + np.exp(-r.troe[3] / T)
                    lFc = np.log10(Fc)
                    c = -0.4 - 0.67 * lFc
                    n = 0.75 - 1.27 * lFc
                    lPr = np.log10(np.maximum(Pr, 1e-300))
                    f1 = (lPr + c) / (n - 0.14 * (lPr + c))
                    F = 10.0 ** (lFc / (1.0 + f1 * f1))
                kf[..., j] = kinf * Pr / (1.0 + Pr) * F
        kr = np.where(self.rev, kf / Kc, 0.0)
        prod_f = np.prod(C[..., None, :] ** self.nuf, axis=-1)
        prod_r = np.prod(C[..., None, :] ** self.nur, axis=-1)
        return kf * prod_f - kr * prod_r

    def wdot(self, C, T):
        return self.rates_of_progress(C, T) @ self.nu

    # ---- reactor RHS ---------------------------------------------------------------------
    def rhs(self, t, y, rhoe0, rhoe_src, rY_src, T_guess=None):
        """y = [rhoY_0..rhoY_{N-1}, T] (cells, N+1). Returns dy/dt and the T used."""
        N = self.N
        rY = y[..., :N]
        rho = rY.sum(-1)
        Y = rY / rho[..., None]
        e = (rhoe0 + t * rhoe_src) / rho
        T = self.T_from_e(e, Y, y[..., N] if T_guess is None else T_guess)
        C = rY / self.W
        f = self.W * self.wdot(C, T) + rY_src
        ek = self.e_mass(T)
        cv = (Y * self.cv_mass(T)).sum(-1)
        fT = (rhoe_src - (ek * f).sum(-1)) / (rho * cv)
        return np.concatenate([f, fT[..., None]], axis=-1), T
