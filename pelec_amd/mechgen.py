"""Mechanism -> HIP device code generator.

Replaces the generated ``PelePhysics/Mechanisms/<model>/mechanism.{H,cpp}`` that the reference
compiles (CMake/BuildPelePhysicsLib.cmake:113-116; productionRate/CKWC, the Fuego thermo and the
analytic Jacobian DWDOT/aJacobian selected by PELE_COMPILE_AJACOBIAN, Exec/Make.PeleC:65-67).
The emitted header is written for gfx950 one-cell-per-lane execution:

* every reaction is straight-line code with its constants folded in (no table lookups, so the
  per-lane arrays stay in VGPRs after unrolling);
* reverse rates use per-species exp(g_k/RT), so a reversible reaction costs products and one
  division instead of an extra exp: 1/Kc = prod_k exp(g_k/RT)^nu_k * (P_atm/RT)^-dnu;
* thermo is emitted as folded polynomials in T of e_k, cv_k, dcv_k/dT (mass units) and as
  per-range mixture-polynomial coefficient sums, so T(e, Y) Newton iterations cost one degree-5
  polynomial per iteration;
* third-body concentrations are ctot + sum (alpha-1) C_k over the non-unit efficiencies only;
* the Jacobian of the production rates w.r.t. concentrations and T is accumulated per reaction
  into the caller's matrix (the dense third-body part as a row update per reacting species), so
  only the concentrations, exp(g/RT) table and rates stay live in registers.

The same header also carries the mechanism as data tables (T_* members: per-species thermo
coefficients, per-reaction Arrhenius/falloff/Troe/stoichiometry records, efficiency sets and a
species -> reaction list of net coefficients) for the cooperative kernels, where one group of lanes
integrates one cell and each lane runs one species or reaction through a single code path.

Output: pelec_amd/csrc/gen/mech_<name>.h, struct ``pelec::mech::<name>``.
"""
from __future__ import annotations

import math
import os
import sys
from collections import OrderedDict
from typing import Dict, List

from . import mechanism as M

GEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "gen")

LN10 = math.log(10.0)


def f(x: float) -> str:
    """C++ double literal, round-trip exact."""
    if x == 0.0:
        return "0.0"
    s = repr(float(x))
    if "e" not in s and "." not in s and "inf" not in s:
        s += ".0"
    return s


def _ident(name: str) -> str:
    out = []
    for ch in name:
        out.append(ch if ch.isalnum() else "_")
    s = "".join(out)
    return s if not s[0].isdigit() else "m_" + s


class Gen:
    def __init__(self, mech: M.Mechanism, chain: int = 0):
        self.m = mech
        self.chain = chain
        self.N = mech.nsp
        self.R = mech.nrxn
        self.lines: List[str] = []
        # distinct efficiency sets (as tuples) -> index
        self.effsets: "OrderedDict[tuple, int]" = OrderedDict()
        for r in mech.reactions:
            if r.efficiencies is not None:
                key = tuple(r.efficiencies)
                if key not in self.effsets:
                    self.effsets[key] = len(self.effsets)
        self.tmids = sorted({s.Tmid for s in mech.species})

    def w(self, s: str = ""):
        self.lines.append(s)

    # ------------------------------------------------------------------ helpers
    def tmid_flag(self, Tmid: float) -> str:
        return f"lo{self.tmids.index(Tmid)}"

    def rate_expr(self, A: float, b: float, EaR: float) -> str:
        """Arrhenius A T^b exp(-EaR/T) using lnT, invT, T in scope."""
        if A == 0.0:
            return "0.0"
        sgn = "-" if A < 0 else ""
        lnA = math.log(abs(A))
        if EaR == 0.0 and b == 0.0:
            return f(A)
        if EaR == 0.0 and b in (1.0, 2.0, 3.0):
            return f"({f(A)} * {'*'.join(['T'] * int(b))})"
        if EaR == 0.0 and b in (-1.0, -2.0):
            return f"({f(A)} * {'*'.join(['invT'] * int(-b))})"
        terms = [f(lnA)]
        if b != 0.0:
            terms.append(f"{f(b)} * lnT")
        if EaR != 0.0:
            terms.append(f"{f(-EaR)} * invT")
        return f"{sgn}exp({' + '.join(terms)})"

    def ln_rate_expr(self, A: float, b: float, EaR: float) -> str:
        return f"({f(math.log(A))} + {f(b)} * lnT + {f(-EaR)} * invT)"

    def third_body_expr(self, eff: List[float]) -> str:
        terms = ["ctot"]
        for k, a in enumerate(eff):
            if a != 1.0:
                d = a - 1.0
                terms.append(f"{f(d)} * C[{k}]")
        return " + ".join(terms)

    @staticmethod
    def prod_expr(side: Dict[int, float], arr: str = "C") -> str:
        terms = []
        for k, nu in sorted(side.items()):
            if nu == int(nu) and 1 <= nu <= 3:
                terms += [f"{arr}[{k}]"] * int(nu)
            else:
                terms.append(f"pow({arr}[{k}], {f(nu)})")
        return " * ".join(terms) if terms else "1.0"

    def dprod_terms(self, side: Dict[int, float]):
        """[(k, expr for d(prod)/dC_k)]"""
        out = []
        items = sorted(side.items())
        for idx, (k, nu) in enumerate(items):
            others = {kk: vv for kk, vv in items if kk != k}
            rest = self.prod_expr(others) if others else None
            if nu == 1.0:
                e = rest if rest else "1.0"
            elif nu == 2.0:
                e = f"2.0 * C[{k}]" + (f" * {rest}" if rest else "")
            elif nu == 3.0:
                e = f"3.0 * C[{k}] * C[{k}]" + (f" * {rest}" if rest else "")
            else:
                e = f"{f(nu)} * pow(C[{k}], {f(nu - 1.0)})" + (f" * {rest}" if rest else "")
            out.append((k, e))
        return out

    def inv_kc_expr(self, r: M.Reaction) -> str:
        """1/Kc = prod exp(g)^nu_prod * prod exp(-g)^nu_reac * refC^-dnu"""
        if not r.reversible:
            return "0.0"
        terms = []
        for k, nu in sorted(r.products.items()):
            if nu == int(nu):
                terms += [f"eg[{k}]"] * int(nu)
            else:
                terms.append(f"pow(eg[{k}], {f(nu)})")
        den = []
        for k, nu in sorted(r.reactants.items()):
            if nu == int(nu):
                den += [f"eg[{k}]"] * int(nu)
            else:
                den.append(f"pow(eg[{k}], {f(nu)})")
        dnu = sum(r.products.values()) - sum(r.reactants.values())
        if dnu == 1.0:
            terms.append("irefC")
        elif dnu == -1.0:
            terms.append("refC")
        elif dnu == 2.0:
            terms += ["irefC", "irefC"]
        elif dnu == -2.0:
            terms += ["refC", "refC"]
        elif dnu != 0.0:
            terms.append(f"pow(refC, {f(-dnu)})")
        num = " * ".join(terms) if terms else "1.0"
        return f"({num}) / ({' * '.join(den)})" if den else num

    @staticmethod
    def troe_const(troe: List[float]):
        """F_cent when it does not depend on T over 1..1e5 K (T*** -> 0, T* -> inf, no T**)."""
        a, T3, T1 = troe[:3]
        if len(troe) > 3 or T3 > 1e-20 or T1 < 1e20:
            return None
        return a  # (1-a)*exp(-T/T***) -> 0 and a*exp(-T/T*) -> a

    def nu_of(self, r: M.Reaction) -> Dict[int, float]:
        nu: Dict[int, float] = {}
        for k, v in r.reactants.items():
            nu[k] = nu.get(k, 0.0) - v
        for k, v in r.products.items():
            nu[k] = nu.get(k, 0.0) + v
        return {k: v for k, v in nu.items() if v != 0.0}

    # ------------------------------------------------------------------ thermo emitters
    def emit_thermo(self):
        N = self.N
        R = M.R_ERG
        self.w("  // ---- thermo: folded NASA-7 polynomials (mass units, erg/g, erg/(g K))")
        self.w("  __host__ __device__ static inline void tmid_flags(double T, bool* lo) {")
        for i, tm in enumerate(self.tmids):
            self.w(f"    lo[{i}] = T < {f(tm)};")
        self.w("  }")
        # e_k(T) = R/W [ (a0-1) T + a1/2 T^2 + a2/3 T^3 + a3/4 T^4 + a4/5 T^5 + a5 ]
        # cv_k(T) = R/W [ (a0-1) + a1 T + a2 T^2 + a3 T^3 + a4 T^4 ]
        self.ecoef = []
        self.cvcoef = []
        for s in self.m.species:
            rw = R / s.W
            ec = []
            cc = []
            for a in (s.low, s.high):
                ec.append([rw * a[5], rw * (a[0] - 1.0), rw * a[1] / 2, rw * a[2] / 3, rw * a[3] / 4, rw * a[4] / 5])
                cc.append([rw * (a[0] - 1.0), rw * a[1], rw * a[2], rw * a[3], rw * a[4]])
            self.ecoef.append(ec)
            self.cvcoef.append(cc)

        def horner(c: List[float]) -> str:
            expr = f(c[-1])
            for x in reversed(c[:-1]):
                expr = f"fma(T, {expr}, {f(x)})"
            return expr

        self.w("  // species e_k, cv_k (mass units) and dcv_k/dT")
        self.w("  __host__ __device__ static inline void e_cv(double T, double* e, double* cv) {")
        self.w(f"    bool lo[{len(self.tmids)}]; tmid_flags(T, lo);")
        for k, s in enumerate(self.m.species):
            fl = f"lo[{self.tmids.index(s.Tmid)}]"
            el, eh = self.ecoef[k]
            cl, ch = self.cvcoef[k]
            self.w(f"    e[{k}] = {fl} ? {horner(el)} : {horner(eh)};")
            self.w(f"    cv[{k}] = {fl} ? {horner(cl)} : {horner(ch)};")
        self.w("  }")
        self.w("  __host__ __device__ static inline void dcv_dT(double T, double* d) {")
        self.w(f"    bool lo[{len(self.tmids)}]; tmid_flags(T, lo);")
        for k, s in enumerate(self.m.species):
            fl = f"lo[{self.tmids.index(s.Tmid)}]"
            cl, ch = self.cvcoef[k]
            dl = [cl[1], 2 * cl[2], 3 * cl[3], 4 * cl[4]]
            dh = [ch[1], 2 * ch[2], 3 * ch[3], 4 * ch[4]]
            self.w(f"    d[{k}] = {fl} ? {horner(dl)} : {horner(dh)};")
        self.w("  }")
        # mixture polynomial coefficients: for each Tmid range set r and branch (lo/hi),
        # E_rb[i] = sum_{k in set r} Y_k ecoef_k[b][i]
        nt = len(self.tmids)
        self.w("  // mixture e(T) = sum Y_k e_k(T) as polynomial coefficients per Tmid group and range:")
        self.w(f"  // pe[(g*2 + b)*6 + i], g < {nt} Tmid groups, b = 0 low / 1 high")
        self.w(f"  static constexpr int NTMID = {nt};")
        self.w("  __host__ __device__ static inline void mix_e_poly(const double* Y, double* pe) {")
        for g in range(nt):
            for b in range(2):
                for i in range(6):
                    terms = []
                    for k, s in enumerate(self.m.species):
                        if s.Tmid != self.tmids[g]:
                            continue
                        c = self.ecoef[k][b][i]
                        if c != 0.0:
                            terms.append(f"{f(c)} * Y[{k}]")
                    expr = " + ".join(terms) if terms else "0.0"
                    self.w(f"    pe[{(g * 2 + b) * 6 + i}] = {expr};")
        self.w("  }")
        self.w("  __host__ __device__ static inline double tmid(int g) {")
        for g, tm in enumerate(self.tmids):
            self.w(f"    if (g == {g}) return {f(tm)};")
        self.w("    return 0.0;")
        self.w("  }")
        # g/RT and h/RT per species
        self.w("  // g_k/RT and h_k/RT")
        self.w("  __host__ __device__ static inline void g_h_RT(double T, double lnT, double invT, double* gRT, double* hRT) {")
        self.w(f"    bool lo[{len(self.tmids)}]; tmid_flags(T, lo);")
        for k, s in enumerate(self.m.species):
            fl = f"lo[{self.tmids.index(s.Tmid)}]"
            exprs_g = []
            exprs_h = []
            for a in (s.low, s.high):
                # g/RT = a0(1-lnT) - a1 T/2 - a2 T^2/6 - a3 T^3/12 - a4 T^4/20 + a5/T - a6
                poly = [a[0] - a[6], -a[1] / 2, -a[2] / 6, -a[3] / 12, -a[4] / 20]
                exprs_g.append(f"({horner(poly)} - {f(a[0])} * lnT + {f(a[5])} * invT)")
                hp = [a[0], a[1] / 2, a[2] / 3, a[3] / 4, a[4] / 5]
                exprs_h.append(f"({horner(hp)} + {f(a[5])} * invT)")
            self.w(f"    gRT[{k}] = {fl} ? {exprs_g[0]} : {exprs_g[1]};")
            self.w(f"    if (hRT) hRT[{k}] = {fl} ? {exprs_h[0]} : {exprs_h[1]};")
        self.w("  }")

    # ------------------------------------------------------------------ kinetics emitters
    def emit_reaction_rate(self, j: int, r: M.Reaction, jac: bool):
        """Emits code computing kf (effective, incl. M / falloff), invKc, net, q for reaction j.
        In jac mode also dkdM (for third-body/falloff) and dlnk (d ln keff/dT)."""
        w = self.w
        w(f"    {{  // R{j}: {r.equation}")
        if r.kind == "falloff":
            w(f"      const double kinf = {self.rate_expr(r.A, r.b, r.Ea_R)};")
            A0, b0, E0 = r.low
            lnratio = f"{f(math.log(A0) - math.log(r.A))} + {f(b0 - r.b)} * lnT + {f(-(E0 - r.Ea_R))} * invT"
            w(f"      const double Mj = M{self.effsets[tuple(r.efficiencies)]};")
            w(f"      const double k0r = exp({lnratio});")
            w(f"      const double Pr = fmax(k0r * Mj, 1e-300);")
            w(f"      const double iPr1 = 1.0 / (1.0 + Pr);")
            if r.troe is not None and self.troe_const(r.troe) is not None:
                fcv = self.troe_const(r.troe)
                lfc = math.log10(fcv)
                w(f"      const double lFc = {f(lfc)};")
                w(f"      const double cc = {f(-0.4 - 0.67 * lfc)}, nn = {f(0.75 - 1.27 * lfc)};")
                w("      const double xx = log10(Pr) + cc;")
                w("      const double DD = 1.0 / (nn - 0.14 * xx);")
                w("      const double f1 = xx * DD;")
                w("      const double iden = 1.0 / (1.0 + f1 * f1);")
                w(f"      const double F = exp({f(LN10 * lfc)} * iden);")
                if jac:
                    w("      const double g = -lFc * 2.0 * f1 * nn * DD * DD * iden * iden;")
                    w("      const double dlnF_T = 0.0;")
            elif r.troe is not None:
                a, T3, T1 = r.troe[:3]
                fc_terms = []
                dfc_terms = []
                if 1.0 - a != 0.0:
                    fc_terms.append(f"{f(1.0 - a)} * exp({f(-1.0 / T3)} * T)")
                    dfc_terms.append(f"{f(-(1.0 - a) / T3)} * exp({f(-1.0 / T3)} * T)")
                if a != 0.0:
                    fc_terms.append(f"{f(a)} * exp({f(-1.0 / T1)} * T)")
                    dfc_terms.append(f"{f(-a / T1)} * exp({f(-1.0 / T1)} * T)")
                if len(r.troe) > 3:
                    T2 = r.troe[3]
                    fc_terms.append(f"exp({f(-T2)} * invT)")
                    dfc_terms.append(f"{f(T2)} * invT * invT * exp({f(-T2)} * invT)")
                w(f"      const double Fc = {' + '.join(fc_terms)};")
                w("      const double lFc = log10(Fc);")
                w("      const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;")
                w("      const double xx = log10(Pr) + cc;")
                w("      const double DD = 1.0 / (nn - 0.14 * xx);")
                w("      const double f1 = xx * DD;")
                w("      const double iden = 1.0 / (1.0 + f1 * f1);")
                w(f"      const double F = exp({f(LN10)} * lFc * iden);")
                if jac:
                    w("      const double g = -lFc * 2.0 * f1 * nn * DD * DD * iden * iden;")
                    w("      const double df1 = (-0.67 * nn + 1.27 * xx) * DD * DD;")
                    w("      const double dlF = iden - lFc * 2.0 * f1 * df1 * iden * iden;")
                    w(f"      const double dFc = {' + '.join(dfc_terms)};")
                    w("      const double dlnF_T = dlF * dFc / Fc;")
            else:
                w("      const double F = 1.0;")
                if jac:
                    w("      const double g = 0.0, dlnF_T = 0.0;")
            w(f"      const double kf = RMUL(kinf * Pr * iPr1 * F, {j});")
            if jac:
                w(f"      const double dlnkinf = ({f(r.b)} + {f(r.Ea_R)} * invT) * invT;")
                w(f"      const double dlnPr = ({f(b0 - r.b)} + {f(E0 - r.Ea_R)} * invT) * invT;")
                w(f"      const double dkdM = RMUL(kinf * k0r * F * (iPr1 * iPr1 + g * iPr1), {j});")
                w("      const double dlnk = dlnkinf + dlnPr * iPr1 + g * dlnPr + dlnF_T;")
        elif r.kind == "three-body":
            w(f"      const double kb = RMUL({self.rate_expr(r.A, r.b, r.Ea_R)}, {j});")
            w(f"      const double kf = kb * M{self.effsets[tuple(r.efficiencies)]};")
            if jac:
                w("      const double dkdM = kb;")
                w(f"      const double dlnk = ({f(r.b)} + {f(r.Ea_R)} * invT) * invT;")
        else:
            w(f"      const double kf = RMUL({self.rate_expr(r.A, r.b, r.Ea_R)}, {j});")
            if jac:
                w(f"      const double dlnk = ({f(r.b)} + {f(r.Ea_R)} * invT) * invT;")
        w(f"      const double Pf = {self.prod_expr(r.reactants)};")
        if r.reversible:
            w(f"      const double iKc = {self.inv_kc_expr(r)};")
            w(f"      const double Pb = {self.prod_expr(r.products)};")
            w("      const double net = Pf - iKc * Pb;")
        else:
            w("      const double net = Pf;")
        w("      const double qj = kf * net;")

    def emit_chain(self, j: int, dep: str):
        """After every CHAIN reactions, pass T, lnT and invT through an empty asm that also reads
        the last rate: the next group's rate constants depend on the asm's outputs, so the
        compiler cannot hoist all exponentials to the top (which costs one register pair each
        and spills). Bounds the live set to one group of reactions."""
        if self.chain > 0 and (j + 1) % self.chain == 0 and j + 1 < self.R:
            self.w(f"      PELEC_REG_CHAIN(T, lnT, invT, {dep});")

    def emit_rates(self):
        N, R = self.N, self.R
        w = self.w
        w("  // common T-dependent terms and per-species exp(g/RT) for reverse rates")
        w("  __host__ __device__ static inline void prelude(double T, double lnT, double invT, double* eg, double* hRT) {")
        w(f"    double gRT[{N}];")
        w("    g_h_RT(T, lnT, invT, gRT, hRT);")
        w("#pragma unroll")
        w(f"    for (int k = 0; k < {N}; ++k) eg[k] = exp(gRT[k]);")
        w("  }")
        # third-body sums
        def emit_M():
            w("    double ctot = 0.0;")
            w("#pragma unroll")
            w(f"    for (int k = 0; k < {N}; ++k) ctot += C[k];")
            for key, idx in self.effsets.items():
                w(f"    const double M{idx} = {self.third_body_expr(list(key))};")

        # rates of progress
        w("  // rates of progress q_j [mol/(cm^3 s)]; scale (may be null) multiplies each q_j")
        w("  template <bool MUL = false>")
        w("  __host__ __device__ static inline void progress(const double* C, double T, double* q, const double* rmul = nullptr) {")
        w("    RMUL_DEF;")
        w("    const double lnT = log(T), invT = 1.0 / T;")
        w(f"    const double refC = {f(M.P_ATM / M.R_ERG)} * invT, irefC = {f(M.R_ERG / M.P_ATM)} * T;")
        w("    (void)refC; (void)irefC;")
        w(f"    double eg[{N}];")
        w("    prelude(T, lnT, invT, eg, nullptr);")
        emit_M()
        for j, r in enumerate(self.m.reactions):
            self.emit_reaction_rate(j, r, False)
            w(f"      q[{j}] = qj;")
            w("    }")
        w("  }")
        # production rates
        w("  // molar production rates wdot_k [mol/(cm^3 s)]")
        w("  template <bool MUL = false>")
        w("  __host__ __device__ static inline void wdot(const double* C, double T, double* wd, const double* rmul = nullptr) {")
        w("    RMUL_DEF;")
        w("    double lnT = log(T), invT = 1.0 / T;")
        w(f"    const double refC = {f(M.P_ATM / M.R_ERG)} * invT, irefC = {f(M.R_ERG / M.P_ATM)} * T;")
        w("    (void)refC; (void)irefC;")
        w(f"    double eg[{N}];")
        w("    prelude(T, lnT, invT, eg, nullptr);")
        emit_M()
        w("#pragma unroll")
        w(f"    for (int k = 0; k < {N}; ++k) wd[k] = 0.0;")
        for j, r in enumerate(self.m.reactions):
            self.emit_reaction_rate(j, r, False)
            for k, nu in sorted(self.nu_of(r).items()):
                w(f"      wd[{k}] += {f(nu)} * qj;")
            self.emit_chain(j, "qj")
            w("    }")
        w("  }")
        # Jacobian
        ne = len(self.effsets)
        w("  // wdot and its Jacobian: dwdC[k*LD+m] = d wdot_k / d C_m, dwdT[k] = d wdot_k / dT")
        w("  template <int LD, typename JT, bool MUL = false>")
        w("  __host__ __device__ static inline void wdot_jac(const double* C, double T, double* wd, JT* dwdC, double* dwdT, const double* rmul = nullptr) {")
        w("    RMUL_DEF;")
        w("    double lnT = log(T), invT = 1.0 / T;")
        w(f"    const double refC = {f(M.P_ATM / M.R_ERG)} * invT, irefC = {f(M.R_ERG / M.P_ATM)} * T;")
        w("    (void)refC; (void)irefC;")
        w(f"    double eg[{N}], hRT[{N}];")
        w("    prelude(T, lnT, invT, eg, hRT);")
        emit_M()
        w("#pragma unroll")
        w(f"    for (int k = 0; k < {N}; ++k) {{ wd[k] = 0.0; dwdT[k] = 0.0; }}")
        w("#pragma unroll")
        w(f"    for (int k = 0; k < {N}; ++k)")
        w("#pragma unroll")
        w(f"      for (int m = 0; m < {N}; ++m) dwdC[k * LD + m] = JT(0);")
        for j, r in enumerate(self.m.reactions):
            self.emit_reaction_rate(j, r, True)
            nu = self.nu_of(r)
            # dlnKc/dT = (sum nu hRT - dnu)/T
            if r.reversible:
                dnu = sum(r.products.values()) - sum(r.reactants.values())
                sh = " + ".join(f"{f(v)} * hRT[{k}]" for k, v in sorted(nu.items())) or "0.0"
                w(f"      const double dlnKc = ({sh} - {f(dnu)}) * invT;")
                w("      const double dqdT = kf * (dlnk * net + iKc * Pb * dlnKc);")
            else:
                w("      const double dqdT = kf * dlnk * net;")
            # mass-action derivatives
            dq: Dict[int, List[str]] = {}
            for k, e in self.dprod_terms(r.reactants):
                dq.setdefault(k, []).append(f"kf * ({e})")
            if r.reversible:
                for k, e in self.dprod_terms(r.products):
                    dq.setdefault(k, []).append(f"-kf * iKc * ({e})")
            for m, exprs in sorted(dq.items()):
                w(f"      const double dq{m} = {' + '.join(exprs)};")
            for k, v in sorted(nu.items()):
                w(f"      wd[{k}] += {f(v)} * qj;")
                w(f"      dwdT[{k}] += {f(v)} * dqdT;")
                for m in sorted(dq):
                    w(f"      dwdC[{k} * LD + {m}] += JT({f(v)} * dq{m});")
            if r.kind in ("three-body", "falloff"):
                w("      const double dqdM = dkdM * net;")
                eff = r.efficiencies
                for k, v in sorted(nu.items()):
                    for m in range(self.N):
                        if eff[m] != 0.0:
                            w(f"      dwdC[{k} * LD + {m}] += JT({f(v * eff[m])} * dqdM);")
            # one reaction's updates reach memory before the next reaction starts: keeps the
            # matrix out of registers (it is accumulated in the caller's storage)
            w("      PELEC_MEMORY_FENCE();")
            self.emit_chain(j, "qj")
            w("    }")
        w("  }")
        # transpose of the net stoichiometric matrix applied to a species vector
        w("  // out[j] = sum_k nu_kj w[k] (net stoichiometric coefficients, products minus reactants)")
        w("  __host__ __device__ static inline void nu_dot(const double* w, double* out) {")
        for j, r in enumerate(self.m.reactions):
            terms = [f"{f(v)} * w[{k}]" for k, v in sorted(self.nu_of(r).items()) if v != 0.0]
            w(f"    out[{j}] = {' + '.join(terms) if terms else '0.0'};")
        w("  }")

    # ------------------------------------------------------------------ data tables
    def table_form_issue(self):
        """Why the data-table form cannot represent this mechanism (None when it can): the
        cooperative kernels then are not built for it, and the one-cell-per-lane kernels run."""
        for r in self.m.reactions:
            if r.A <= 0.0:
                return f"{r.equation}: A <= 0"
            # the product side enters the rate only through the reverse rate: an irreversible
            # reaction may have any products (its product slots stay empty)
            for side in ((r.reactants, r.products) if r.reversible else (r.reactants,)):
                if len(side) > 3 or any(v != int(v) or not 1 <= v <= 3 for v in side.values()):
                    return f"{r.equation}: more than 3 species or a non-integer order on one side"
            d = sum(r.products.values()) - sum(r.reactants.values())
            if r.reversible and (d != int(d) or abs(d) > 2):
                return f"{r.equation}: |dnu| > 2"
            if len(self.nu_of(r)) > 6:
                return f"{r.equation}: more than 6 species change"
        if len(self.effsets) > 32:
            return "more than 32 efficiency sets"
        return None

    def emit_tables(self):
        """The mechanism as data, for the cooperative (one lane group per cell) kernels, which run
        one code path for every species / reaction lane instead of the straight-line code above.
        Coefficients are the same Python values the straight-line emitters print."""
        issue = self.table_form_issue()
        if issue is not None:
            self.w(f"  // no data tables: {issue}")
            self.w("  static constexpr bool T_OK = false;")
            return
        self.w("  static constexpr bool T_OK = true;")
        N, R = self.N, self.R
        w = self.w
        sp = self.m.species
        rx = self.m.reactions

        def arr(name, ctype, vals, shape):
            dims = "".join(f"[{d}]" for d in shape)

            def fmt(v):
                if isinstance(v, list):
                    return "{" + ", ".join(fmt(x) for x in v) + "}"
                return f(v) if ctype == "double" else str(int(v))
            w(f"  static constexpr {ctype} {name}{dims} = {fmt(vals)};")

        w("  // ---- data tables (cooperative kernels)")
        w(f"  static constexpr double T_REFC = {f(M.P_ATM / M.R_ERG)};  // P_atm / R, times 1/T")
        w(f"  static constexpr double T_IREFC = {f(M.R_ERG / M.P_ATM)};")
        w(f"  static constexpr double T_LN10 = {f(LN10)};")
        arr("T_TMID", "double", [s.Tmid for s in sp], [N])
        arr("T_TMIDG", "int", [self.tmids.index(s.Tmid) for s in sp], [N])
        gco, hco = [], []
        for s in sp:
            gb, hb = [], []
            for a in (s.low, s.high):
                gb.append([a[0] - a[6], -a[1] / 2, -a[2] / 6, -a[3] / 12, -a[4] / 20, a[0], a[5]])
                hb.append([a[0], a[1] / 2, a[2] / 3, a[3] / 4, a[4] / 5, a[5]])
            gco.append(gb)
            hco.append(hb)
        arr("T_G", "double", gco, [N, 2, 7])
        arr("T_H", "double", hco, [N, 2, 6])
        arr("T_E", "double", [self.ecoef[k] for k in range(N)], [N, 2, 6])
        arr("T_CV", "double", [self.cvcoef[k] for k in range(N)], [N, 2, 5])
        arr("T_DCV", "double", [[[c[1], 2 * c[2], 3 * c[3], 4 * c[4]] for c in self.cvcoef[k]]
                                for k in range(N)], [N, 2, 4])
        # efficiency sets: d = eff - 1 (third-body sums) and eff (Jacobian)
        ns = max(1, len(self.effsets))
        effd = [[0.0] * N for _ in range(ns)]
        effv = [[0.0] * N for _ in range(ns)]
        for key, idx in self.effsets.items():
            effd[idx] = [a - 1.0 for a in key]
            effv[idx] = list(key)
        w(f"  static constexpr int T_NSETS = {len(self.effsets)};")
        arr("T_EFFD", "double", effd, [ns, N])
        arr("T_EFFV", "double", effv, [ns, N])
        # reactions: lane order puts falloff, then three-body reactions first, so the rounds of
        # the reaction phase after the first are uniform elementary reactions
        kind_id = {"elementary": 0, "three-body": 1, "falloff": 2}
        order = sorted(range(R), key=lambda j: (-kind_id[rx[j].kind], j))
        arr("T_RORD", "int", order, [R])
        kind, rev, kconst, A, lnA, b, EaR, low, tk, troe, eff = ([] for _ in range(11))
        rsp, rnu, psp, pnu, dnu, nsp_, nnu = ([] for _ in range(7))
        for r in rx:
            if r.A <= 0.0:
                raise ValueError(f"{r.equation}: table form needs A > 0")
            kind.append(kind_id[r.kind])
            rev.append(1 if r.reversible else 0)
            kconst.append(1 if (r.Ea_R == 0.0 and r.b == 0.0) else 0)
            A.append(r.A)
            lnA.append(math.log(r.A))
            b.append(r.b)
            EaR.append(r.Ea_R)
            if r.kind == "falloff":
                A0, b0, E0 = r.low
                low.append([math.log(A0) - math.log(r.A), b0 - r.b, E0 - r.Ea_R])
            else:
                low.append([0.0, 0.0, 0.0])
            if r.kind == "falloff" and r.troe is not None and self.troe_const(r.troe) is not None:
                lfc = math.log10(self.troe_const(r.troe))
                tk.append(1)
                troe.append([lfc, -0.4 - 0.67 * lfc, 0.75 - 1.27 * lfc, LN10 * lfc, 0.0, 0.0, 0.0, 0.0])
            elif r.kind == "falloff" and r.troe is not None:
                a, T3, T1 = r.troe[:3]
                T2 = r.troe[3] if len(r.troe) > 3 else 0.0
                tk.append(2)
                troe.append([1.0 - a, -1.0 / T3, a, -1.0 / T1, T2, 1.0 if len(r.troe) > 3 else 0.0,
                             -(1.0 - a) / T3, -a / T1])
            else:
                tk.append(0)
                troe.append([0.0] * 8)
            eff.append(self.effsets[tuple(r.efficiencies)] if r.efficiencies is not None else -1)
            for side, si, ni in ((r.reactants, rsp, rnu), (r.products, psp, pnu)):
                items = sorted(side.items()) if (side is r.reactants or r.reversible) else []
                if len(items) > 3 or any(v != int(v) or not 1 <= v <= 3 for _, v in items):
                    raise ValueError(f"{r.equation}: table form needs <= 3 species, nu in 1..3")
                si.append([k for k, _ in items] + [-1] * (3 - len(items)))
                ni.append([int(v) for _, v in items] + [0] * (3 - len(items)))
            d = sum(r.products.values()) - sum(r.reactants.values()) if r.reversible else 0
            if d != int(d) or abs(d) > 2:
                raise ValueError(f"{r.equation}: table form needs |dnu| <= 2")
            dnu.append(int(d))
            net = sorted(self.nu_of(r).items())
            nsp_.append([k for k, _ in net] + [-1] * (6 - len(net)))
            nnu.append([v for _, v in net] + [0.0] * (6 - len(net)))
        arr("T_KIND", "int", kind, [R])
        arr("T_REV", "int", rev, [R])
        arr("T_KCONST", "int", kconst, [R])
        arr("T_A", "double", A, [R])
        arr("T_LNA", "double", lnA, [R])
        arr("T_B", "double", b, [R])
        arr("T_EAR", "double", EaR, [R])
        arr("T_LOW", "double", low, [R, 3])
        arr("T_TROEK", "int", tk, [R])
        arr("T_TROE", "double", troe, [R, 8])
        arr("T_EFF", "int", eff, [R])
        arr("T_RSP", "int", rsp, [R, 3])
        arr("T_RNU", "int", rnu, [R, 3])
        arr("T_PSP", "int", psp, [R, 3])
        arr("T_PNU", "int", pnu, [R, 3])
        arr("T_DNU", "int", dnu, [R])
        arr("T_NSPC", "int", nsp_, [R, 6])
        arr("T_NNU", "double", nnu, [R, 6])
        # species -> (reaction, net nu) in reaction order (the order wdot accumulates in)
        rows = [[] for _ in range(N)]
        for j, r in enumerate(rx):
            for k, v in sorted(self.nu_of(r).items()):
                rows[k].append((j, v))
        ptr = [0]
        for rr in rows:
            ptr.append(ptr[-1] + len(rr))
        flat = [e for rr in rows for e in rr]
        w(f"  static constexpr int T_CSR_MAX = {max(len(rr) for rr in rows)};")
        arr("T_CSR_PTR", "int", ptr, [N + 1])
        # per entry also the reaction's six mass-action slot species and efficiency set, so a
        # lane's gather loads depend only on the entry index; PAD entries past the end keep
        # batched loads in bounds
        pad = 8
        w(f"  static constexpr int T_CSR_PAD = {pad};")
        arr("T_CSR_J", "int", [j for j, _ in flat] + [0] * pad, [len(flat) + pad])
        arr("T_CSR_NU", "double", [v for _, v in flat] + [0.0] * pad, [len(flat) + pad])
        slots = [rsp[j] + psp[j] for j, _ in flat] + [[-1] * 6] * pad
        arr("T_CSR_SLOT", "int", slots, [len(flat) + pad, 6])
        arr("T_CSR_EFF", "int", [eff[j] for j, _ in flat] + [-1] * pad, [len(flat) + pad])

    # ------------------------------------------------------------------ top level
    def generate(self) -> str:
        m = self.m
        N = self.N
        ident = _ident(m.name)
        w = self.w
        w("// GENERATED by pelec_amd/mechgen.py -- do not edit.")
        w(f"// mechanism: {m.name}  source sha256: {m.source_sha256}")
        w(f"// species: {N}  reactions: {self.R}")
        w("#pragma once")
        w("#include <cmath>")
        w("#ifndef PELEC_MEMORY_FENCE")
        w("#define PELEC_MEMORY_FENCE() __asm__ __volatile__(\"\" ::: \"memory\")")
        w("#endif")
        w("#ifndef PELEC_REG_CHAIN")
        w("#if defined(__HIP_DEVICE_COMPILE__)")
        w("#define PELEC_REG_CHAIN(a, b, c, d) __asm__ __volatile__(\"\" : \"+v\"(a), \"+v\"(b), \"+v\"(c) : \"v\"(d))")
        w("#else")
        w("#define PELEC_REG_CHAIN(a, b, c, d) ((void)(d))")
        w("#endif")
        w("#endif")
        w("// RMUL(k, j): the rate constant of reaction j times its multiplier rmul[j] when MUL (the")
        w("// per-reaction pre-exponential scaling exp(log A_j) of the sensitivity and adjoint paths)")
        w("#ifndef RMUL")
        w("#define RMUL_DEF (void)rmul")
        w("#define RMUL(k, j) (MUL ? (k) * rmul[j] : (k))")
        w("#endif")
        w("#ifndef __HIPCC__")
        w("#define __host__")
        w("#define __device__")
        w("#endif")
        w("namespace pelec {")
        w("namespace mech {")
        w(f"struct {ident} {{")
        w(f"  static constexpr int NSP = {N};")
        w(f"  static constexpr int NRXN = {self.R};")
        w(f"  static constexpr const char* NAME = \"{m.name}\";")
        w(f"  static constexpr const char* SHA256 = \"{m.source_sha256}\";")
        w(f"  static constexpr double W[{N}] = {{{', '.join(f(s.W) for s in m.species)}}};")
        w(f"  static constexpr double IW[{N}] = {{{', '.join(f(1.0 / s.W) for s in m.species)}}};")
        names = ", ".join(f'"{s.name}"' for s in m.species)
        w(f"  static constexpr const char* SPECIES[{N}] = {{{names}}};")
        self.emit_thermo()
        self.emit_rates()
        self.emit_tables()
        w("};")
        w("}  // namespace mech")
        w("}  // namespace pelec")
        return "\n".join(self.lines) + "\n"


def generate(name: str, out_dir: str = GEN_DIR) -> str:
    mech = M.load(name)
    code = Gen(mech).generate()
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"mech_{mech.name}.h")
    old = None
    if os.path.exists(path):
        with open(path) as fh:
            old = fh.read()
    if old != code:
        with open(path, "w") as fh:
            fh.write(code)
    return path


if __name__ == "__main__":
    for nm in sys.argv[1:] or M.available():
        print(generate(nm))
