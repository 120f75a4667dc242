"""Oracle (test infrastructure only): ctypes binding of oracle/_build/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker. The product library (pelec_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
#: the same source built with FMA contraction (g++ -mfma -ffp-contract=fast): a second FP64
#: evaluation of the same algorithm, whose distance from the first measures how far rounding
#: alone moves an integration (its step decisions) on given inputs
FMA_LIB_PATH = os.path.join(HERE, "_build", "liboracle_fma.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH) or not os.path.exists(FMA_LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_libs = {}


def lib(variant: str = ""):
    """The oracle library: variant "" (no contraction) or "fma"."""
    if variant not in _libs:
        path = FMA_LIB_PATH if variant == "fma" else LIB_PATH
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_mech_new.restype = ctypes.c_void_p
        L.orc_mech_new.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 4 + [_ip, _ip] + [_dp] * 8 + [_ip]
        L.orc_mech_free.argtypes = [ctypes.c_void_p]
        L.orc_thermo.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp, _dp]
        L.orc_T_from_e.restype = ctypes.c_double
        L.orc_T_from_e.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_double, _ip]
        L.orc_progress.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_wdot.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.orc_rhs.restype = ctypes.c_double
        L.orc_rhs.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_jac.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_getrf.restype = ctypes.c_int
        L.orc_getrf.argtypes = [ctypes.c_int, _dp, _ip]
        L.orc_getrs.argtypes = [ctypes.c_int, _dp, _ip, _dp]
        L.orc_react.restype = ctypes.c_int
        L.orc_react.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_react_rk64.restype = ctypes.c_int
        L.orc_react_rk64.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                     _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _ip, _ip]
        L.orc_replay.restype = ctypes.c_int
        L.orc_replay.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                 ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_double,
                                 _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        _libs[variant] = L
    return _libs[variant]


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


STAT_NAMES = ("nst", "nfe", "nje", "nsetups", "nni", "netf", "ncfn")


class OracleMech:
    """CPU FP64 oracle of one mechanism (tables built from pelec_amd.mechanism data)."""

    def __init__(self, mech, variant: str = ""):
        self.mech = mech
        self.L = lib(variant)
        N, R = mech.nsp, mech.nrxn
        self.N, self.R = N, R
        self.W = np.ascontiguousarray(mech.weights, dtype=np.float64)
        Tmid = np.array([s.Tmid for s in mech.species])
        low = np.array([s.low for s in mech.species]).ravel()
        high = np.array([s.high for s in mech.species]).ravel()
        kind = np.array([{"elementary": 0, "three-body": 1, "falloff": 2}[r.kind] for r in mech.reactions], dtype=np.int32)
        rev = np.array([1 if r.reversible else 0 for r in mech.reactions], dtype=np.int32)
        A = np.array([r.A for r in mech.reactions])
        b = np.array([r.b for r in mech.reactions])
        EaR = np.array([r.Ea_R for r in mech.reactions])
        nuf = np.zeros((R, N))
        nur = np.zeros((R, N))
        eff = np.zeros((R, N))
        lowp = np.zeros((R, 3))
        troe = np.zeros((R, 4))
        ntroe = np.zeros(R, dtype=np.int32)
        for j, r in enumerate(mech.reactions):
            for k, v in r.reactants.items():
                nuf[j, k] += v
            for k, v in r.products.items():
                nur[j, k] += v
            if r.efficiencies:
                eff[j] = r.efficiencies
            if r.low:
                lowp[j] = r.low
            if r.troe:
                ntroe[j] = len(r.troe)
                troe[j, : len(r.troe)] = r.troe
        self._keep = [self.W, Tmid, low, high, kind, rev, A, b, EaR, nuf, nur, eff, lowp, troe, ntroe]
        self.h = self.L.orc_mech_new(N, R, _d(self.W), _d(Tmid), _d(low), _d(high), _i(kind), _i(rev),
                                    _d(A), _d(b), _d(EaR), _d(nuf.ravel()), _d(nur.ravel()),
                                    _d(eff.ravel()), _d(lowp.ravel()), _d(troe.ravel()), _i(ntroe))
        self._keep += [nuf.ravel(), nur.ravel(), eff.ravel(), lowp.ravel(), troe.ravel()]

    def __del__(self):
        try:
            self.L.orc_mech_free(self.h)
        except Exception:
            pass

    def thermo(self, T):
        cp, h, s = (np.zeros(self.N) for _ in range(3))
        self.L.orc_thermo(self.h, float(T), _d(cp), _d(h), _d(s))
        return cp, h, s

    def T_from_e(self, e, Y, Tguess):
        it = np.zeros(1, dtype=np.int32)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        T = self.L.orc_T_from_e(self.h, float(e), _d(Y), float(Tguess), _i(it))
        return T, int(it[0])

    def progress(self, C, T, jac=False):
        C = np.ascontiguousarray(C, dtype=np.float64)
        q = np.zeros(self.R)
        if jac:
            dqdC = np.zeros((self.R, self.N))
            dqdT = np.zeros(self.R)
            self.L.orc_progress(self.h, _d(C), float(T), _d(q), _d(dqdC), _d(dqdT))
            return q, dqdC, dqdT
        self.L.orc_progress(self.h, _d(C), float(T), _d(q), None, None)
        return q

    def wdot(self, C, T):
        C = np.ascontiguousarray(C, dtype=np.float64)
        w = np.zeros(self.N)
        self.L.orc_wdot(self.h, _d(C), float(T), _d(w))
        return w

    def rhs(self, t, y, rhoe0, rhoesrc=0.0, rYsrc=None):
        y = np.ascontiguousarray(y, dtype=np.float64)
        src = np.zeros(self.N) if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        out = np.zeros(self.N + 1)
        T = self.L.orc_rhs(self.h, float(t), _d(y), float(rhoe0), float(rhoesrc), _d(src), _d(out))
        return out, T

    def jac(self, t, y, rhoe0, rhoesrc=0.0, rYsrc=None):
        y = np.ascontiguousarray(y, dtype=np.float64)
        src = np.zeros(self.N) if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        J = np.zeros((self.N + 1, self.N + 1))
        self.L.orc_jac(self.h, float(t), _d(y), float(rhoe0), float(rhoesrc), _d(src), _d(J))
        return J

    def react(self, rY, T, rE, dt, rYsrc=None, rEsrc=None, mask=None, rtol=1e-6, atol=1e-10,
              typ=None, max_steps=100000, max_order=5, nthreads=0):
        """rY: (N, ncells) component-major; T, rE: (ncells,). Returns dict of outputs (copies)."""
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, copy=True)
        rE = np.array(rE, dtype=np.float64, copy=True)
        nc = T.shape[0]
        assert rY.shape == (self.N, nc)
        src = None if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        esrc = None if rEsrc is None else np.ascontiguousarray(rEsrc, dtype=np.float64)
        msk = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        tv = None if typ is None else np.ascontiguousarray(typ, dtype=np.float64)
        FC = np.zeros(nc)
        status = np.zeros(nc, dtype=np.int32)
        stats = np.zeros((nc, 7), dtype=np.int32)
        nfail = self.L.orc_react(self.h, nc, _d(rY), _d(src), _d(T), _d(rE), _d(esrc), _d(FC), _i(msk),
                                float(dt), float(rtol), float(atol), _d(tv), int(max_steps),
                                int(max_order), int(nthreads), _i(status), _i(stats))
        return {"rY": rY, "T": T, "rE": rE, "FC": FC, "status": status, "stats": stats, "nfail": nfail}

    def replay(self, tape, y0, rhoe0, dt, logA=None, rhoesrc=0.0, rYsrc=None, tangent=None):
        """Replay one cell's recorded BDF tape (nrec, 16 + n) at (y0, rhoe0, logA). Returns
        (y_end, T_end); with tangent = (dy0, drhoe0, dlogA) also (dy_end, dT_end)."""
        tape = np.ascontiguousarray(tape, dtype=np.float64)
        n = self.N + 1
        assert tape.ndim == 2 and tape.shape[1] == 16 + n
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        la = None if logA is None else np.ascontiguousarray(logA, dtype=np.float64)
        src = None if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        yend = np.zeros(n)
        Tend = np.zeros(1)
        dy0 = de = dla = None
        dyend = dTend = None
        if tangent is not None:
            dy0 = np.ascontiguousarray(tangent[0], dtype=np.float64)
            de = float(tangent[1])
            dla = None if tangent[2] is None else np.ascontiguousarray(tangent[2], dtype=np.float64)
            dyend = np.zeros(n)
            dTend = np.zeros(1)
        self.L.orc_replay(self.h, _d(tape), tape.shape[0], tape.shape[1], _d(y0), float(rhoe0),
                         float(rhoesrc), _d(src), _d(la), float(dt), _d(yend), _d(Tend), _d(dy0),
                         de if de is not None else 0.0, _d(dla), _d(dyend), _d(dTend))
        if tangent is None:
            return yend, Tend[0]
        return yend, Tend[0], dyend, dTend[0]

    def react_rk64(self, rY, T, rE, dt, rYsrc=None, rEsrc=None, mask=None, rtol=1e-6, atol=1e-10,
                   typ=None, nsub_guess=10, nsub_min=5, nsub_max=100000, max_steps=100000,
                   nthreads=0):
        """ReactorRK64 restatement; stats columns: steps, RHS evaluations, rejected steps."""
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, copy=True)
        rE = np.array(rE, dtype=np.float64, copy=True)
        nc = T.shape[0]
        assert rY.shape == (self.N, nc)
        src = None if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        esrc = None if rEsrc is None else np.ascontiguousarray(rEsrc, dtype=np.float64)
        msk = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        tv = None if typ is None else np.ascontiguousarray(typ, dtype=np.float64)
        FC = np.zeros(nc)
        status = np.zeros(nc, dtype=np.int32)
        stats = np.zeros((nc, 3), dtype=np.int32)
        nfail = self.L.orc_react_rk64(self.h, nc, _d(rY), _d(src), _d(T), _d(rE), _d(esrc), _d(FC),
                                     _i(msk), float(dt), float(rtol), float(atol), _d(tv),
                                     int(nsub_guess), int(nsub_min), int(nsub_max), int(max_steps),
                                     int(nthreads), _i(status), _i(stats))
        return {"rY": rY, "T": T, "rE": rE, "FC": FC, "status": status, "stats": stats, "nfail": nfail}
