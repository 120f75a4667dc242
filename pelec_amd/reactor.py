"""Host-side mirror of ReactorBase over the C-ABI (include/pelec_amd.h).

Python counterpart of the C++ shim in include/pelec_amd_reactor.H, used by the tests and the
benchmark. Device memory and streams come from torch (ROCm); every call goes through the gfx950
library libpelec_amd_<mechanism>.so, and constructing a Reactor fails loudly if that library is not
built -- there is no CPU fallback.

Reference interface (Source/PeleC.cpp:2032-2047, Source/React.cpp:7-27, 222-228):
    reactor = ReactorBase::create(chem_integrator); reactor->init(1, 1);
    reactor->set_typ_vals_ode(typ);
    reactor->react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt, time, stream);
    reactor->close();
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

from . import build as _build

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32

DIAG_NAMES = ("status", "nst", "nfe", "nje", "nsetups", "nni", "netf", "ncfn")


class Array4(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("jstride", _i64), ("kstride", _i64), ("nstride", _i64),
                ("begin", _i32 * 3), ("end", _i32 * 3), ("ncomp", _i32)]


class IArray4(ctypes.Structure):
    _fields_ = Array4._fields_


class BoxArgs(ctypes.Structure):
    _fields_ = [("lo", _i32 * 3), ("hi", _i32 * 3), ("rY", Array4), ("rYsrc_ext", Array4),
                ("T", Array4), ("rEner", Array4), ("rEner_src_ext", Array4), ("FC", Array4),
                ("mask", IArray4)]


class ReactorError(RuntimeError):
    pass


_ERR = {-1: "PRC_EINVAL", -2: "PRC_EHIP", -3: "PRC_ESTATE", -4: "PRC_ENOMEM"}

_LIBS: Dict[str, ctypes.CDLL] = {}

#: exported functions and their ctypes signatures (restype, argtypes)
_SIGS = {
    "prc_num_species": (ctypes.c_int, []),
    "prc_num_reactions": (ctypes.c_int, []),
    "prc_species_name": (ctypes.c_char_p, [ctypes.c_int]),
    "prc_mechanism_name": (ctypes.c_char_p, []),
    "prc_mechanism_sha256": (ctypes.c_char_p, []),
    "prc_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "prc_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "prc_set_typ_vals_ode": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "prc_set_param_real": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]),
    "prc_set_param_int": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "prc_set_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _i64]),
    "prc_get_param_int": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "prc_autotune": (ctypes.c_int, [ctypes.c_void_p, _i64, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "prc_close": (ctypes.c_int, [ctypes.c_void_p]),
    "prc_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "prc_react_box": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                     Array4, Array4, Array4, Array4, Array4, Array4, IArray4,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                     ctypes.c_void_p]),
    "prc_react_flat": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 7 + [_i64,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_void_p]),
    "prc_react_flat_noalloc": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 7 + [_i64,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.c_void_p]),
    "prc_react_boxes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(BoxArgs),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.c_void_p]),
    "prc_last_failures": (_i64, [ctypes.c_void_p, ctypes.c_void_p]),
    "prc_kernel_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(_i64)]),
    "prc_react_state_box": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i32),
                                           ctypes.POINTER(_i32), Array4, Array4, Array4, Array4,
                                           IArray4, Array4, Array4, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "prc_compute_temp": (ctypes.c_int, [ctypes.POINTER(_i32), ctypes.POINTER(_i32), Array4,
                                        IArray4, ctypes.c_int, ctypes.c_void_p]),
    "prc_adjoint": (ctypes.c_int, [ctypes.c_void_p, _i64] + [ctypes.c_void_p] * 5 + [ctypes.c_double]
                    + [ctypes.c_void_p] * 10 + [_i64, ctypes.c_void_p, ctypes.c_void_p]),
    "prc_eos_pyt2re": (ctypes.c_int, [_i64] + [ctypes.c_void_p] * 6),
    "prc_eos_rey2t": (ctypes.c_int, [_i64] + [ctypes.c_void_p] * 5),
    "prc_eos_rty2hi": (ctypes.c_int, [_i64] + [ctypes.c_void_p] * 3),
    "prc_eos_rty2wdot": (ctypes.c_int, [_i64] + [ctypes.c_void_p] * 5),
    "prc_model_eval": (ctypes.c_int, [_i64, ctypes.c_double] + [ctypes.c_void_p] * 6
                       + [ctypes.c_int, ctypes.c_void_p]),
}


def load_library(mech: str) -> ctypes.CDLL:
    """Load libpelec_amd_<mech>.so (raises if it has not been built)."""
    if mech in _LIBS:
        return _LIBS[mech]
    path = _build.lib_path(mech)
    if not os.path.exists(path):
        raise ReactorError(f"{path} is not built; run __graft_entry__.build() or "
                           f"python -m pelec_amd.build {mech}")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[mech] = lib
    return lib


def _check(code: int, what: str):
    if code != 0:
        raise ReactorError(f"{what} failed: {_ERR.get(code, code)}")


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    return t.data_ptr()


def _stream_handle(stream) -> Optional[int]:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def array4(t, lo=(0, 0, 0)) -> Array4:
    """Array4 view of a contiguous torch tensor shaped (ncomp, nz, ny, nx) (component-major)."""
    a = Array4()
    if t is None:
        return a
    if t.dim() == 3:
        t = t.unsqueeze(0)
    nc, nz, ny, nx = t.shape
    st = t.stride()
    a.p = t.data_ptr()
    a.nstride, a.kstride, a.jstride = st[0], st[1], st[2]
    assert st[3] == 1 or nx == 1, "x must be the unit-stride dimension"
    a.begin[:] = lo
    a.end[:] = (lo[0] + nx, lo[1] + ny, lo[2] + nz)
    a.ncomp = nc
    return a


def iarray4(t, lo=(0, 0, 0)) -> IArray4:
    a = IArray4()
    if t is None:
        return a
    b = array4(t, lo)
    for f, _ in Array4._fields_:
        setattr(a, f, getattr(b, f))
    return a


class Reactor:
    """ReactorBase-shaped handle on one mechanism's gfx950 reactor."""

    def __init__(self, mech: str = "drm19", chem_integrator: str = "ReactorCvode",
                 reactor_type: int = 1, ncells: int = 1):
        self.mech = mech
        self.lib = load_library(mech)
        self.nsp = self.lib.prc_num_species()
        self.species = [self.lib.prc_species_name(k).decode() for k in range(self.nsp)]
        h = ctypes.c_void_p()
        _check(self.lib.prc_create(chem_integrator.encode(), ctypes.byref(h)), "prc_create")
        self.h = h
        _check(self.lib.prc_init(self.h, reactor_type, ncells), "prc_init")
        self._diag = None

    # -- configuration -----------------------------------------------------------------
    def set_typ_vals_ode(self, typ: Optional[Sequence[float]]):
        if typ is None:
            _check(self.lib.prc_set_typ_vals_ode(self.h, None, 0), "set_typ_vals_ode")
            return
        arr = (ctypes.c_double * len(typ))(*typ)
        _check(self.lib.prc_set_typ_vals_ode(self.h, arr, len(typ)), "set_typ_vals_ode")

    def set_param(self, key: str, value):
        if isinstance(value, float):
            _check(self.lib.prc_set_param_real(self.h, key.encode(), value), f"set {key}")
        else:
            _check(self.lib.prc_set_param_int(self.h, key.encode(), int(value)), f"set {key}")

    def configure(self, rtol=None, atol=None, max_order=None, max_steps=None, form=None,
                  lu_fp32=None, kernel_timing=None, rk64_nsubsteps=None, cooperative=None,
                  lu_refine=None, max_slots=None):
        """rk64_nsubsteps: (guess, min, max) substep counts of ReactorRK64. cooperative: which of
        the batched form's kernels run one cell per 32-lane group -- True / 3 both, 1 the Newton
        kernel, 2 the Jacobian kernel, False / 0 neither. lu_refine: FP64 refinement steps (1 or 2)
        after each FP32-LU solve. max_slots: slots of the batched form's work queue (0: one per
        cell while device memory allows); a call with more cells refills freed slots."""
        if rtol is not None:
            self.set_param("ode.rtol", float(rtol))
        if atol is not None:
            self.set_param("ode.atol", float(atol))
        if max_order is not None:
            self.set_param("cvode.max_order", int(max_order))
        if max_steps is not None:
            self.set_param("ode.max_steps", int(max_steps))
        if form is not None:
            self.set_param("prc.form", {"plain": 0, "batched": 1}.get(form, form))
        if lu_fp32 is not None:
            self.set_param("prc.lu_fp32", 1 if lu_fp32 else 0)
        if kernel_timing is not None:
            self.set_param("prc.kernel_timing", 1 if kernel_timing else 0)
        if lu_refine is not None:
            self.set_param("prc.lu_refine", int(lu_refine))
        if max_slots is not None:
            self.set_param("prc.max_slots", int(max_slots))
        if cooperative is not None:
            self.set_param("prc.cooperative",
                           3 if cooperative is True else int(cooperative))
        if rk64_nsubsteps is not None:
            g, lo, hi = (int(x) for x in rk64_nsubsteps)
            self.set_param("rk64.nsubsteps_guess", g)
            self.set_param("rk64.nsubsteps_min", lo)
            self.set_param("rk64.nsubsteps_max", hi)
        return self

    SHAPE_KEYS = ("prc.iters_per_graph", "prc.cooperative", "prc.blocks_newton", "prc.blocks_jac",
                  "prc.blocks_lu", "prc.compact_pct", "prc.compact_every", "prc.coop_tail_slots",
                  "prc.blocks_tail")

    def get_param(self, key: str) -> int:
        v = ctypes.c_int(0)
        _check(self.lib.prc_get_param_int(self.h, key.encode(), ctypes.byref(v)), f"get {key}")
        return v.value

    def shape(self):
        """Launch shape of the batched form in use."""
        return {k: self.get_param(k) for k in self.SHAPE_KEYS}

    def autotune(self, rY, T, rE, dt: float, stream=None):
        """Pick the batched launch shape by timing react calls on a sample (rY (N, n), T, rE (n,)
        cuda float64; not modified). Returns (shape dict, best ms)."""
        ms = ctypes.c_double(0.0)
        _check(self.lib.prc_autotune(self.h, T.shape[0], _ptr(rY), _ptr(T), _ptr(rE), float(dt),
                                     ctypes.byref(ms), _stream_handle(stream)), "prc_autotune")
        return self.shape(), ms.value

    def set_diagnostics(self, diag_tensor):
        """diag_tensor: int32 cuda tensor shaped (8, capacity) or None."""
        self._diag = diag_tensor
        if diag_tensor is None:
            _check(self.lib.prc_set_diagnostics(self.h, None, 0), "set_diagnostics")
        else:
            assert diag_tensor.is_contiguous() and diag_tensor.shape[0] == 8
            _check(self.lib.prc_set_diagnostics(self.h, diag_tensor.data_ptr(), diag_tensor.shape[1]),
                   "set_diagnostics")

    # -- reaction step -----------------------------------------------------------------
    def react_flat(self, rY, T, rE, dt: float, rYsrc=None, rEsrc=None, FC=None, mask=None,
                   time: float = 0.0, stream=None):
        """rY (N, ncells), T/rE/rEsrc/FC (ncells,) float64 cuda tensors, updated in place."""
        nc = T.shape[0]
        assert rY.shape == (self.nsp, nc) and rY.is_contiguous()
        dtc = ctypes.c_double(dt)
        tm = ctypes.c_double(time)
        _check(self.lib.prc_react_flat(self.h, _ptr(rY), _ptr(rYsrc), _ptr(T), _ptr(rE), _ptr(rEsrc),
                                       _ptr(FC), _ptr(mask), nc, ctypes.byref(dtc), ctypes.byref(tm),
                                       _stream_handle(stream)), "prc_react_flat")
        return tm.value

    def react_flat_noalloc(self, rY, T, rE, dt: float, rYsrc=None, rEsrc=None, FC=None, mask=None,
                           time: float = 0.0, stream=None):
        """react_flat through prc_react_flat_noalloc: the batched form with its workspace in the
        library's static device array (no device allocation during the call)."""
        nc = T.shape[0]
        assert rY.shape == (self.nsp, nc) and rY.is_contiguous()
        dtc = ctypes.c_double(dt)
        tm = ctypes.c_double(time)
        _check(self.lib.prc_react_flat_noalloc(self.h, _ptr(rY), _ptr(rYsrc), _ptr(T), _ptr(rE),
                                               _ptr(rEsrc), _ptr(FC), _ptr(mask), nc,
                                               ctypes.byref(dtc), ctypes.byref(tm),
                                               _stream_handle(stream)), "prc_react_flat_noalloc")
        return tm.value

    def react_box(self, lo, hi, rY, rYsrc, T, rE, rEsrc, FC, mask, dt: float, time: float = 0.0,
                  stream=None):
        """Each array: torch tensor (ncomp, nz, ny, nx) covering box [lo, hi] (or None)."""
        lo_c = (_i32 * 3)(*lo)
        hi_c = (_i32 * 3)(*hi)
        dtc = ctypes.c_double(dt)
        tm = ctypes.c_double(time)
        _check(self.lib.prc_react_box(self.h, lo_c, hi_c, array4(rY, lo), array4(rYsrc, lo),
                                      array4(T, lo), array4(rE, lo), array4(rEsrc, lo),
                                      array4(FC, lo), iarray4(mask, lo), ctypes.byref(dtc),
                                      ctypes.byref(tm), _stream_handle(stream)), "prc_react_box")
        return tm.value

    def react_boxes(self, boxes, dt: float, time: float = 0.0, stream=None):
        """boxes: list of dicts with lo, hi, rY, rYsrc, T, rE, rEsrc, FC, mask (tensors/None)."""
        arr = (BoxArgs * len(boxes))()
        for b, d in zip(arr, boxes):
            lo = d["lo"]
            b.lo[:] = lo
            b.hi[:] = d["hi"]
            b.rY = array4(d["rY"], lo)
            b.rYsrc_ext = array4(d.get("rYsrc"), lo)
            b.T = array4(d["T"], lo)
            b.rEner = array4(d["rE"], lo)
            b.rEner_src_ext = array4(d.get("rEsrc"), lo)
            b.FC = array4(d.get("FC"), lo)
            b.mask = iarray4(d.get("mask"), lo)
        dtc = ctypes.c_double(dt)
        tm = ctypes.c_double(time)
        _check(self.lib.prc_react_boxes(self.h, len(boxes), arr, ctypes.byref(dtc), ctypes.byref(tm),
                                        _stream_handle(stream)), "prc_react_boxes")
        return tm.value

    def react_state_box(self, lo, hi, S_old, S_new, nonreact_src, I_R, dt: float, ufs: int = 7,
                        react_init: bool = False, mask=None, rot_energy=None, cost=None,
                        time: float = 0.0, stream=None):
        """PeleC::react_state on one box. State arrays: torch tensors (NVAR, nz, ny, nx);
        I_R: (NUM_SPECIES + 2, nz, ny, nx); mask (int32), rot_energy, cost: (nz, ny, nx) or None."""
        lo_c = (_i32 * 3)(*lo)
        hi_c = (_i32 * 3)(*hi)
        dtc = ctypes.c_double(dt)
        tm = ctypes.c_double(time)
        _check(self.lib.prc_react_state_box(self.h, lo_c, hi_c, array4(S_old, lo), array4(S_new, lo),
                                            array4(nonreact_src, lo), array4(I_R, lo),
                                            iarray4(mask, lo), array4(rot_energy, lo),
                                            array4(cost, lo), int(ufs), 1 if react_init else 0,
                                            ctypes.byref(dtc), ctypes.byref(tm),
                                            _stream_handle(stream)), "prc_react_state_box")
        return tm.value

    def adjoint(self, rY0, T0, rE0, dt: float, logA=None, ct_rY=None, ct_T=None, rYsrc=None,
                rEsrc=None, tape_cap: int = 0, stream=None):
        """Discrete adjoint of the BDF reactor (prc_adjoint). rY0 (N, ncells), T0/rE0 (ncells,),
        logA (NRXN,) or None, cotangents ct_rY (N, ncells) / ct_T (ncells,) or None: cuda float64.
        Returns a dict: rY, T (end state), g_rY0, g_rE0, g_logA (NRXN, ncells), status, and with
        tape_cap > 0 the complete forward tapes (tape, tape_count)."""
        import torch
        nc = T0.shape[0]
        assert rY0.shape == (self.nsp, nc) and rY0.is_contiguous()
        dev = T0.device
        out = {"rY": torch.empty_like(rY0), "T": torch.empty_like(T0),
               "g_rY0": torch.empty_like(rY0), "g_rE0": torch.empty_like(T0),
               "g_logA": torch.empty((self.lib.prc_num_reactions(), nc), dtype=torch.float64, device=dev),
               "status": torch.empty(nc, dtype=torch.int32, device=dev)}
        tape = cnt = None
        if tape_cap > 0:
            tape = torch.empty((nc, tape_cap, 16 + self.nsp + 1), dtype=torch.float64, device=dev)
            cnt = torch.empty(nc, dtype=torch.int32, device=dev)
            out["tape"], out["tape_count"] = tape, cnt
        _check(self.lib.prc_adjoint(self.h, nc, _ptr(rY0), _ptr(T0), _ptr(rE0), _ptr(rYsrc),
                                    _ptr(rEsrc), float(dt), _ptr(logA), _ptr(ct_rY), _ptr(ct_T),
                                    _ptr(out["rY"]), _ptr(out["T"]), _ptr(out["g_rY0"]),
                                    _ptr(out["g_rE0"]), _ptr(out["g_logA"]), _ptr(out["status"]),
                                    _ptr(tape), int(tape_cap), _ptr(cnt), _stream_handle(stream)),
               "prc_adjoint")
        return out

    def last_failures(self, stream=None) -> int:
        n = self.lib.prc_last_failures(self.h, _stream_handle(stream))
        if n < 0:
            raise ReactorError(f"prc_last_failures: {_ERR.get(n, n)}")
        return n

    def kernel_times(self):
        """{kernel: total ms} and launches of the previous batched call (kernel_timing mode)."""
        ms = (ctypes.c_double * 3)()
        n = _i64(0)
        _check(self.lib.prc_kernel_times(self.h, ms, ctypes.byref(n)), "prc_kernel_times")
        return {"k_newton": ms[0], "k_jac": ms[1], "k_lu": ms[2]}, int(n.value)

    def close(self):
        if self.h is not None:
            _check(self.lib.prc_destroy(self.h), "prc_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Eos:
    """Boundary EOS helpers of one mechanism library (all arrays cuda float64)."""

    def __init__(self, mech: str):
        self.lib = load_library(mech)
        self.nsp = self.lib.prc_num_species()

    def pyt2re(self, p, T, Y, stream=None):
        import torch
        nc = T.shape[0]
        rho = torch.empty_like(T)
        e = torch.empty_like(T)
        _check(self.lib.prc_eos_pyt2re(nc, _ptr(p), _ptr(T), _ptr(Y), _ptr(rho), _ptr(e),
                                       _stream_handle(stream)), "prc_eos_pyt2re")
        return rho, e

    def rey2t(self, rho, e, Y, T, stream=None):
        _check(self.lib.prc_eos_rey2t(T.shape[0], _ptr(rho), _ptr(e), _ptr(Y), _ptr(T),
                                      _stream_handle(stream)), "prc_eos_rey2t")
        return T

    def compute_temp(self, lo, hi, S, ufs: int = 7, mask=None, stream=None):
        """PeleC::computeTemp's pc_cmpTemp over a box; S (NVAR, nz, ny, nx) updated in place."""
        lo_c = (_i32 * 3)(*lo)
        hi_c = (_i32 * 3)(*hi)
        _check(self.lib.prc_compute_temp(lo_c, hi_c, array4(S, lo), iarray4(mask, lo), int(ufs),
                                         _stream_handle(stream)), "prc_compute_temp")
        return S

    def rty2hi(self, T, stream=None):
        import torch
        hi = torch.empty((self.nsp, T.shape[0]), dtype=T.dtype, device=T.device)
        _check(self.lib.prc_eos_rty2hi(T.shape[0], _ptr(T), _ptr(hi), _stream_handle(stream)),
               "prc_eos_rty2hi")
        return hi

    def model(self, t, y, rhoe0, rhoesrc=None, rYsrc=None, jacobian=True, cooperative=False,
              stream=None):
        """Reactor RHS (and exact Jacobian) at states y (n x ncells): (ydot, J or None)."""
        import torch
        nc = y.shape[1]
        ydot = torch.empty_like(y)
        J = torch.empty((nc, y.shape[0], y.shape[0]), dtype=y.dtype, device=y.device) \
            if jacobian else None
        _check(self.lib.prc_model_eval(nc, float(t), _ptr(y), _ptr(rhoe0), _ptr(rhoesrc),
                                       _ptr(rYsrc), _ptr(ydot), _ptr(J), int(bool(cooperative)),
                                       _stream_handle(stream)), "prc_model_eval")
        return ydot, J

    def rty2wdot(self, rho, T, Y, stream=None):
        import torch
        w = torch.empty((self.nsp, T.shape[0]), dtype=T.dtype, device=T.device)
        _check(self.lib.prc_eos_rty2wdot(T.shape[0], _ptr(rho), _ptr(T), _ptr(Y), _ptr(w),
                                         _stream_handle(stream)), "prc_eos_rty2wdot")
        return w
