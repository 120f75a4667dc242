"""ORACLE -- test infrastructure only. numpy restatement of PeleC::react_state's per-box body
(Source/React.cpp:29-362) around the oracle reactor, and of computeTemp's pc_cmpTemp
(Source/Utilities.H:68-85). Only tests/ use it, as the checker of prc_react_state_box and
prc_compute_temp.

Layout (Source/IndexDefines.H:6-13,60-69): URHO 0, UMX..UMZ 1..3, UEDEN 4, UEINT 5, UTEMP 6,
species from ufs = 7 + NUM_ADV. State arrays are (NVAR, ncells) component-major.
"""
import numpy as np

URHO, UMX, UMY, UMZ, UEDEN, UEINT, UTEMP = range(7)
R_ERG = 8.31446261815324e7


def react_state(om, npm, S_old, S_new, nonrs, dt, ufs=7, react_init=False, mask=None, rot=None,
                rtol=1e-6, atol=1e-10, nthreads=0):
    """Returns (S_new_out, I_R, FC). Follows React.cpp step by step:
    pack (99-121), forcing (181-220), react (222-228), unpack (233-314), heat release (325-343)."""
    N = npm.N
    S_new = np.array(S_new, dtype=float, copy=True)
    so = S_new.copy() if react_init else np.asarray(S_old, dtype=float)
    nc = S_new.shape[1]
    rot = np.zeros(nc) if rot is None else np.asarray(rot, dtype=float)
    act = np.ones(nc, bool) if mask is None else np.asarray(mask) != 0
    # old internal energy and the external energy forcing
    rho_old = so[URHO]
    e_old = (so[UEDEN] - 0.5 * (so[UMX] ** 2 + so[UMY] ** 2 + so[UMZ] ** 2) / rho_old) / rho_old + rot
    rho_new = S_new[URHO]
    ke_new = 0.5 * (S_new[UMX] ** 2 + S_new[UMY] ** 2 + S_new[UMZ] ** 2) / rho_new
    redot = (S_new[UEDEN] - ke_new + rho_new * rot - rho_old * e_old) / dt
    # STemp = [rhoY, T, rho e] of the (old) state; species forcing from the non-reacting source
    rY = so[ufs:ufs + N].copy()
    out = om.react(rY, so[UTEMP].copy(), so[UEINT].copy(), dt, rYsrc=nonrs[ufs:ufs + N],
                   rEsrc=redot, mask=act.astype(np.int32), rtol=rtol, atol=atol, nthreads=nthreads)
    rYn, Tn = out["rY"], out["T"]
    # unpack
    um = so[UMX] + dt * nonrs[UMX]
    vm = so[UMY] + dt * nonrs[UMY]
    wm = so[UMZ] + dt * nonrs[UMZ]
    rhonew = rYn.sum(0)
    reint = rho_old * e_old + dt * redot
    ke = 0.5 * (um * um + vm * vm + wm * wm) / rhonew - rhonew * rot
    IR = np.zeros((N + 2, nc))
    IR[:N] = (rYn - so[ufs:ufs + N]) / dt - nonrs[ufs:ufs + N]
    IR[N] = (reint + ke - so[UEDEN]) / dt - nonrs[UEDEN]
    Snew = S_new.copy()
    if not react_init:
        Snew[URHO] = rhonew
        Snew[UMX], Snew[UMY], Snew[UMZ] = um, vm, wm
        Snew[ufs:ufs + N] = rYn
        Snew[UTEMP] = Tn
        Snew[UEINT] = reint
        Snew[UEDEN] = reint + ke
    # heat release with h_k(T) of the updated state
    hk = npm.h_RT(Snew[UTEMP]) * R_ERG * Snew[UTEMP][:, None] / npm.W    # (nc, N) erg/g
    IR[N + 1] = -(hk.T * IR[:N]).sum(0)
    # covered cells: state untouched, I_R = 0
    Snew[:, ~act] = S_new[:, ~act]
    IR[:, ~act] = 0.0
    return Snew, IR, out["FC"]


def compute_temp(npm, S, ufs=7, mask=None):
    """pc_cmpTemp: T = T(e = rho e / rho, Y) with the stored T as the guess."""
    S = np.array(S, dtype=float, copy=True)
    N = npm.N
    act = np.ones(S.shape[1], bool) if mask is None else np.asarray(mask) != 0
    Y = (S[ufs:ufs + N] / S[URHO]).T
    T = npm.T_from_e(S[UEINT] / S[URHO], Y, S[UTEMP])
    S[UTEMP, act] = T[act]
    return S
