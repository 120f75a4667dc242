"""Seeded synthetic reactor inputs for the workloads this project measures.

Inputs are produced on the host from a counter-based generator (splitmix64 of seed, stream and
cell index), so any subset of cells can be regenerated bit-identically and the GPU and the CPU
baseline always consume the same states. Cell states follow Exec/RegTests/zeroD/prob.cpp:34-39
(PYT2RE: rho from p, T, Y; rho*e = rho * sum_k Y_k e_k(T)) with zero external sources, as
PeleC's react_state passes them on the first call (React.cpp:59-62).

Workloads (species order of the mechanism files):
  lidryer_32: LiDryer H2/air, 32^3 cells, T ~ U[1000,1600] K, p = 1 atm, phi ~ U[0.5,2];
              H, O, OH, HO2 seeded at Y = 1e-6 then renormalised.
  drm19_64:   drm19 CH4/air, 64^3 cells, T ~ U[900,2200] K, p ~ U[1,10] atm, phi ~ U[0.6,1.6];
              10 % of cells ignition-seeded: T ~ U[1400,1600] K with Y_H = Y_OH = 1e-4.
  drm19_adj:  drm19 CH4/air, 16384 cells, T ~ U[1000,1600] K, p = 1 atm, phi ~ U[0.5,2].
  dodecane_64: dodecane_lu n-C12H26/air, 64^3 cells, T ~ U[700,1800] K, p ~ U[10,60] atm,
              phi ~ U[0.5,2] (n_C12H26 = phi / 18.5 per O2), dt = 1e-5 s, rtol 1e-6.
  gri30_amr:  GRI-Mech 3.0 CH4/air AMR-shaped batch (make_amr_batch): boxes with edge lengths
              drawn from {8, 16, 32, 48, 64} per axis, 20 % of the cells masked out,
              T ~ U[800,2400] K, p ~ U[1,20] atm, phi ~ U[0.5,1.8], dt = 5e-6 s.
  drm19_amr:  the same batch with drm19 (same boxes, mask and states, drm19 species).
  gri30_flat: cells of the gri30_amr state distribution, no mask, as flat arrays.
Air is O2 + 3.76 N2 (no Ar); H2/air uses n_H2 = 2 phi, CH4/air n_CH4 = phi / 2 per O2,
n-dodecane/air n_C12H26 = phi / 18.5 per O2.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Optional

import numpy as np

from . import mechanism as M

ATM = 1013250.0
_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK64
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """U[0,1) doubles, one per cell index, independent per (seed, stream)."""
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        z = _splitmix64(idx.astype(np.uint64) ^ key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Thermo:
    """Host NASA-7 evaluation used to build the inputs (PYT2RE)."""

    def __init__(self, mech: M.Mechanism):
        self.W = np.array(mech.weights)
        self.low = np.array([s.low for s in mech.species])
        self.high = np.array([s.high for s in mech.species])
        self.Tmid = np.array([s.Tmid for s in mech.species])

    def e_mass(self, T: np.ndarray) -> np.ndarray:
        """(ncells, N) mass-specific internal energies."""
        T = T[:, None]
        a = np.where((T < self.Tmid)[..., None], self.low, self.high)
        hRT = (a[..., 0] + T * (a[..., 1] / 2 + T * (a[..., 2] / 3 + T * (a[..., 3] / 4 + T * a[..., 4] / 5)))
               + a[..., 5] / T)
        return (hRT - 1.0) * M.R_ERG * T / self.W

    def pyt2re(self, p, T, Y):
        Wbar = 1.0 / (Y / self.W).sum(-1)
        rho = p * Wbar / (M.R_ERG * T)
        e = (Y * self.e_mass(T)).sum(-1)
        return rho, e


@dataclasses.dataclass
class Workload:
    name: str
    mech: str
    ncells: int
    dt: float
    rtol: float
    atol: float
    seed: int


WORKLOADS: Dict[str, Workload] = {
    "lidryer_32": Workload("lidryer_32", "LiDryer", 32 ** 3, 1e-6, 1e-10, 1e-10, 1001),
    "drm19_64": Workload("drm19_64", "drm19", 64 ** 3, 1e-5, 1e-6, 1e-10, 2002),
    "drm19_adj": Workload("drm19_adj", "drm19", 16384, 1e-4, 1e-6, 1e-10, 4004),
    "gri30_flat": Workload("gri30_flat", "gri30", 262144, 5e-6, 1e-6, 1e-10, 5005),
    "dodecane_64": Workload("dodecane_64", "dodecane_lu", 64 ** 3, 1e-5, 1e-6, 1e-10, 3003),
}


def _fuel_air_Y(mech: M.Mechanism, fuel: str, nfuel: np.ndarray) -> np.ndarray:
    N = mech.nsp
    X = np.zeros((nfuel.shape[0], N))
    X[:, mech.species_index(fuel)] = nfuel
    X[:, mech.species_index("O2")] = 1.0
    X[:, mech.species_index("N2")] = 3.76
    W = np.array(mech.weights)
    Y = X * W
    return Y / Y.sum(-1, keepdims=True)


def make_inputs(workload: str, ncells: Optional[int] = None, first: int = 0) -> Dict[str, np.ndarray]:
    """Cells [first, first+ncells) of a workload. Returns rY (N, n), T, rE, p, phi, ign."""
    wl = WORKLOADS[workload]
    n = wl.ncells if ncells is None else ncells
    mech = M.load(wl.mech)
    th = Thermo(mech)
    idx = np.arange(first, first + n, dtype=np.int64)
    u = lambda s: uniform(wl.seed, s, idx)
    ign = np.zeros(n, dtype=bool)
    if wl.name == "lidryer_32":
        T = 1000.0 + 600.0 * u(0)
        p = np.full(n, ATM)
        phi = 0.5 + 1.5 * u(2)
        Y = _fuel_air_Y(mech, "H2", 2.0 * phi)
        for sp in ("H", "O", "OH", "HO2"):
            Y[:, mech.species_index(sp)] = 1e-6
        Y /= Y.sum(-1, keepdims=True)
    elif wl.name == "drm19_64":
        T = 900.0 + 1300.0 * u(0)
        p = ATM * (1.0 + 9.0 * u(1))
        phi = 0.6 + 1.0 * u(2)
        Y = _fuel_air_Y(mech, "CH4", 0.5 * phi)
        ign = u(3) < 0.1
        T = np.where(ign, 1400.0 + 200.0 * u(4), T)
        for sp in ("H", "OH"):
            Y[ign, mech.species_index(sp)] = 1e-4
        Y /= Y.sum(-1, keepdims=True)
    elif wl.name == "dodecane_64":
        T = 700.0 + 1100.0 * u(0)
        p = ATM * (10.0 + 50.0 * u(1))
        phi = 0.5 + 1.5 * u(2)
        Y = _fuel_air_Y(mech, "NC12H26", phi / 18.5)
    elif wl.name == "gri30_flat":
        T = 800.0 + 1600.0 * u(0)
        p = ATM * (1.0 + 19.0 * u(1))
        phi = 0.5 + 1.3 * u(2)
        Y = _fuel_air_Y(mech, "CH4", 0.5 * phi)
    elif wl.name == "drm19_adj":
        T = 1000.0 + 600.0 * u(0)
        p = np.full(n, ATM)
        phi = 0.5 + 1.5 * u(2)
        Y = _fuel_air_Y(mech, "CH4", 0.5 * phi)
    else:  # pragma: no cover
        raise KeyError(workload)
    rho, e = th.pyt2re(p, T, Y)
    rY = np.ascontiguousarray((rho[:, None] * Y).T)
    return {"rY": rY, "T": T.copy(), "rE": rho * e, "p": p, "phi": phi, "ign": ign,
            "rho": rho, "Y": Y}


def typical_values(rY: np.ndarray, T: np.ndarray, rhoY_min: float = 1e-10) -> np.ndarray:
    """PeleC::set_typical_values_chem (Source/React.cpp:13-25): typ rhoY_k = max(mid-range,
    typical_rhoY_val_min), typ T = mid-range of T."""
    typ = np.maximum(0.5 * (rY.min(axis=1) + rY.max(axis=1)), rhoY_min)
    return np.concatenate([typ, [0.5 * (T.min() + T.max())]])


def make_conservative_state(wl, n, seed=7, hydro=True, ufs=7):
    """Synthetic PeleC conservative state (NVAR = ufs + N, n) built on make_inputs(wl, n): random
    velocities (|u| <= 3e3 cm/s), rho E = rho e + KE; and, when hydro, a non-reacting source S_nr
    perturbing rhoY, momenta and rho E by up to 1e-3 (1e-4 for energy) per dt. Layout of
    Source/IndexDefines.H: URHO 0, UMX..UMZ 1..3, UEDEN 4, UEINT 5, UTEMP 6, species at ufs."""
    W = WORKLOADS[wl]
    d = make_inputs(wl, n)
    N = d["rY"].shape[0]
    rng = np.random.default_rng(seed)
    rho = d["rY"].sum(0)
    vel = rng.uniform(-3e3, 3e3, (3, n)) if hydro else np.zeros((3, n))
    S = np.zeros((ufs + N, n))
    S[0] = rho
    S[1:4] = rho * vel
    S[5] = d["rE"]
    S[4] = d["rE"] + 0.5 * rho * (vel ** 2).sum(0)
    S[6] = d["T"]
    S[ufs:ufs + N] = d["rY"]
    nonrs = np.zeros_like(S)
    if hydro:
        dt = W.dt
        nonrs[ufs:ufs + N] = 1e-3 * d["rY"] * rng.uniform(-1, 1, (N, n)) / dt
        nonrs[0] = nonrs[ufs:ufs + N].sum(0)
        nonrs[1:4] = 1e-3 * S[1:4] * rng.uniform(-1, 1, (3, n)) / dt
        nonrs[4] = 1e-4 * S[4] * rng.uniform(-1, 1, n) / dt
    return S, nonrs


AMR_EDGES = (8, 16, 32, 48, 64)
AMR_WORKLOADS = {
    "gri30": Workload("gri30_amr", "gri30", 0, 5e-6, 1e-6, 1e-10, 5005),
    "drm19": Workload("drm19_amr", "drm19", 0, 5e-6, 1e-6, 1e-10, 5005),
}
AMR = AMR_WORKLOADS["drm19"]


def make_amr_batch(nboxes: int = 256, mech_name: str = "drm19") -> Dict[str, object]:
    """The <mech>_amr batch: per box (lo, hi, nx, ny, nz, first cell); all cells concatenated box
    by box, x fastest, as rY (N, ncells), T, rE and mask (int32, 0 = covered: not integrated).
    Boxes, mask, T, p and phi do not depend on the mechanism."""
    wl = AMR_WORKLOADS[mech_name]
    mech = M.load(wl.mech)
    th = Thermo(mech)
    bi = np.arange(nboxes, dtype=np.int64)
    shapes = np.stack([np.array(AMR_EDGES)[np.minimum((uniform(wl.seed, 100 + ax, bi) * 5).astype(int), 4)]
                       for ax in range(3)], axis=1)
    boxes = []
    first = 0
    for b in range(nboxes):
        nx, ny, nz = (int(v) for v in shapes[b])
        lo = (0, 0, 100 * b)
        boxes.append({"lo": lo, "hi": (nx - 1, ny - 1, lo[2] + nz - 1), "shape": (nx, ny, nz),
                      "first": first})
        first += nx * ny * nz
    idx = np.arange(first, dtype=np.int64)
    u = lambda st: uniform(wl.seed, st, idx)
    T = 800.0 + 1600.0 * u(0)
    p = ATM * (1.0 + 19.0 * u(1))
    phi = 0.5 + 1.3 * u(2)
    mask = (u(3) >= 0.2).astype(np.int32)
    Y = _fuel_air_Y(mech, "CH4", 0.5 * phi)
    rho, e = th.pyt2re(p, T, Y)
    return {"boxes": boxes, "ncells": first, "rY": (rho[:, None] * Y).T.copy(), "T": T,
            "rE": rho * e, "mask": mask, "workload": wl}
