"""Algorithmic FP64 work per reactor call, counted from the generated mechanism code.

W_cell = n_rhs F_rhs + n_jac F_jac + n_lu (2/3 n^3 + n^2) + n_solve (2 n^2 + 8 n) + n_step F_step

F_rhs and F_jac are the add/sub/mul/div operations of the emitted straight-line code (an fma counts
as 2), plus the generic per-cell code around it (T(e) inversion, energy closure, Jacobian change of
variables). Transcendental calls (exp, log, log10, pow, sqrt) are counted separately and not as
flops. The n_* are the per-cell counters the kernels return (RHS evaluations, Jacobian evaluations,
LU factorisations, Newton iterations, steps).
"""
from __future__ import annotations

import functools
import os
import re
from typing import Dict

from . import mechgen

_FUNC = re.compile(r"static inline void (\w+)\(")


def _split_functions(text: str) -> Dict[str, str]:
    out = {}
    pos = [(m.group(1), m.start()) for m in _FUNC.finditer(text)]
    for i, (name, start) in enumerate(pos):
        end = pos[i + 1][1] if i + 1 < len(pos) else len(text)
        out[name] = text[start:end]
    return out


def _count_expr(line: str):
    line = re.sub(r"\[[^\]]*\]", "", line)           # index arithmetic is folded
    line = re.sub(r"//.*", "", line)
    flops = 2 * line.count("fma(")
    trans = sum(line.count(f) for f in ("exp(", "log(", "log10(", "pow(", "sqrt("))
    body = line.replace("fma(", "(")
    # compound assignments
    flops += len(re.findall(r"[-+*/]=", body))
    body = re.sub(r"[-+*/]=", "=", body)
    # binary operators: a sign after '(' , ',' , '=' or at start is unary
    for m in re.finditer(r"[-+*/]", body):
        prev = body[: m.start()].rstrip()
        if m.group(0) in "+-" and (not prev or prev[-1] in "(,=?:e" or prev.endswith("return")):
            # unary sign or exponent of a literal (1e-05)
            continue
        flops += 1
    return flops, trans


def _count_function(src: str):
    flops = trans = 0
    trip = 1
    for raw in src.splitlines():
        line = raw.strip()
        if not line or line.startswith("#") or line.startswith("//"):
            continue
        m = re.match(r"for \(int \w+ = 0; \w+ < (\d+); \+\+\w+\)(.*)", line)
        if m:
            trip_here = int(m.group(1))
            rest = m.group(2).strip()
            if rest.startswith("for"):
                m2 = re.match(r"for \(int \w+ = 0; \w+ < (\d+); \+\+\w+\)(.*)", rest)
                trip_here *= int(m2.group(1))
                rest = m2.group(2).strip()
            if rest and rest != "{":
                f, t = _count_expr(rest)
                flops += trip_here * f
                trans += trip_here * t
            elif rest == "":
                trip = trip_here   # applies to the next line
            continue
        if line.startswith(("bool", "double eg", "double u", "double gRT", "(void)")) and "=" not in line:
            continue
        f, t = _count_expr(line)
        flops += trip * f
        trans += trip * t
        trip = 1
    return flops, trans


@functools.lru_cache(maxsize=None)
def mechanism_costs(mech: str) -> Dict[str, float]:
    path = os.path.join(mechgen.GEN_DIR, f"mech_{mech}.h")
    if not os.path.exists(path):
        path = mechgen.generate(mech)
    fns = _split_functions(open(path).read())
    c = {k: _count_function(v) for k, v in fns.items()}
    N = int(re.search(r"NSP = (\d+);", open(path).read()).group(1))
    n = N + 1
    # generic per-cell code (cell_model.h): rho/Y/C (3N), T(e) with the mixture polynomial
    # (2 Newton iterations x 22 + setup), energy closure (6N + 2)
    t_inv = c["mix_e_poly"][0] + 2 * 22
    pre = c["prelude"][0] + c["g_h_RT"][0]       # exp(g/RT) table (called by wdot / wdot_jac)
    pre_tr = c["prelude"][1] + c["g_h_RT"][1]
    rhs = c["wdot"][0] + pre + c["e_cv"][0] + t_inv + 3 * N + 6 * N + 2
    rhs_tr = c["wdot"][1] + pre_tr + 1
    jac = (c["wdot_jac"][0] + pre + c["e_cv"][0] + c["dcv_dT"][0] + t_inv + 3 * N + 10 * N
           + N * (6 * N + 5) + 3 * n)
    jac_tr = c["wdot_jac"][1] + pre_tr + 1
    return {"N": N, "n": n, "F_rhs": float(rhs), "F_rhs_transcendental": float(rhs_tr),
            "F_jac": float(jac), "F_jac_transcendental": float(jac_tr),
            "F_lu": 2.0 / 3.0 * n ** 3 + n * n, "F_solve": 2.0 * n * n + 8 * n,
            "F_step": 30.0 * n}


def work(mech: str, nfe, nje, nsetups, nni, nst) -> float:
    """Total algorithmic flops for per-cell counter arrays (or sums)."""
    c = mechanism_costs(mech)
    return float(c["F_rhs"] * nfe + c["F_jac"] * nje + c["F_lu"] * nsetups + c["F_solve"] * nni
                 + c["F_step"] * nst)
