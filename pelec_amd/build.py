"""Build the per-mechanism gfx950 reactor libraries in-tree (pelec_amd/lib/).

One shared library per mechanism, like the reference's one-mechanism-per-build
(Chemistry_Model in Exec/Make.PeleC). The device mechanism code is generated first
(pelec_amd.mechgen), then reactor.hip is compiled with hipcc for gfx950.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from typing import Iterable, List

from . import mechgen
from . import mechanism as M

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

#: mechanisms built by default
DEFAULT_MECHS = ("LiDryer", "drm19", "drm19_inert40")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def lib_path(mech: str) -> str:
    return os.path.join(LIBDIR, f"libpelec_amd_{mech}.so")


def _ident(name: str) -> str:
    return mechgen._ident(name)


def build_one(mech: str, verbose: bool = False) -> str:
    header = mechgen.generate(mech)
    os.makedirs(LIBDIR, exist_ok=True)
    out = lib_path(mech)
    tmp = out + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + [
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        f"-DPELEC_MECH={_ident(mech)}",
        f'-DPELEC_MECH_HEADER="gen/{os.path.basename(header)}"',
        os.path.join(CSRC, "reactor.hip"), "-o", tmp,
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {mech}:\n{res.stderr[-4000:]}")
    if verbose and res.stderr:
        sys.stderr.write(res.stderr)
    os.replace(tmp, out)
    return out


def build(mechs: Iterable[str] = DEFAULT_MECHS, jobs: int = 4) -> List[str]:
    mechs = list(mechs)
    for m in mechs:
        if m not in M.available():
            raise ValueError(f"unknown mechanism {m}")
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(mechs)))) as ex:
        return list(ex.map(build_one, mechs))


if __name__ == "__main__":
    for p in build(sys.argv[1:] or DEFAULT_MECHS):
        print(p)
