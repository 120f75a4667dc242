"""Build the per-mechanism gfx950 reactor libraries in-tree (pelec_amd/lib/).

One shared library per mechanism, like the reference's one-mechanism-per-build
(Chemistry_Model in Exec/Make.PeleC). The device mechanism code is generated first
(pelec_amd.mechgen), then reactor.hip is compiled with hipcc for gfx950.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys
from typing import Dict, Iterable, List

from . import mechgen
from . import mechanism as M

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

#: mechanisms built by default
DEFAULT_MECHS = ("LiDryer", "drm19", "drm19_inert40", "gri30", "dodecane_lu")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    # per-kernel register / scratch / LDS figures, checked below
    "-Rpass-analysis=kernel-resource-usage",
]

#: Largest private segment (bytes per lane) a kernel may need. The runtime backs scratch for every
#: wave the GPU can hold: 29,568 B per lane (the plain FP32-LU kernel at drm19, n = 22) allocates
#: and runs; 30,528 B (the plain kernel at n = 41) and 64 KB were refused with
#: HSA_STATUS_ERROR_OUT_OF_RESOURCES, which aborts the queue. A build whose kernels need more than
#: the size shown to run fails here, on the CPU.
MAX_SCRATCH_PER_LANE = 29568

_REMARK = re.compile(r"remark:\s+(?:Function Name: (\S+)|([A-Za-z ]+?)(?: \[[^\]]*\])?: "
                     r"(\S+)) \[-Rpass-analysis=kernel-resource-usage\]")


def kernel_resources(stderr: str) -> Dict[str, Dict[str, str]]:
    """Per-kernel resource figures from hipcc's kernel-resource-usage remarks (device pass)."""
    out: Dict[str, Dict[str, str]] = {}
    cur = None
    for m in _REMARK.finditer(stderr):
        if m.group(1):
            cur = out.setdefault(m.group(1), {})
        elif cur is not None:
            cur[m.group(2).strip()] = m.group(3)
    return out


def check_resources(mech: str, res: Dict[str, Dict[str, str]]) -> List[str]:
    """Kernels whose private segment exceeds MAX_SCRATCH_PER_LANE or that need a dynamic stack."""
    bad = []
    for fn, r in res.items():
        if "ScratchSize" not in r:
            continue
        if int(r["ScratchSize"]) > MAX_SCRATCH_PER_LANE or r.get("Dynamic Stack") == "True":
            bad.append(f"{mech}: {fn} scratch {r['ScratchSize']} B/lane, dynamic stack "
                       f"{r.get('Dynamic Stack')}")
    return bad


def resources_path(mech: str) -> str:
    return os.path.join(LIBDIR, f"libpelec_amd_{mech}.resources.txt")


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
    kr = kernel_resources(res.stderr)
    if not kr:
        raise RuntimeError(f"{mech}: no kernel-resource-usage remarks in hipcc's output")
    with open(resources_path(mech), "w") as f:
        for fn in sorted(kr):
            r = kr[fn]
            f.write(f"{fn} VGPRs {r.get('VGPRs')} AGPRs {r.get('AGPRs')} "
                    f"scratch {r.get('ScratchSize')} LDS {r.get('LDS Size')} "
                    f"occupancy {r.get('Occupancy')}\n")
    bad = check_resources(mech, kr)
    if bad:
        os.remove(tmp)
        raise RuntimeError("private segment above the limit the runtime can back:\n" + "\n".join(bad))
    if verbose:
        sys.stderr.write(open(resources_path(mech)).read())
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
