"""CPU checks of the drop-in boundary: every library loads and exports exactly what
include/*.h declares (no GPU calls are made here)."""
import ctypes
import os
import re

import pytest

from pelec_amd import build, reactor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(prc_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_boundary():
    names = declared_functions()
    for required in ("prc_create", "prc_init", "prc_set_typ_vals_ode", "prc_react_box", "prc_close",
                     "prc_react_flat", "prc_react_boxes"):
        assert required in names


@pytest.mark.parametrize("mech", build.DEFAULT_MECHS)
def test_library_exports_all_declared_symbols(mech):
    path = build.lib_path(mech)
    assert os.path.exists(path), f"{path} not built"
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes binding covers every declared function
    assert declared_functions() <= set(reactor._SIGS)


@pytest.mark.parametrize("mech", build.DEFAULT_MECHS)
def test_mechanism_metadata(mech):
    from pelec_amd import mechanism as M
    lib = reactor.load_library(mech)
    m = M.load(mech)
    assert lib.prc_num_species() == m.nsp
    assert lib.prc_num_reactions() == m.nrxn
    assert lib.prc_mechanism_name().decode() == m.name
    assert lib.prc_mechanism_sha256().decode() == m.source_sha256
    assert [lib.prc_species_name(k).decode() for k in range(m.nsp)] == [s.name for s in m.species]
    assert lib.prc_species_name(m.nsp) is None


def test_argument_validation_without_gpu():
    lib = reactor.load_library("LiDryer")
    h = ctypes.c_void_p()
    assert lib.prc_create(b"NoSuchReactor", ctypes.byref(h)) == -1
    assert lib.prc_create(b"ReactorCvode", None) == -1
    assert lib.prc_create(b"ReactorCvode", ctypes.byref(h)) == 0
    assert lib.prc_set_param_int(h, b"cvode.max_order", 9) == -1
    assert lib.prc_set_param_int(h, b"cvode.max_order", 3) == 0
    assert lib.prc_set_param_real(h, b"ode.rtol", -1.0) == -1
    assert lib.prc_set_param_real(h, b"no.such.key", 1.0) == -1
    # react before init is a state error, not a crash
    dt = ctypes.c_double(1e-6)
    tm = ctypes.c_double(0.0)
    assert lib.prc_react_flat(h, None, None, None, None, None, None, None, 4, ctypes.byref(dt),
                              ctypes.byref(tm), None) == -3
    assert lib.prc_destroy(h) == 0


SHIM_HARNESS = r"""
// AMReX-shaped stand-ins: the shim binds to any Box / Array4 with these members.
#include <cstdio>
#include "pelec_amd_reactor.H"
struct Dim3 { int x, y, z; };
struct Box {
  int lo[3], hi[3];
  int smallEnd(int d) const { return lo[d]; }
  int bigEnd(int d) const { return hi[d]; }
};
template <class T> struct Array4 {
  T* p; long jstride, kstride, nstride; Dim3 begin, end; int ncomp;
};
int main() {
  pelec_amd::Reactor* r = pelec_amd::Reactor::create("ReactorCvode");
  r->set("ode.rtol", 1e-6);
  r->set("cvode.max_order", 5);
  static double buf[4 * 4 * 4 * 32];
  static int msk[4 * 4 * 4];
  Box bx{{0, 0, 0}, {3, 3, 3}};
  Array4<double> a{buf, 4, 16, 64, {0, 0, 0}, {4, 4, 4}, 1};
  Array4<int> m{msk, 4, 16, 64, {0, 0, 0}, {4, 4, 4}, 1};
  double dt = 1e-6, time = 0.0;
  // react before init(1, 1) is a state error (no device work is attempted)
  const int rc = r->react(bx, a, a, a, a, a, a, m, dt, time);
  std::printf("%d %d\n", rc, prc_num_species());
  bool threw = false;
  try {
    r->set("no.such.key", 1);
  } catch (const std::exception&) {
    threw = true;
  }
  delete r;
  return (rc == PRC_ESTATE && threw) ? 0 : 1;
}
"""


def test_cxx_shim_compiles_and_forwards(tmp_path):
    """include/pelec_amd_reactor.H: the ReactorBase-shaped C++ class binds AMReX-shaped types and
    forwards to the library (host-only calls, no GPU)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "shim.cpp"
    src.write_text(SHIM_HARNESS)
    lib = build.lib_path("drm19")
    exe = tmp_path / "shim"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src),
                    lib, f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split() == ["-3", "21"]
