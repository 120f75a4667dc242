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
