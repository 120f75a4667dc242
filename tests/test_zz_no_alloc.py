"""prc_react_flat_noalloc: the react_flat variant whose workspace, argument block and counters
live in the library's static device array, checked through the C-ABI against prc_react_flat (the
batched form with its own allocations) on two calls in a row, the second with fewer cells, and
the first of them against the FP64 oracle as well. Named to be collected last: it must pass
after every other GPU test, and alone in a fresh process."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pelec_amd import inputs, mechanism as M  # noqa: E402
from pelec_amd.reactor import Reactor  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


@pytest.mark.parametrize("lu_fp32", [True, False])
def test_noalloc_matches_react_flat(lu_fp32, oracle_lib):
    W = inputs.WORKLOADS["drm19_64"]
    # 20000 cells need more slots than the scratch array holds (~13000 at drm19), so the first
    # call also refills; the second, smaller call reuses the array left by the first
    sizes = (20000, 3000)
    d = inputs.make_inputs(W.name, sizes[0])
    mask = (inputs.uniform(91, 0, np.arange(sizes[0])) >= 0.1).astype(np.int32)
    r = Reactor(W.mech)
    r.configure(rtol=W.rtol, atol=W.atol, form="batched", lu_fp32=lu_fp32)
    for call, n in enumerate(sizes):
        outs = []
        for variant in ("react_flat", "noalloc"):
            rY, T, rE = _dev(d["rY"][:, :n]), _dev(d["T"][:n]), _dev(d["rE"][:n])
            FC = torch.zeros_like(T)
            m = _dev(mask[:n], torch.int32)
            print(f"[gpu-test] call {call} n={n} {variant} lu_fp32={lu_fp32}", flush=True)
            fn = r.react_flat if variant == "react_flat" else r.react_flat_noalloc
            tnew = fn(rY, T, rE, W.dt, FC=FC, mask=m, time=1.0)
            torch.cuda.synchronize()
            assert abs(tnew - (1.0 + W.dt)) < 1e-15
            assert r.last_failures() == 0
            outs.append({"rY": rY.cpu().numpy(), "T": T.cpu().numpy(), "rE": rE.cpu().numpy(),
                         "FC": FC.cpu().numpy()})
        # same kernels, same per-cell arithmetic: bit-identical results
        for k in ("rY", "T", "rE", "FC"):
            np.testing.assert_array_equal(outs[1][k], outs[0][k])
        off = mask[:n] == 0
        np.testing.assert_array_equal(outs[1]["T"][off], d["T"][:n][off])
        assert (outs[1]["FC"][off] == 0).all() and (outs[1]["FC"][~off] > 0).all()
        if call == 0:
            idx = np.flatnonzero(~off)[::40]
            ref = oracle_lib.OracleMech(M.load(W.mech)).react(
                d["rY"][:, idx], d["T"][idx], d["rE"][idx], W.dt, rtol=W.rtol, atol=W.atol,
                nthreads=8)
            dT = np.abs(outs[1]["T"][idx] - ref["T"]).max()
            dY = (np.abs(outs[1]["rY"][:, idx] - ref["rY"]) / ref["rY"].sum(0)).max()
            print(f"[gpu-test]   vs oracle on {idx.size} cells: max|dT| {dT:.2e} K, "
                  f"max|dY| {dY:.2e}", flush=True)
            assert dT <= 1e-3 and dY <= 1e-8
    r.close()


def test_noalloc_rejects_bad_arguments():
    """Errors come back as status codes: null state pointers and ReactorRK64 (no batched form)
    give PRC_EINVAL; nothing is launched."""
    import ctypes
    r = Reactor("drm19")
    r.configure(form="batched")
    dt, tm = ctypes.c_double(1e-6), ctypes.c_double(0.0)
    rc = r.lib.prc_react_flat_noalloc(r.h, None, None, None, None, None, None, None, 5,
                                      ctypes.byref(dt), ctypes.byref(tm), None)
    assert rc == -1 and tm.value == 0.0
    rk = Reactor("drm19", "ReactorRK64")
    n = 64
    with pytest.raises(Exception):
        rk.react_flat_noalloc(torch.ones((21, n), dtype=torch.float64, device=DEV),
                              torch.ones(n, dtype=torch.float64, device=DEV),
                              torch.ones(n, dtype=torch.float64, device=DEV), 1e-6)
    r.close()
    rk.close()
