"""lee_update_boundary_noalloc (device arrays, static staging area) against
lee_update_boundary (host arrays, staging allocated by the handle).

Both run the same BC and overset kernels, so the results must be
bit-identical.  Two calls in a row share the device's staging area: the
cylinder case (666 receivers, wall-slip face) first, then the two-block
case with fewer receivers (501), so stale staging values from the first
call would show up in the second."""
import ctypes

import numpy as np
import pytest

import overset_cases
from camille_amd import FP64, PATH_FUSED, PATH_PLAIN, Solver

pytestmark = pytest.mark.gpu


def _solver(name, path):
    if name == "overset_small":
        c = overset_cases.case_overset_small()
        tables = overset_cases.abi_tables(c)
    else:
        c = overset_cases.case_cylinder_small()
        tables = overset_cases.cylinder_abi_tables(c)
    s = Solver([dict(x=b["x"], y=b["y"], mean=b["mean"]) for b in c["blocks"]], mode=c["mode"], value=c["cfl"],
               precision=FP64, path=path, boundaries=c["bc"], overset=tables)
    return c, s


def _hip():
    """The HIP runtime the library itself links (the test allocates the
    caller's device arrays with it; torch's bundled runtime is a different
    library and is not mixed into the process)."""
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc(s, ys):
    hip = _hip()
    h2d, d2h = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost
    host = [np.ascontiguousarray(y, np.float64) for y in ys]
    dev = []
    try:
        for a in host:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            dev.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, h2d) == 0
        ptrs = (ctypes.c_void_p * len(dev))(*[p.value for p in dev])
        rc = s.L.lee_update_boundary_noalloc(s.h, ptrs)
        assert rc == 0, s.L.lee_last_error(s.h).decode()
        s.synchronize()
        out = [np.empty_like(a) for a in host]
        for o, p in zip(out, dev):
            assert hip.hipMemcpy(o.ctypes.data, p, o.nbytes, d2h) == 0
        return out
    finally:
        for p in dev:
            hip.hipFree(p)


@pytest.mark.parametrize("path", [PATH_FUSED, PATH_PLAIN])
def test_noalloc_boundary_update_matches_the_allocating_one(path):
    rng = np.random.default_rng(21)
    for name in ("cylinder_small", "overset_small"):
        c, s = _solver(name, path)
        ys = [rng.standard_normal((5,) + b["x"].shape) for b in c["blocks"]]
        ref = s.update_boundary(ys)
        got = _noalloc(s, ys)
        changed = 0
        for ib in range(len(ys)):
            assert np.array_equal(got[ib], ref[ib]), (name, ib)
            changed += int(np.sum(ref[ib] != ys[ib]))
        assert changed > 0
        print(f"noalloc {name} path {path}: bit-identical, {changed} values changed", flush=True)
        s.close()
