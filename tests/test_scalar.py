"""Passive scalars (PHYSICAL_MODEL = LEE_SCALAR) on the GPU, through the C
ABI, against the oracle (pinned bit for bit to the reference's scalar
fixture in test_oracle.py).

The golden `scalar` case itself runs in test_gpu.py's fixture tests; here
the unit operators and the scalar's path through the other features: the
optimized 11-point schemes, wall-slip faces, and overset interpolation with
holes.  Bars as elsewhere: plain FP64 bit-identical, fused FP64 1e-12,
fused FP32 1e-5 (relative to each field's max magnitude).
"""
import math

import numpy as np
import pytest

import cases
import overset_cases
from camille_amd import FP32, FP64, PATH_FUSED, PATH_PLAIN, Solver
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

PATHS = (("plain", FP64, PATH_PLAIN, 0.0), ("fused64", FP64, PATH_FUSED, 1e-12), ("fused32", FP32, PATH_FUSED, 1e-5))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _with_scalar(mean, ic, x, y):
    """One passive scalar: a tanh layer as its mean (so u' . grad Zbar acts)
    and a Gaussian blob of amplitude 0.01 as its initial value."""
    zbar = 0.5 * (1.0 + np.tanh((y + 0.3 * x) / 2.0))
    blob = 0.01 * np.exp(-math.log(2.0) * ((x - 1.5) ** 2 + (y + 1.0) ** 2) / 1.5 ** 2)
    return np.concatenate([mean, zbar[None]]), np.concatenate([ic, blob[None]])


def _state(c, seed):
    """A smooth state of every variable (w' = 0) for the unit operators."""
    x, y = c["x"], c["y"]
    rng = np.random.default_rng(seed)
    out = np.zeros((5 + c["num_scalar"],) + x.shape)
    for v in [0, 1, 2, 4] + list(range(5, out.shape[0])):
        a, b, ph = rng.uniform(0.05, 0.4, 3)
        out[v] = 1e-3 * np.sin(a * x + ph) * np.cos(b * y - ph) * np.exp(-(x * x + y * y) / 400.0)
    return out + 1e-5 * rng.standard_normal(out.shape) * (np.arange(out.shape[0]) != 3)[:, None, None]


def test_scalar_rhs_and_filter_unit_parity():
    """compute_RHS and apply_filter of a LEE_SCALAR state: the scalar rows
    (ubar . grad Z' + u' . grad Zbar, wall-slip and Dirichlet faces do not
    enter the unit calls) and the filtered scalars match the oracle."""
    c = cases.case_scalar()
    o = Oracle.from_case(c)
    y = _state(c, 0)
    t = 0.41
    ref = o.rhs(y[:, None], t)[:, 0]
    fref = o.filter(y[:, None])[:, 0]
    act = [0, 1, 2, 4, 5, 6]
    unit_tol = {"plain": (0.0, 0.0), "fused64": (1e-13, 1e-14), "fused32": (1e-5, 1e-6)}
    for label, prec, path, _ in PATHS:
        s = Solver.from_case(c, precision=prec, path=path)
        r = s.compute_rhs(y, t)
        f = s.apply_filter(y)
        s.close()
        er = [_rel(r[v], ref[v]) for v in act]
        ef = [_rel(f[v], fref[v]) for v in act]
        print(f"scalar unit {label}: rhs {max(er):.2e} (scalars {er[4]:.2e}, {er[5]:.2e}), filter {max(ef):.2e}",
              flush=True)
        assert max(er) <= unit_tol[label][0] and max(ef) <= unit_tol[label][1]


@pytest.mark.parametrize("base", ["opt11", "wall", "centerline"])
def test_scalar_with_other_features(base):
    """A scalar added to the FDO11P/SFO11P case (its rows and closures,
    Neumann and wall faces), the wall-slip channel and the centerline case:
    every path against the oracle after the case's steps."""
    c = cases.ALL[base]()
    c["mean"], c["ic"] = _with_scalar(c["mean"], c["ic"], c["x"], c["y"])
    c["num_scalar"] = 1
    o = Oracle.from_case(c)
    o.run(c["nsteps"])
    ref = o.solution()[:, 0]
    for label, prec, path, tol in PATHS:
        s = Solver.from_case(c, precision=prec, path=path)
        s.solve(c["nsteps"])
        sol = s.solution()
        assert s.time == o.time
        s.close()
        if tol == 0.0:
            assert np.array_equal(sol, ref), (base, label)
            print(f"scalar {base} {label}: bit-identical to the oracle", flush=True)
            continue
        errs = [_rel(sol[v], ref[v]) for v in (0, 1, 2, 4, 5)]
        print(f"scalar {base} {label}: rel errors {['%.2e' % e for e in errs]}", flush=True)
        assert max(errs) <= tol


def test_scalar_through_overset_and_holes():
    """A scalar across the two-block overset case (holes, interpolated
    fringe, fluid-to-hole rows): interpolated like every variable
    (overset.cpp:886-974) and zeroed in the holes; 30 steps."""
    c = overset_cases.case_overset_small()
    for b in c["blocks"]:
        b["mean"], b["ic"] = _with_scalar(b["mean"], b["ic"], b["x"], b["y"])
    scheme = {k: c[k] for k in ("fd_order", "filter_order", "filter_blend", "rk", "sponge_order", "sponge_constant")}
    tables = overset_cases.abi_tables(c)
    o = Oracle(c["blocks"], scheme, c["bc"], None, c["mode"], c["cfl"], overset=tables, num_scalar=1)
    nsteps = 30
    o.run(nsteps)
    for label, prec, path, tol in PATHS:
        s = Solver([dict(x=b["x"], y=b["y"], mean=b["mean"]) for b in c["blocks"]], mode=c["mode"], value=c["cfl"],
                   precision=prec, path=path, boundaries=c["bc"], overset=tables, num_scalar=1)
        for ib, b in enumerate(c["blocks"]):
            s.set_solution(b["ic"], ib)
        s.solve(nsteps)
        for ib in range(2):
            sol, ref = s.solution(ib), o.solution(ib)[:, 0]
            holes = s.iblank(ib) == 1
            assert holes.sum() == 0 or np.all(sol[5][holes] == 0.0)
            if tol == 0.0:
                assert np.array_equal(sol, ref), (label, ib)
            else:
                errs = [_rel(sol[v], ref[v]) for v in (0, 1, 2, 4, 5)]
                print(f"scalar overset {label} block {ib}: rel errors {['%.2e' % e for e in errs]}", flush=True)
                assert max(errs) <= tol
        s.close()
    print("scalar overset: plain bit-identical to the oracle", flush=True)
