"""The 3-D shear-layer adjoint workload at its full size, 256^3 (no reference
counterpart: the reference stops at the 3-D Jacobian, metrics.cpp:290-292).

Grid, mean flow, observer cylinder and control embedding are those of the
benchmark's 3-D leg (bench.c4_inputs): 256^3 Cartesian on [-2.5, 2.5]^3,
u(r) = 0.45 (1 - tanh((r - 0.5) / 0.05)), 16-point Freund sponges on all
six faces, observers within half a cell of r = 1.5, a 32^3
source-amplitude control embedded trilinearly.  The horizon
is shortened to 60 steps (window 41..60, checkpoints every 20) so that the
shear layer's instability has not yet amplified round-off:

  * transposes: <L a, k> = <a, L^T k> and <F a, k> = <a, F^T k> at 256^3,
    FP64 to 1e-12 and FP32 to 1e-5;
  * the FP64 adjoint gradient against a central difference (eps 1e-4) of J
    along a random control direction, to 1e-6;
  * FP32 against FP64 over the same run: J to 1e-5, the control gradient
    to 1e-4, p' after 60 steps to 1e-5 inside the sponges and to 1e-4 over
    the whole box.  The box's corner points take one-sided closure rows and
    identity filter rows in all three directions; FP32 round-off, like a
    6e-8 relative perturbation of the source in FP64, reaches ~5e-5 of max
    |p'| there and stays at that level.
"""
import math

import numpy as np
import pytest

from camille_amd import FP32, FP64, Solver

pytestmark = pytest.mark.gpu

N = 256
NSTEPS, WINDOW, EVERY = 60, 40, 20


@pytest.fixture(scope="module")
def shear3d():
    import bench
    return bench.c4_inputs(N)


def _solver(inp, prec):
    s = Solver([dict(x=inp["x"], y=inp["y"], z=inp["z"], mean=inp["mean"])], num_dim=3, precision=prec,
               simulation="SOURCE_FIELD", source_params=(1e-3, 2.0 * math.pi * 0.5), mode="FIX_CFL", value=0.5,
               boundaries=inp["bc"])
    s.set_observer_points(np.zeros(inp["obs"].size), inp["obs"], np.full(inp["obs"].size, inp["h"] ** 2))
    return s


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_shear3d_transposes_at_full_size(shear3d, prec, tol):
    s = _solver(shear3d, prec)
    shape = shear3d["x"].shape
    rng = np.random.default_rng(41)
    a = rng.standard_normal((5,) + shape)
    k = rng.standard_normal((5,) + shape)
    t = 0.3
    La = s.compute_rhs(a, t) - s.compute_rhs(np.zeros_like(a), t)
    e1 = abs(float(np.sum(La * k)) - float(np.sum(a * s.compute_rhs_adjoint(k)))) / abs(float(np.sum(La * k)))
    del La
    Fa_k = float(np.sum(s.apply_filter(a) * k))
    e2 = abs(Fa_k - float(np.sum(a * s.apply_filter_adjoint(k)))) / abs(Fa_k)
    print(f"shear3d 256^3 dot products prec {prec}: rhs {e1:.2e}, filter {e2:.2e}", flush=True)
    s.close()
    assert e1 <= tol and e2 <= tol


def _control(seed):
    import torch
    rng = np.random.default_rng(seed)
    a0 = torch.tensor(rng.uniform(0.5, 1.5, (32, 32, 32)), dtype=torch.float64)
    e = torch.tensor(rng.standard_normal((32, 32, 32)), dtype=torch.float64)
    return a0, e


def test_shear3d_gradient_and_fp32_at_full_size(shear3d):
    import torch

    from camille_amd.adjoint import embed_control3, observer_energy
    shape = shear3d["x"].shape
    a0, e = _control(4)
    y0 = torch.zeros((5,) + shape, dtype=torch.float64)
    out = {}
    for label, prec in (("fp64", FP64), ("fp32", FP32)):
        s = _solver(shear3d, prec)
        a = a0.clone().requires_grad_(True)
        J = observer_energy(s, embed_control3(a, shape), y0, NSTEPS, WINDOW, EVERY)
        p_end = s.solution()[4].copy()  # the state after the forward run
        J.backward()
        out[label] = dict(J=float(J.item()), g=a.grad.detach().numpy().copy(), p=p_end, s=s)
        print(f"shear3d {label}: J {J.item():.10e} after {NSTEPS} steps", flush=True)
    s = out["fp64"]["s"]
    eps = 1e-4
    with torch.no_grad():
        jp = observer_energy(s, embed_control3(a0 + eps * e, shape), y0, NSTEPS, WINDOW, EVERY).item()
        jm = observer_energy(s, embed_control3(a0 - eps * e, shape), y0, NSTEPS, WINDOW, EVERY).item()
    fd = (jp - jm) / (2 * eps)
    ad = float(np.sum(out["fp64"]["g"] * e.numpy()))
    err_fd = abs(fd - ad) / abs(fd)
    p64, p32 = out["fp64"]["p"], out["fp32"]["p"]
    scale = float(np.max(np.abs(p64)))
    err_p = float(np.max(np.abs(p32 - p64))) / scale
    w = shear3d["sponge"]
    inner = (slice(w, -w),) * 3
    err_pi = float(np.max(np.abs(p32[inner] - p64[inner]))) / scale
    err_J = abs(out["fp32"]["J"] - out["fp64"]["J"]) / out["fp64"]["J"]
    g64, g32 = out["fp64"]["g"], out["fp32"]["g"]
    err_g = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    print(f"shear3d 256^3: FD {fd:.12e} adjoint {ad:.12e} rel {err_fd:.2e}; FP32 vs FP64: p inside the sponges "
          f"{err_pi:.2e}, whole box {err_p:.2e}, J {err_J:.2e}, gradient {err_g:.2e}", flush=True)
    for v in out.values():
        v["s"].close()
    assert err_fd <= 1e-6
    assert err_pi <= 1e-5 and err_p <= 1e-4 and err_J <= 1e-5
    assert err_g <= 1e-4
