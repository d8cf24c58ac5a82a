"""Discrete adjoint of the fused time advance (no reference counterpart).

Pinned by mathematics rather than by the reference, which has no adjoint:
  * dot products: <L y, k> = <y, L^T k> and <F y, z> = <y, F^T z> for the
    right-hand side (affine part removed: L y - L 0) and the filter, on
    fixture cases with curvilinear metrics, sponges, non-uniform means and
    the 4th- and 8th-order filters;
  * gradients of the observer energy against central finite differences:
    J is exactly quadratic in (q, y0), so central differences carry only
    rounding error;
  * checkpoint-interval invariance (bitwise) and FP32 against FP64;
  * the torch.autograd.Function against finite differences through a
    differentiable embedding of a coarse control.
"""
import math

import numpy as np
import pytest

import cases
from camille_amd import FP32, FP64, LeeError, Solver

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _dot(a, b):
    return float(np.sum(a * b))


def _rand_state(shape, rng):
    y = rng.standard_normal((5,) + shape)
    y[3] = 0.0
    return y


@pytest.mark.parametrize("name", ["filt", "variants", "c1"])
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-13), (FP32, 1e-5)])
def test_filter_transpose_dot_product(name, prec, tol):
    c = cases.ALL[name]()
    s = Solver.from_case(c, precision=prec)
    rng = np.random.default_rng(1)
    y, z = _rand_state(c["x"].shape, rng), _rand_state(c["x"].shape, rng)
    lhs = _dot(s.apply_filter(y), z)
    rhs = _dot(y, s.apply_filter_adjoint(z))
    err = abs(lhs - rhs) / abs(lhs)
    print(f"filter dot {name} prec {prec}: {lhs:.12e} {rhs:.12e} rel {err:.2e}", flush=True)
    assert err <= tol
    s.close()


@pytest.mark.parametrize("name", ["curv", "shear", "jet", "c1", "axisym", "roundjet"])
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_rhs_transpose_dot_product(name, prec, tol):
    """The axisymmetric case adds the term -(gamma pbar v' + gamma vbar p')
    / (r + 1e-99) of the pressure equation; u_r' = v' is 0 on the axis (the
    axis face sets it), as in every state the stages see.  The round-jet
    entropy wave's harmonic sponge is affine: L y - L 0 keeps its linear
    part, -sigma s'."""
    c = cases.ALL[name]()
    if c.get("axisym"):
        # a radial mean velocity as well (0 on the axis), so that both of the
        # term's coefficients, gamma pbar and gamma vbar, enter
        c["mean"] = c["mean"].copy()
        c["mean"][2] = 0.05 * c["y"] * np.exp(-(c["x"] ** 2 + (c["y"] - 3.0) ** 2) / 20.0)
    s = Solver.from_case(c, precision=prec)
    rng = np.random.default_rng(2)
    y, k = _rand_state(c["x"].shape, rng), _rand_state(c["x"].shape, rng)
    if c.get("axisym"):
        # on the axis r = 1e-99: the term's weight on v' there is ~1e99
        # (FP64) or ~6e37 (FP32), and v' = 0 is what makes it vanish; k_p = 0
        # there keeps the FP32 transpose finite
        y[2][0] = 0.0
        k[4][0] = 0.0
    t = 0.37
    Ly = s.compute_rhs(y, t) - s.compute_rhs(np.zeros_like(y), t)
    lhs = _dot(Ly[[0, 1, 2, 4]], k[[0, 1, 2, 4]])
    rhs = _dot(y, s.compute_rhs_adjoint(k))
    err = abs(lhs - rhs) / max(abs(lhs), 1e-300)
    print(f"rhs dot {name} prec {prec}: {lhs:.12e} {rhs:.12e} rel {err:.2e}", flush=True)
    assert err <= tol
    s.close()


@pytest.mark.parametrize("name", ["opt11"])
def test_adjoint_rejects_uncovered_terms(name):
    """The adjoint covers sponges (harmonic-wave zones included), sources,
    holes, overset interpolation, the axisymmetric term, passive scalars
    and the dirichlet_allzero, harmonic-wave, Neumann, wall-slip, Cartesian
    and axisymmetric centerline faces; the 11-point schemes are refused."""
    c = cases.ALL[name]()
    s = Solver.from_case(c)
    with pytest.raises(LeeError):
        s.objective_forward(5, 0, 5)
    s.close()


@pytest.mark.parametrize("name", ["wall", "dirichlet", "centerline", "opt11", "axisym", "entropy", "roundjet"])
def test_boundary_face_transposes_dot_products(name):
    """<B y - B 0, z> = <y, B^T z> for the BC faces of the fixtures
    (wall-slip and Neumann on curvilinear faces, dirichlet_allzero,
    Cartesian centerlines, the axis of an axisymmetric run, the
    harmonic-wave faces, whose data do not depend on the state);
    the opt11 case runs its faces with the standard operator's closure row
    here (the fused path takes OA4)."""
    c = cases.ALL[name]()
    kw = dict(fd_order=4, filter_order=8, filter_strength=1.0) if name == "opt11" else {}
    c = dict(c, **kw)
    s = Solver.from_case(c)
    rng = np.random.default_rng(14)
    y, z = _rand_state(c["x"].shape, rng), _rand_state(c["x"].shape, rng)
    By = s.update_boundary([y])[0] - s.update_boundary([np.zeros_like(y)])[0]
    Btz = s.update_boundary_adjoint([z])[0]
    err = abs(_dot(By, z) - _dot(y, Btz)) / abs(_dot(By, z))
    print(f"face transposes {name}: rel {err:.2e}", flush=True)
    s.close()
    # FP64 round-off of sums of random products (a wrong transpose gives O(1))
    assert err <= 1e-13


def _problem(ni=97, nj=81, shear=True):
    """Curvilinear block, a shear layer u(y) (or uniform flow), sponges, a
    field source and observers on a circle."""
    dx = 0.25
    I, J = np.meshgrid(np.arange(ni), np.arange(nj))
    lx, ly = dx * (ni - 1), dx * (nj - 1)
    x = dx * I - lx / 2 + 0.02 * lx * np.sin(2 * np.pi * J / (nj - 1))
    y = dx * J - ly / 2 + 0.02 * ly * np.sin(2 * np.pi * I / (ni - 1))
    mean = np.zeros((5, nj, ni))
    mean[1] = 0.45 * 0.5 * (1 - np.tanh((np.abs(y) - 3.0) / 0.8)) if shear else 0.3
    mean[4] = 1.0 / GAMMA
    w = 9
    bc = [(0, "sponge_freund_ambient", "+xi", "0", str(w), "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "-xi", f"-{w}", "-0", "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "+eta", "0", "-0", "0", str(w), "0", "0"),
          (0, "sponge_freund_ambient", "-eta", "0", "-0", f"-{w}", "-0", "0", "0")]
    r = np.sqrt(x * x + y * y)
    obs = [(0, i, j, dx * dx) for j in range(nj) for i in range(ni) if abs(r[j, i] - 5.0) < 0.5 * dx]
    q = np.exp(-((x + 2.0) ** 2 + (y - 1.0) ** 2) / 2.0)
    return dict(x=x, y=y, mean=mean, bc=bc, obs=obs, q=q)


def _solver(p, prec):
    s = Solver([dict(x=p["x"], y=p["y"], mean=p["mean"])], boundaries=p["bc"], precision=prec,
               simulation="SOURCE_FIELD", source_params=(1.0, 0.4 * math.pi))
    s.set_observer(p["obs"])
    return s


def _J(s, q, y0, nsteps, window, every):
    s.set_source_field(q)
    s.set_solution(y0)
    s.set_time(0, 0.0, 1e-6)
    return s.objective_forward(nsteps, window, every)


def _grads(s, q, y0, nsteps, window, every):
    J = _J(s, q, y0, nsteps, window, every)
    s.objective_backward(1.0)
    return J, s.source_gradient(), s.state_gradient()


@pytest.mark.parametrize("shear", [True, False])
def test_gradient_matches_central_differences(shear):
    p = _problem(shear=shear)
    s = _solver(p, FP64)
    rng = np.random.default_rng(4)
    shape = p["x"].shape
    y0 = 1e-3 * _rand_state(shape, rng) * np.exp(-(p["x"] ** 2 + p["y"] ** 2) / 8.0)
    nsteps, window, every = 60, 40, 16
    J0, gq, gy = _grads(s, p["q"], y0, nsteps, window, every)
    assert J0 > 0 and np.any(gq != 0) and np.any(gy != 0)
    for trial in range(3):
        e = rng.standard_normal(shape) * np.exp(-(p["x"] ** 2 + p["y"] ** 2) / 20.0)
        eps = 1e-4
        jp = _J(s, p["q"] + eps * e, y0, nsteps, window, every)
        jm = _J(s, p["q"] - eps * e, y0, nsteps, window, every)
        fd = (jp - jm) / (2 * eps)
        ad = _dot(gq, e)
        err = abs(fd - ad) / abs(fd)
        print(f"shear {shear} q-direction {trial}: fd {fd:.12e} adjoint {ad:.12e} rel {err:.2e}", flush=True)
        assert err <= 1e-6
    ey = 1e-3 * _rand_state(shape, rng)
    eps = 1e-2
    jp = _J(s, p["q"], y0 + eps * ey, nsteps, window, every)
    jm = _J(s, p["q"], y0 - eps * ey, nsteps, window, every)
    fd = (jp - jm) / (2 * eps)
    ad = _dot(gy, ey)
    err = abs(fd - ad) / abs(fd)
    print(f"shear {shear} y0-direction: fd {fd:.12e} adjoint {ad:.12e} rel {err:.2e}", flush=True)
    assert err <= 1e-6
    s.close()


def test_axisymmetric_gradient_matches_central_differences():
    """The axisymmetric fixture's box (axis at j = 0 with its centerline
    face, a jet-like mean, sponges) driven by a field source, observers on
    a half circle: the adjoint gradient in the source and initial-state
    directions against central differences (initial states with u_r' = 0
    on the axis, as the axis face leaves every state)."""
    c = cases.ALL["axisym"]()
    x, y = c["x"], c["y"]
    dx = 0.25
    s = Solver([dict(x=x, y=y, mean=c["mean"])], boundaries=c["bc"], precision=FP64, axisym=1,
               simulation="SOURCE_FIELD", source_params=(1.0, 0.4 * math.pi))
    r = np.sqrt(x * x + (y - 0.0) ** 2)
    obs = [(0, i, j, dx * dx) for j in range(x.shape[0]) for i in range(x.shape[1]) if abs(r[j, i] - 6.0) < 0.5 * dx]
    s.set_observer(obs)
    q = np.exp(-((x + 1.0) ** 2 + (y - 2.0) ** 2) / 2.0)
    rng = np.random.default_rng(21)
    y0 = 1e-3 * _rand_state(x.shape, rng) * np.exp(-(x ** 2 + (y - 3.0) ** 2) / 8.0)
    y0[2][0] = 0.0
    nsteps, window, every = 40, 25, 10
    J0, gq, gy = _grads(s, q, y0, nsteps, window, every)
    assert J0 > 0 and np.any(gq != 0) and np.any(gy != 0)
    e = rng.standard_normal(x.shape) * np.exp(-(x ** 2 + (y - 2.0) ** 2) / 20.0)
    eps = 1e-4
    fd = (_J(s, q + eps * e, y0, nsteps, window, every) - _J(s, q - eps * e, y0, nsteps, window, every)) / (2 * eps)
    ad = _dot(gq, e)
    err_q = abs(fd - ad) / abs(fd)
    ey = 1e-3 * _rand_state(x.shape, rng)
    ey[2][0] = 0.0
    eps = 1e-2
    fdy = (_J(s, q, y0 + eps * ey, nsteps, window, every) - _J(s, q, y0 - eps * ey, nsteps, window, every)) / (2 * eps)
    ady = _dot(gy, ey)
    err_y = abs(fdy - ady) / abs(fdy)
    print(f"axisym gradient: q fd {fd:.12e} adjoint {ad:.12e} rel {err_q:.2e}; "
          f"y0 fd {fdy:.12e} adjoint {ady:.12e} rel {err_y:.2e}", flush=True)
    s.close()
    assert err_q <= 1e-6 and err_y <= 1e-6


def test_harmonic_wave_gradient_matches_central_differences():
    """The round-jet entropy wave fixture (dirichlet_harmonicwave face and
    harmonic sponge on +xi, a jet-like mean, a wavy grid) driven by a field
    source as well, observers on a circle: J is quadratic in (q, y0) with
    the wave's response as an affine part, and the adjoint gradient matches
    central differences in the source and initial-state directions."""
    c = cases.ALL["roundjet"]()
    x, y = c["x"], c["y"]
    s = Solver([dict(x=x, y=y, mean=c["mean"])], boundaries=c["bc"], precision=FP64, harmonic=c["harmonic"],
               simulation="SOURCE_FIELD", source_params=(1.0, 0.4 * math.pi))
    # observers, source and initial pulse next to the wave's face, so that
    # the harmonic sponge (x > 8.7) and face lie in the adjoint's reach
    r = np.sqrt((x - 6.0) ** 2 + y ** 2)
    dx = 0.25
    obs = [(0, i, j, dx * dx) for j in range(x.shape[0]) for i in range(x.shape[1]) if abs(r[j, i] - 3.0) < 0.6 * dx]
    assert len(obs) > 20
    s.set_observer(obs)
    q = np.exp(-((x - 5.0) ** 2 + (y - 1.0) ** 2) / 2.0)
    rng = np.random.default_rng(23)
    y0 = 1e-3 * _rand_state(x.shape, rng) * np.exp(-((x - 7.0) ** 2 + y ** 2) / 8.0)
    nsteps, window, every = 40, 25, 10
    J0, gq, gy = _grads(s, q, y0, nsteps, window, every)
    assert J0 > 0 and np.any(gq != 0) and np.any(gy != 0)
    e = rng.standard_normal(x.shape) * np.exp(-((x - 5.0) ** 2 + y ** 2) / 20.0)
    eps = 1e-4
    fd = (_J(s, q + eps * e, y0, nsteps, window, every) - _J(s, q - eps * e, y0, nsteps, window, every)) / (2 * eps)
    ad = _dot(gq, e)
    err_q = abs(fd - ad) / abs(fd)
    ey = 1e-3 * _rand_state(x.shape, rng)
    eps = 1e-2
    fdy = (_J(s, q, y0 + eps * ey, nsteps, window, every) - _J(s, q, y0 - eps * ey, nsteps, window, every)) / (2 * eps)
    ady = _dot(gy, ey)
    err_y = abs(fdy - ady) / abs(fdy)
    print(f"harmonic-wave gradient: q fd {fd:.12e} adjoint {ad:.12e} rel {err_q:.2e}; "
          f"y0 fd {fdy:.12e} adjoint {ady:.12e} rel {err_y:.2e}", flush=True)
    s.close()
    assert err_q <= 1e-6 and err_y <= 1e-6


def test_passive_scalars_leave_the_gradient_unchanged():
    """LEE_SCALAR: the scalars feed nothing back into s', u', v', p', so J
    and its gradients (q and the five LEE fields of y0) are those of the
    same run without them, bit for bit; a central difference checks the
    run with scalars on its own.  The scalar fixture's faces (wall-slip,
    Neumann, dirichlet_allzero) and sponges, a field source, observers on a
    circle."""
    c = cases.ALL["scalar"]()
    x, y = c["x"], c["y"]
    xc, yc = float(x.mean()), float(y.mean())
    dx = float(np.hypot(x[0, 1] - x[0, 0], y[0, 1] - y[0, 0]))
    r = np.hypot(x - xc, y - yc)
    obs = [(0, i, j, dx * dx) for j in range(x.shape[0]) for i in range(x.shape[1]) if abs(r[j, i] - 5.0) < 0.6 * dx]
    assert len(obs) > 20
    q = np.exp(-((x - xc - 1.0) ** 2 + (y - yc) ** 2) / 2.0)
    rng = np.random.default_rng(29)
    y0 = 1e-3 * rng.standard_normal((5 + c["num_scalar"],) + x.shape) * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / 8.0)
    y0[3] = 0.0
    nsteps, window, every = 40, 25, 10
    out = {}
    for ns in (c["num_scalar"], 0):
        s = Solver([dict(x=x, y=y, mean=c["mean"][:5 + ns])], boundaries=c["bc"], precision=FP64, num_scalar=ns,
                   simulation="SOURCE_FIELD", source_params=(1.0, 0.4 * math.pi))
        s.set_observer(obs)
        out[ns] = _grads(s, q, y0[:5 + ns], nsteps, window, every)
        if ns:
            e = rng.standard_normal(x.shape) * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / 20.0)
            eps = 1e-4
            fd = (_J(s, q + eps * e, y0, nsteps, window, every) - _J(s, q - eps * e, y0, nsteps, window, every)) / (2 * eps)
            err = abs(fd - _dot(out[ns][1], e)) / abs(fd)
            print(f"scalars: J {out[ns][0]:.12e}, q-direction fd {fd:.12e} rel {err:.2e}", flush=True)
            assert err <= 1e-6
        s.close()
    a, b = out[c["num_scalar"]], out[0]
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    print("scalars: J and gradients bit-identical to the run without them", flush=True)


def test_checkpoint_interval_does_not_change_the_gradient():
    p = _problem()
    s = _solver(p, FP64)
    y0 = np.zeros((5,) + p["x"].shape)
    ref = None
    for every in (1, 16, 60):
        J, gq, gy = _grads(s, p["q"], y0, 60, 40, every)
        if ref is None:
            ref = (J, gq, gy)
        assert J == ref[0]
        assert np.array_equal(gq, ref[1]) and np.array_equal(gy, ref[2]), every
    # the state after the reverse pass is the end state of the forward run
    s2 = _solver(p, FP64)
    s2.set_source_field(p["q"])
    s2.set_solution(y0)
    s2.set_time(0, 0.0, 1e-6)
    s2.solve(60)
    assert np.array_equal(s2.solution(), s.solution())
    assert s2.clock() == s.clock()
    s.close()
    s2.close()


def test_fp32_adjoint_against_fp64():
    p = _problem()
    y0 = np.zeros((5,) + p["x"].shape)
    out = {}
    for prec in (FP64, FP32):
        s = _solver(p, prec)
        out[prec] = _grads(s, p["q"], y0, 60, 40, 16)
        s.close()
    errJ = abs(out[FP32][0] - out[FP64][0]) / out[FP64][0]
    g64, g32 = out[FP64][1], out[FP32][1]
    errg = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    print(f"FP32 vs FP64: J rel {errJ:.2e}, dJ/dq rel {errg:.2e}", flush=True)
    assert errJ <= 1e-4 and errg <= 1e-4


def test_autograd_function_matches_central_differences():
    import torch

    from camille_amd.adjoint import embed_control, observer_energy
    p = _problem()
    s = _solver(p, FP64)
    shape = p["x"].shape
    rng = np.random.default_rng(9)
    a = torch.tensor(rng.uniform(0.5, 1.5, (8, 8)), dtype=torch.float64, requires_grad=True)
    y0 = torch.zeros((5,) + shape, dtype=torch.float64)
    nsteps, window, every = 50, 30, 10

    def J_of(av):
        return observer_energy(s, embed_control(av, shape), y0, nsteps, window, every)

    J = J_of(a)
    J.backward()
    g = a.grad.detach().numpy().copy()
    for trial in range(2):
        e = torch.tensor(rng.standard_normal((8, 8)), dtype=torch.float64)
        eps = 1e-4
        with torch.no_grad():
            fd = (J_of(a + eps * e).item() - J_of(a - eps * e).item()) / (2 * eps)
        ad = float(np.sum(g * e.numpy()))
        err = abs(fd - ad) / abs(fd)
        print(f"autograd direction {trial}: fd {fd:.12e} adjoint {ad:.12e} rel {err:.2e}", flush=True)
        assert err <= 1e-6
    s.close()


def test_autograd_two_forwards_then_one_backward():
    """J1 + J2 from two forward runs on one solver (different controls), then
    one backward: each term gets the gradient of its own run, as separate
    backward passes on fresh forward runs give them."""
    import torch

    from camille_amd.adjoint import embed_control, observer_energy
    p = _problem()
    s = _solver(p, FP64)
    shape = p["x"].shape
    rng = np.random.default_rng(11)
    a1 = torch.tensor(rng.uniform(0.5, 1.5, (8, 8)), dtype=torch.float64, requires_grad=True)
    a2 = torch.tensor(rng.uniform(-1.0, 1.0, (8, 8)), dtype=torch.float64, requires_grad=True)
    y0 = torch.zeros((5,) + shape, dtype=torch.float64)
    nsteps, window, every = 40, 20, 10

    def J_of(av):
        return observer_energy(s, embed_control(av, shape), y0, nsteps, window, every)

    sep = []
    for a in (a1, a2):
        J_of(a).backward()
        sep.append(a.grad.detach().numpy().copy())
        a.grad = None
    (J_of(a1) + J_of(a2)).backward()
    for k, a in enumerate((a1, a2)):
        g = a.grad.detach().numpy()
        err = float(np.max(np.abs(g - sep[k])) / np.max(np.abs(sep[k])))
        print(f"two forwards, control {k}: gradient rel diff {err:.2e}", flush=True)
        assert err <= 1e-13
    s.close()


def test_repeated_observer_points_add_their_weights():
    """An observer list that names a point twice with weights w1, w2 gives the
    J and gradient of the point named once with w1 + w2 (the duplicates are
    merged when the observers are set), and sampling still returns p' for
    every entry of the caller's list."""
    p = _problem()
    y0 = np.zeros((5,) + p["x"].shape)
    obs = p["obs"][:40]
    dup = [obs[3][:3] + (0.25 * obs[3][3],), obs[17][:3] + (0.5 * obs[17][3],)]
    once = list(obs)
    once[3] = obs[3][:3] + (1.25 * obs[3][3],)
    once[17] = obs[17][:3] + (1.5 * obs[17][3],)
    out = []
    for pts in (obs + dup, once):
        s = _solver(p, FP64)
        s.set_observer(pts)
        J, gq, gy = _grads(s, p["q"], y0, 40, 20, 10)
        samples = s.sample_observers()
        out.append((J, gq, samples))
        s.close()
    (Ja, ga, sa), (Jb, gb, sb) = out
    assert abs(Ja - Jb) <= 1e-14 * abs(Jb)
    assert float(np.max(np.abs(ga - gb))) <= 1e-14 * float(np.max(np.abs(gb)))
    assert len(sa) == len(obs) + 2 and len(sb) == len(obs)
    assert sa[len(obs)] == sa[3] and sa[len(obs) + 1] == sa[17]
    assert np.array_equal(sa[:len(obs)], sb)


# ----------------------------------------------------------- overset, BCs
# The two overset fixtures: holes (per-point closure codes), interpolated
# fringes, for the cylinder an O-grid seam interpolated from its own block,
# a wall-slip face and a potential-flow mean.

def _overset_adjoint_solver(name, prec):
    import overset_cases
    if name == "overset_small":
        c = overset_cases.case_overset_small()
        tables = overset_cases.abi_tables(c)
    else:
        c = overset_cases.case_cylinder_small()
        tables = overset_cases.cylinder_abi_tables(c)
    s = Solver([dict(x=b["x"], y=b["y"], mean=b["mean"]) for b in c["blocks"]], mode=c["mode"], value=c["cfl"],
               precision=prec, boundaries=c["bc"], overset=tables, simulation="SOURCE_FIELD",
               source_params=(1.0, 0.6 * math.pi))
    return c, s


@pytest.mark.parametrize("name", ["overset_small", "cylinder_small"])
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-13), (FP32, 1e-5)])
def test_overset_transposes_dot_products(name, prec, tol):
    """<B y, z> = <y, B^T z> for the boundary update (BC faces, then overset
    interpolation, on every block), and the right-hand side and filter
    transposes on blocks with holes."""
    c, s = _overset_adjoint_solver(name, prec)
    rng = np.random.default_rng(12)
    shapes = [b["x"].shape for b in c["blocks"]]
    ys = [_rand_state(sh, rng) for sh in shapes]
    zs = [_rand_state(sh, rng) for sh in shapes]
    By = s.update_boundary(ys)
    Btz = s.update_boundary_adjoint(zs)
    lhs = sum(_dot(By[k], zs[k]) for k in range(len(ys)))
    rhs = sum(_dot(ys[k], Btz[k]) for k in range(len(ys)))
    eb = abs(lhs - rhs) / abs(lhs)
    errs = [eb]
    for ib, sh in enumerate(shapes):
        a, k = _rand_state(sh, rng), _rand_state(sh, rng)
        La = s.compute_rhs(a, 0.3, ib) - s.compute_rhs(np.zeros_like(a), 0.3, ib)
        l1, r1 = _dot(La, k), _dot(a, s.compute_rhs_adjoint(k, ib))
        l2, r2 = _dot(s.apply_filter(a, ib), k), _dot(a, s.apply_filter_adjoint(k, ib))
        errs += [abs(l1 - r1) / abs(l1), abs(l2 - r2) / abs(l2)]
    print(f"{name} prec {prec}: boundary {eb:.2e}, rhs/filter per block {['%.2e' % e for e in errs[1:]]}",
          flush=True)
    s.close()
    assert max(errs) <= tol


@pytest.mark.parametrize("name,nsteps,window,every", [("overset_small", 30, 15, 10), ("cylinder_small", 40, 20, 10)])
def test_overset_gradient_matches_central_differences(name, nsteps, window, every):
    """Observer energy on fluid points of the background block, gradient
    with respect to a source field on both blocks, through holes, fringes,
    the O-grid seam and the slip wall; central differences in FP64."""
    c, s = _overset_adjoint_solver(name, FP64)
    rng = np.random.default_rng(13)
    ib0 = s.iblank(0)
    x0, y0_ = c["blocks"][0]["x"], c["blocks"][0]["y"]
    cand = np.flatnonzero((ib0 == 0).reshape(-1))
    pick = rng.choice(cand, 40, replace=False)
    s.set_observer_points(np.zeros(40), pick, np.full(40, 1.0))
    qs = [np.exp(-((b["x"] - 0.5) ** 2 + (b["y"] - 0.3) ** 2) / 0.5) * (s.iblank(k) != 1)
          for k, b in enumerate(c["blocks"])]
    ics = [b["ic"] for b in c["blocks"]]

    def J(qq):
        for k in range(2):
            s.set_source_field(qq[k], k)
            s.set_solution(ics[k], k)
        s.set_time(0, 0.0, 1e-6)
        return s.objective_forward(nsteps, window, every)

    J0 = J(qs)
    s.objective_backward(1.0)
    g = [s.source_gradient(k) for k in range(2)]
    assert J0 > 0 and all(np.any(gk != 0) for gk in g)
    for trial in range(2):
        e = [rng.standard_normal(q.shape) * (s.iblank(k) != 1) for k, q in enumerate(qs)]
        eps = 1e-3
        fd = (J([qs[k] + eps * e[k] for k in range(2)]) - J([qs[k] - eps * e[k] for k in range(2)])) / (2 * eps)
        ad = sum(_dot(g[k], e[k]) for k in range(2))
        err = abs(fd - ad) / abs(fd)
        print(f"{name} gradient direction {trial}: J {J0:.6e} fd {fd:.12e} adjoint {ad:.12e} rel {err:.2e}",
              flush=True)
        assert err <= 1e-6
    s.close()
