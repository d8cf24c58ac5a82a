"""3-D fused path (the reference never runs 3-D: its 3-D Jacobian exits,
metrics.cpp:292).  Pinned against the 2-D path, which is pinned against the
reference, and by symmetry and exactness properties:

  * a z-invariant problem on an extruded grid reproduces the 2-D solution
    in every plane (FP64 1e-12, FP32 1e-5);
  * exchanging the x and z axes of a problem exchanges the axes of the
    solution (the xi, eta, zeta code paths agree);
  * the 3-D derivative is exact on quadratics (closure rows included).
"""
import math

import numpy as np
import pytest

import cases
from camille_amd import FP32, FP64, LeeError, Solver

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _sponge3(w, sides=("+xi", "-xi", "+eta", "-eta")):
    w -= 1
    spec = {"+xi": ("0", str(w), "0", "-0", "0", "-0"), "-xi": (f"-{w}", "-0", "0", "-0", "0", "-0"),
            "+eta": ("0", "-0", "0", str(w), "0", "-0"), "-eta": ("0", "-0", f"-{w}", "-0", "0", "-0")}
    return [(0, "sponge_freund_ambient", s) + spec[s] for s in sides]


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_z_invariant_problem_reproduces_2d(prec, tol):
    c = cases.case_curv()
    nk, dz = 9, 0.37
    nj, ni = c["x"].shape
    x3 = np.repeat(c["x"][None], nk, 0)
    y3 = np.repeat(c["y"][None], nk, 0)
    z3 = dz * np.arange(nk)[:, None, None] + 0.0 * x3
    mean3 = np.repeat(c["mean"][:, None], nk, 1)
    ic3 = np.repeat(c["ic"][:, None], nk, 1)
    bc3 = [(0,) + tuple(b[1:3]) + tuple(b[3:7]) + ("0", "-0") for b in c["bc"]]
    dt = 0.05
    s2 = Solver([dict(x=c["x"], y=c["y"], mean=c["mean"])], boundaries=c["bc"], precision=prec, mode="FIX_DT",
                value=dt)
    s2.set_solution(c["ic"])
    s2.solve(40)
    ref = s2.solution()
    s3 = Solver([dict(x=x3, y=y3, z=z3, mean=mean3)], boundaries=bc3, precision=prec, mode="FIX_DT", value=dt,
                num_dim=3)
    s3.set_solution(ic3)
    s3.solve(40)
    q = s3.solution()
    errs = [max(_rel(q[v, k], ref[v]) for v in (0, 1, 2, 4)) for k in range(nk)]
    wmax = float(np.max(np.abs(q[3])))
    print(f"z-invariant prec {prec}: max plane error {max(errs):.2e}, |w'| {wmax:.1e}", flush=True)
    assert max(errs) <= tol
    assert wmax <= tol * np.max(np.abs(ref[4]))
    s2.close()
    s3.close()


def _box(ni, nj, nk, h=0.25):
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    return h * i, h * j - 0.1 * h * nj, h * k + 0.05 * h * nk


def test_axis_exchange_symmetry():
    """A shear layer u(y, z) along x with an off-centre pulse, against the
    same problem with x and z exchanged (w along z', then u along x)."""
    ni, nj, nk = 26, 22, 18
    x, y, z = _box(ni, nj, nk)
    yc, zc = y.mean(), z.mean()
    r = np.sqrt((y - yc) ** 2 + (z - zc) ** 2)
    mean = np.zeros((5, nk, nj, ni))
    mean[1] = 0.3 * 0.5 * (1 - np.tanh((r - 1.2) / 0.4))
    mean[4] = 1.0 / GAMMA
    ic = np.zeros((5, nk, nj, ni))
    ic[4] = 1e-3 * np.exp(-((x - 2.5) ** 2 + (y - yc - 0.4) ** 2 + (z - zc + 0.3) ** 2) / 0.5)
    ic[2] = 2e-4 * np.exp(-((x - 3.0) ** 2 + (y - yc) ** 2 + (z - zc) ** 2) / 0.6)
    out = []
    for swap in (False, True):
        if not swap:
            X, Y, Z, M, I = x, y, z, mean, ic
        else:  # exchange the i and k axes and the x and z components
            T = lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -3))
            X, Y, Z = T(z), T(y), T(x)
            M = np.stack([T(mean[0]), T(mean[3]), T(mean[2]), T(mean[1]), T(mean[4])])
            I = np.stack([T(ic[0]), T(ic[3]), T(ic[2]), T(ic[1]), T(ic[4])])
        s = Solver([dict(x=X, y=Y, z=Z, mean=M)], boundaries=(), precision=FP64, mode="FIX_DT", value=0.04, num_dim=3)
        s.set_solution(I)
        s.solve(30)
        q = s.solution()
        if swap:
            T = lambda a: np.swapaxes(a, -1, -3)
            q = np.stack([T(q[0]), T(q[3]), T(q[2]), T(q[1]), T(q[4])])
        out.append(q)
        s.close()
    err = max(_rel(out[1][v], out[0][v]) for v in (1, 2, 3, 4))
    print(f"axis exchange: max rel difference {err:.2e}", flush=True)
    assert np.max(np.abs(out[0][4])) > 1e-5
    assert err <= 1e-12


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("dims", [(130, 80, 12), (70, 22, 140), (70, 22, 77)])
def test_3d_plane_filter_forms_bit_identical(prec, dims):
    """Every form of the 3-D filter gives the same bits: the two-pass form
    with its plane pass as tiles or as stacks of 4 waves of 4, 8 or 16 rows,
    and the single pass (tiles of 8 or 16 rows marching through chunks of
    64 planes, two planes per loop trip).  The 130 x 80 planes have
    interior tiles and stacks besides edge ones; the 140-plane box spans
    three chunks of the single pass; the 77-plane box ends with a chunk of
    13 planes, an odd count of planes with their halo."""
    ni, nj, nk = dims
    x, y, z = _box(ni, nj, nk)
    yc, zc = y.mean(), z.mean()
    mean = np.zeros((5, nk, nj, ni))
    mean[1] = 0.3 * 0.5 * (1 - np.tanh((np.abs(y - yc) - 1.2) / 0.4))
    mean[4] = 1.0 / GAMMA
    ic = np.zeros((5, nk, nj, ni))
    ic[4] = 1e-3 * np.exp(-((x - 9.0) ** 2 + (y - yc - 0.4) ** 2 + (z - zc + 0.3) ** 2) / 2.0)
    ic[1] = 2e-4 * np.exp(-((x - 20.0) ** 2 + (y - yc) ** 2 + (z - zc) ** 2) / 3.0)
    ref = None
    for form in ((0, 0), (4, 4), (4, 8), (4, 16), (1, 8), (1, 16)):
        s = Solver([dict(x=x, y=y, z=z, mean=mean)], boundaries=(), precision=prec, mode="FIX_DT", value=0.04,
                   num_dim=3)
        s.set_filter_march(*form)
        s.set_solution(ic)
        s.solve(12)
        q = s.solution()
        s.close()
        if ref is None:
            ref = q
            assert np.max(np.abs(q[4])) > 1e-5
        assert np.array_equal(q, ref), (prec, form)
    print(f"3-D filter forms {dims} prec {prec}: bit-identical", flush=True)
    s = Solver([dict(x=x, y=y, z=z, mean=mean)], boundaries=(), precision=prec, mode="FIX_DT", value=0.04,
               num_dim=3)
    for bad in ((1, 4), (8, 8)):
        with pytest.raises(LeeError):
            s.set_filter_march(*bad)
    s.close()


def test_derivative_exact_on_quadratics():
    ni, nj, nk = 20, 18, 16
    x, y, z = _box(ni, nj, nk)
    mean = np.zeros((5, nk, nj, ni))
    mean[4] = 1.0 / GAMMA
    s = Solver([dict(x=x, y=y, z=z, mean=mean)], boundaries=(), num_dim=3)
    f = z * z + 0.5 * x * z - 2.0 * y * y + x
    for d, exact in enumerate((0.5 * z + 1.0, -4.0 * y, 2.0 * z + 0.5 * x)):
        g = s.take_derivative_xyz(f, d)
        err = float(np.max(np.abs(g - exact)))
        print(f"d/dx_{d}: max error {err:.2e}", flush=True)
        assert err <= 1e-11
    met, jac = s.metrics()
    assert np.allclose(jac, 1.0 / 0.25 ** 3) and np.allclose(met[2, 2] * jac, 1.0 / 0.25)
    assert s.compute_cfl() > 0.0
    s.close()


def test_3d_rejects_other_boundaries():
    ni = nj = nk = 12
    x, y, z = _box(ni, nj, nk)
    mean = np.zeros((5, nk, nj, ni))
    mean[4] = 1.0 / GAMMA
    with pytest.raises(LeeError):
        Solver([dict(x=x, y=y, z=z, mean=mean)], num_dim=3,
               boundaries=[(0, "wall_slip_kinematic", "+eta", "0", "-0", "0", "0", "0", "-0")])


def _shear_box(ni=22, nj=20, nk=18, h=0.25):
    x, y, z = _box(ni, nj, nk, h)
    yc, zc = y.mean(), z.mean()
    r = np.sqrt((y - yc) ** 2 + (z - zc) ** 2)
    mean = np.zeros((5, nk, nj, ni))
    mean[1] = 0.45 * 0.5 * (1 - np.tanh((r - 1.0) / 0.3))
    mean[4] = 1.0 / GAMMA
    bc = [(0, "sponge_freund_ambient", "+xi", "0", "4", "0", "-0", "0", "-0"),
          (0, "sponge_freund_ambient", "-xi", "-4", "-0", "0", "-0", "0", "-0")]
    return x, y, z, mean, bc, r


@pytest.mark.parametrize("dims", [(22, 20, 18), (14, 12, 70)])
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_3d_transposes_dot_products(prec, tol, dims):
    """<L a, k> = <a, L^T k> for the right-hand side and the filter; the
    70-plane box spans three 32-plane chunks of the z passes."""
    x, y, z, mean, bc, _ = _shear_box(*dims)
    s = Solver([dict(x=x, y=y, z=z, mean=mean)], boundaries=bc, precision=prec, num_dim=3)
    rng = np.random.default_rng(5)
    a, k = rng.standard_normal((2, 5) + x.shape)
    t = 0.2
    La = s.compute_rhs(a, t) - s.compute_rhs(np.zeros_like(a), t)
    lhs, rhs = float(np.sum(La * k)), float(np.sum(a * s.compute_rhs_adjoint(k)))
    e1 = abs(lhs - rhs) / abs(lhs)
    lhs2, rhs2 = float(np.sum(s.apply_filter(a) * k)), float(np.sum(a * s.apply_filter_adjoint(k)))
    e2 = abs(lhs2 - rhs2) / abs(lhs2)
    print(f"3-D dot products prec {prec}: rhs {e1:.2e}, filter {e2:.2e}", flush=True)
    assert e1 <= tol and e2 <= tol
    s.close()


@pytest.mark.parametrize("dims", [(22, 20, 18), (16, 14, 70)])
def test_3d_gradient_matches_central_differences(dims):
    """Observer energy on a cylindrical surface around the shear layer,
    gradient with respect to a field source, checkpoints every 8 steps; the
    70-plane box spans three 32-plane chunks of the stage kernel."""
    x, y, z, mean, bc, r = _shear_box(*dims)
    h = 0.25
    s = Solver([dict(x=x, y=y, z=z, mean=mean)], boundaries=bc, precision=FP64, num_dim=3,
               simulation="SOURCE_FIELD", source_params=(1.0, 0.5 * math.pi))
    nk, nj, ni = x.shape
    obs = [(0, i, j, k, h * h) for k in range(nk) for j in range(nj) for i in range(ni)
           if abs(r[k, j, i] - 1.8) < 0.5 * h and 4 < i < ni - 5]
    s.set_observer(obs)
    q = np.exp(-((x - 2.0) ** 2 + (y - y.mean()) ** 2 + (z - z.mean()) ** 2) / 0.5)
    y0 = np.zeros((5,) + x.shape)

    def J(qq):
        s.set_source_field(qq)
        s.set_solution(y0)
        s.set_time(0, 0.0, 1e-6)
        return s.objective_forward(30, 18, 8)

    J0 = J(q)
    s.objective_backward(1.0)
    g = s.source_gradient()
    rng = np.random.default_rng(8)
    for trial in range(2):
        e = rng.standard_normal(x.shape)
        eps = 1e-4
        fd = (J(q + eps * e) - J(q - eps * e)) / (2 * eps)
        ad = float(np.sum(g * e))
        err = abs(fd - ad) / abs(fd)
        print(f"3-D gradient direction {trial}: J {J0:.6e} fd {fd:.12e} adjoint {ad:.12e} rel {err:.2e}", flush=True)
        assert err <= 1e-6
    s.close()
