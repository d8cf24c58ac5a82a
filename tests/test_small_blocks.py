"""Small and ragged blocks through the C ABI, against the oracle.

Blocks from 9 points a side (the narrowest on which the 8th-order filter
keeps one interior row: four boundary rows each side) up to thin strips of
9 x 131 and 131 x 9 points: narrower or shorter than every tile (64 wide,
4-64 rows), filter strip (56 outputs plus a 4-point halo each side) and
stack the kernels use, so every workgroup straddles the block's edges and
the one-sided FD closures and boundary filter rows of opposite faces meet
in the middle. The first test has no BC faces or sponges: every face
takes the scheme's default closures (spatial_discretization.cpp:1078-1336
for the FD rows, 348-492 and 2191-2246 for the filter), on a curvilinear
grid with a uniform mean flow at an angle, a pressure pulse and seeded
noise at every point.

On blocks this small the closures of opposite faces amplify small
perturbations: the oracle's own response to a 1e-7 relative perturbation
of the initial state reaches 3e-5 (relative) after 30 steps on the
9 x 14 block, 3e-6 after 8. FP32 bulk arithmetic perturbs the state at
that level on every step, so it is held to its bar after 8 steps.

Bars of the first test (relative to max |.| of each field): after 30
steps the plain path equal to the oracle bit for bit and the fused FP64
path to 1e-12 (rho', u', v', p'); after 8 steps FP32 bulk to 1e-5 (p').
The clock (step, time, dt) equals the oracle's. The other tests put the same
blocks in a batch, extrude them to 3-D, check the adjoint's transposes on
them and give them a boundary condition on every face.
"""
import numpy as np
import pytest

from camille_amd import FP32, FP64, PATH_FUSED, PATH_PLAIN, Solver

pytestmark = pytest.mark.gpu

GAMMA = 1.4
SHAPES = [(9, 9), (12, 9), (9, 14), (17, 10), (131, 9), (9, 131), (66, 11)]
SCHEME = dict(fd_order=4, filter_order=8, filter_blend=1.0, rk=4, sponge_order=2, sponge_constant=5.0)


def _inputs(ni, nj, seed, shift_cells=None):
    """Grid lines displaced sinusoidally by 3 % of the block's extent
    (shift_cells None: up to ~4 cells on the 131-point strips) or by
    `shift_cells` cells."""
    rng = np.random.default_rng(seed)
    dx = 0.5
    I, J = np.meshgrid(np.arange(ni), np.arange(nj))
    lx, ly = dx * (ni - 1), dx * (nj - 1)
    ax = 0.03 * lx if shift_cells is None else shift_cells * dx
    ay = 0.03 * ly if shift_cells is None else shift_cells * dx
    x = dx * I + ax * np.sin(2 * np.pi * J / (nj - 1))
    y = dx * J + ay * np.sin(2 * np.pi * I / (ni - 1))
    mean = np.zeros((5, nj, ni))
    mean[1] = 0.3
    mean[2] = -0.1
    mean[4] = 1.0 / GAMMA
    ic = 1e-6 * rng.standard_normal((5, nj, ni))
    ic[3] = 0.0  # 2-D: w' stays 0
    ic[4] += 1e-3 * np.exp(-((x - 0.4 * lx) ** 2 + (y - 0.6 * ly) ** 2) / (2 * dx) ** 2)
    return x, y, mean, ic


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("shape", SHAPES)
def test_small_block_matches_oracle(shape):
    from oracle.pyoracle import Oracle
    ni, nj = shape
    x, y, mean, ic = _inputs(ni, nj, seed=ni * 1000 + nj)
    o = Oracle([dict(x=x, y=y, mean=mean, ic=ic)], SCHEME, (), None, "FIX_CFL", 0.5)
    refs = {}
    o.run(8)
    refs[8] = (o.solution(0)[:, 0], o.time, o.dt)
    o.run(22)
    refs[30] = (o.solution(0)[:, 0], o.time, o.dt)
    assert np.all(np.isfinite(refs[30][0])) and np.max(np.abs(refs[30][0][4])) > 1e-4
    for label, prec, path, nsteps, tol in (("plain", FP64, PATH_PLAIN, 30, 0.0),
                                           ("fused64", FP64, PATH_FUSED, 30, 1e-12),
                                           ("fused32", FP32, PATH_FUSED, 8, 1e-5)):
        print(f"small block {shape} {label}", flush=True)
        ref, t_ref, dt_ref = refs[nsteps]
        s = Solver([dict(x=x, y=y, mean=mean)], boundaries=(), precision=prec, path=path, mode="FIX_CFL", value=0.5)
        s.set_solution(ic)
        s.solve(nsteps)
        q = s.solution()
        step, time, dt = s.clock()
        s.close()
        assert step == nsteps and time == t_ref and dt == dt_ref, (label, shape)
        if tol == 0.0:
            assert np.array_equal(q, ref), (label, shape)
            continue
        vars_checked = (4,) if prec == FP32 else (0, 1, 2, 4)
        err = max(_rel(q[v], ref[v]) for v in vars_checked)
        print(f"small block {shape} {label}: {err:.2e}", flush=True)
        assert err <= tol, (label, shape, err)


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_small_blocks_in_a_batch_equal_single_handles(prec):
    """The seven small blocks and one 70 x 50 block as one batched ensemble
    (one launch per RK stage over every solve's tiles, so most workgroups
    hold a tile that straddles a small block's edges): every solve equals
    its own single handle bit for bit after 12 steps, in the tuner's launch
    form, the persistent + graph form, direct launches of 16-row tiles and
    the stacked filter."""
    from camille_amd import GRID_RESIDENT, Batch
    shapes = SHAPES + [(70, 50)]
    solves = []
    for ni, nj in shapes:
        x, y, mean, ic = _inputs(ni, nj, seed=ni * 1000 + nj + 1)
        solves.append(dict(x=x, y=y, mean=mean, ic=ic, mode="FIX_CFL", value=0.5))
    nsteps = 12
    singles = []
    for sv in solves:
        s = Solver([dict(x=sv["x"], y=sv["y"], mean=sv["mean"])], boundaries=(), precision=prec, path=PATH_FUSED,
                   mode="FIX_CFL", value=0.5)
        s.set_solution(sv["ic"])
        s.solve(nsteps)
        singles.append((s.solution(), s.clock()))
        s.close()
    for form in ("tuned", "persistent", "direct16", "stacked"):
        b = Batch(solves, precision=prec)
        for k, sv in enumerate(solves):
            b.set_solution(sv["ic"], k)
        if form == "persistent":
            b.set_launch(True, GRID_RESIDENT, GRID_RESIDENT)
        elif form == "direct16":
            b.set_tile_shape(4, 16)
            b.set_launch(False, 0, 0)
        elif form == "stacked":
            b.set_filter_march(4, 8)
        b.solve(nsteps)
        for k in range(len(solves)):
            assert b.clock(k) == singles[k][1], (form, shapes[k])
            assert np.array_equal(b.solution(k), singles[k][0]), (form, shapes[k])
        print(f"small blocks batch prec {prec} {form} {b.launch()}: bit-identical to single handles", flush=True)
        b.close()


@pytest.mark.parametrize("shape", [(9, 9, 9), (12, 9, 70), (9, 14, 13), (131, 9, 10), (9, 131, 9)])
def test_small_3d_boxes_reproduce_2d(shape):
    """Small 3-D boxes (the oracle, like the reference, has no 3-D; the
    3-D path is pinned to the 2-D one): a small block extruded in z with a
    z-invariant state reproduces the 2-D fused solution in every plane,
    the z closures and boundary filter rows acting on constants. 9 planes
    is the fewest with an interior filter row in z; 70 planes span two
    chunks of the single-pass filter and three of the stage. FP64 to 1e-12
    after 30 steps, FP32 (against 2-D FP32) to 1e-5 after 8 steps, w' at
    most the bar times max |p'|.

    The 3-D metrics are formed by another formula than the 2-D ones (the
    adjugate of the coordinate Jacobian and J = 1/det), so the two agree
    to round-off only. The grid lines here are displaced by half a cell:
    with the displacement of the 2-D tests (x_eta up to 2.8 cells per cell
    on the 131 x 9 strip, so the y derivative cancels terms ~3x its size)
    the 131 x 9 x 10 box differed from 2-D by 2.5e-12 in FP64 after 30
    steps, the other boxes by at most 5e-14."""
    ni, nj, nk = shape
    x, y, mean, ic = _inputs(ni, nj, seed=ni * 1000 + nj + 2, shift_cells=0.5)
    x3 = np.repeat(x[None], nk, 0)
    y3 = np.repeat(y[None], nk, 0)
    z3 = 0.45 * np.arange(nk)[:, None, None] + 0.0 * x3
    mean3 = np.repeat(mean[:, None], nk, 1)
    ic3 = np.repeat(ic[:, None], nk, 1)
    dt = 0.1
    for prec, nsteps, tol in ((FP64, 30, 1e-12), (FP32, 8, 1e-5)):
        print(f"small 3-D box {shape} prec {prec}", flush=True)
        s2 = Solver([dict(x=x, y=y, mean=mean)], boundaries=(), precision=prec, mode="FIX_DT", value=dt)
        s2.set_solution(ic)
        s2.solve(nsteps)
        ref = s2.solution()
        s2.close()
        s3 = Solver([dict(x=x3, y=y3, z=z3, mean=mean3)], boundaries=(), precision=prec, mode="FIX_DT", value=dt,
                    num_dim=3)
        s3.set_solution(ic3)
        s3.solve(nsteps)
        q = s3.solution()
        s3.close()
        vars_checked = (0, 1, 2, 4) if prec == FP64 else (4,)
        err = max(_rel(q[v, k], ref[v]) for v in vars_checked for k in range(nk))
        wmax = float(np.max(np.abs(q[3])))
        print(f"small 3-D box {shape} prec {prec}: max plane error {err:.2e}, |w'| {wmax:.1e}", flush=True)
        assert np.all(np.isfinite(q))
        assert err <= tol
        assert wmax <= tol * np.max(np.abs(ref[4]))


def _rand_state(shape, rng):
    y = rng.standard_normal((5,) + shape)
    y[3] = 0.0
    return y


@pytest.mark.parametrize("shape", SHAPES + [(9, 9, 9), (12, 9, 70), (131, 9, 10)])
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_small_block_transposes_dot_products(shape, prec, tol):
    """The adjoint's transposed right-hand side and filter on small blocks,
    where the closure rows of opposite faces overlap in the transposes:
    <L a, k> = <a, L^T k> and <F a, k> = <a, F^T k> for seeded random a, k."""
    rng = np.random.default_rng(sum(shape))
    if len(shape) == 2:
        x, y, mean, _ = _inputs(*shape, seed=7)
        blk, kw = dict(x=x, y=y, mean=mean), {}
        a, k = _rand_state(x.shape, rng), _rand_state(x.shape, rng)
    else:
        ni, nj, nk = shape
        x, y, mean, _ = _inputs(ni, nj, seed=7, shift_cells=0.5)
        x3 = np.repeat(x[None], nk, 0)
        blk = dict(x=x3, y=np.repeat(y[None], nk, 0), z=0.45 * np.arange(nk)[:, None, None] + 0.0 * x3,
                   mean=np.repeat(mean[:, None], nk, 1))
        kw = dict(num_dim=3)
        a, k = rng.standard_normal((2, 5) + x3.shape)
    s = Solver([blk], boundaries=(), precision=prec, **kw)
    t = 0.3
    La = s.compute_rhs(a, t) - s.compute_rhs(np.zeros_like(a), t)
    v = [0, 1, 2, 4] if len(shape) == 2 else [0, 1, 2, 3, 4]
    lhs, rhs = float(np.sum(La[v] * k[v])), float(np.sum(a * s.compute_rhs_adjoint(k)))
    e1 = abs(lhs - rhs) / abs(lhs)
    lhs2, rhs2 = float(np.sum(s.apply_filter(a) * k)), float(np.sum(a * s.apply_filter_adjoint(k)))
    e2 = abs(lhs2 - rhs2) / abs(lhs2)
    s.close()
    print(f"small block {shape} prec {prec} dot products: rhs {e1:.2e}, filter {e2:.2e}", flush=True)
    assert e1 <= tol and e2 <= tol


BC_FACES = [(0, "wall_slip_kinematic", "+eta", "0", "-0", "0", "0", "0", "0"),
            (0, "wall_slip_kinematic", "-eta", "0", "-0", "-0", "-0", "0", "0"),
            (0, "neumann", "+xi", "0", "0", "0", "-0", "0", "0"),
            (0, "dirichlet_allzero", "-xi", "-0", "-0", "0", "-0", "0", "0")]


@pytest.mark.parametrize("shape", SHAPES)
def test_small_block_with_boundary_faces_matches_oracle(shape):
    """The small blocks with a BC on every face (wall-slip on both eta
    faces, Neumann on the first xi face, Dirichlet-zero on the last;
    boundary_condition.cpp:98-112, 348-494), so faces of different types
    meet at corners a few points apart and the face kernels cover whole
    sides of blocks narrower than their launches. With the faces the
    oracle stays bounded (its response to a 1e-7 relative perturbation of
    the initial state stays below 4e-7 over 30 steps), so every path is
    held after 30 steps: plain bit for bit, fused FP64 to 1e-12, FP32 to
    1e-5 (p')."""
    from oracle.pyoracle import Oracle
    ni, nj = shape
    x, y, mean, ic = _inputs(ni, nj, seed=ni * 1000 + nj + 3)
    nsteps = 30
    o = Oracle([dict(x=x, y=y, mean=mean, ic=ic)], SCHEME, BC_FACES, None, "FIX_CFL", 0.5)
    o.run(nsteps)
    ref = o.solution(0)[:, 0]
    assert np.all(np.isfinite(ref)) and np.max(np.abs(ref[4])) > 1e-5
    for label, prec, path, tol in (("plain", FP64, PATH_PLAIN, 0.0), ("fused64", FP64, PATH_FUSED, 1e-12),
                                   ("fused32", FP32, PATH_FUSED, 1e-5)):
        print(f"small block {shape} with faces {label}", flush=True)
        s = Solver([dict(x=x, y=y, mean=mean)], boundaries=BC_FACES, precision=prec, path=path, mode="FIX_CFL",
                   value=0.5)
        s.set_solution(ic)
        s.solve(nsteps)
        q = s.solution()
        step, time, dt = s.clock()
        s.close()
        assert step == nsteps and time == o.time and dt == o.dt, (label, shape)
        if tol == 0.0:
            assert np.array_equal(q, ref), (label, shape)
            continue
        vars_checked = (4,) if prec == FP32 else (0, 1, 2, 4)
        err = max(_rel(q[v], ref[v]) for v in vars_checked)
        print(f"small block {shape} with faces {label}: {err:.2e}", flush=True)
        assert err <= tol, (label, shape, err)


@pytest.mark.parametrize("shape", [(11, 11), (14, 11), (11, 17), (131, 11), (11, 131), (66, 13)])
def test_small_block_optimized_schemes_match_oracle(shape):
    """FDO11P and SFO11P (strength 0.6; spatial_discretization.cpp:228-344,
    348-492) on small blocks with a BC on every face, from 11 points a
    side (the narrowest with an interior row: five closure rows each
    side). Without the faces these blocks grow fast under the 11-point
    schemes (the oracle's max |p'| reaches 5 from 1e-3 in 30 steps), with
    them they stay bounded; the oracle's response to a 1e-7 relative
    perturbation of the initial state is up to 3.3e-6 after 30 steps and
    1.1e-6 after 8. Plain bit for bit and fused FP64 to 1e-12 after 30
    steps; FP32 to 1e-5 (p') after 8."""
    from oracle.pyoracle import Oracle
    ni, nj = shape
    x, y, mean, ic = _inputs(ni, nj, seed=ni * 1000 + nj + 4)
    scheme = dict(SCHEME, fd_order=11, filter_order=11, filter_strength=0.6)
    o = Oracle([dict(x=x, y=y, mean=mean, ic=ic)], scheme, BC_FACES, None, "FIX_CFL", 0.5)
    refs = {}
    o.run(8)
    refs[8] = (o.solution(0)[:, 0], o.time, o.dt)
    o.run(22)
    refs[30] = (o.solution(0)[:, 0], o.time, o.dt)
    assert np.all(np.isfinite(refs[30][0])) and np.max(np.abs(refs[30][0][4])) > 1e-5
    for label, prec, path, nsteps, tol in (("plain", FP64, PATH_PLAIN, 30, 0.0),
                                           ("fused64", FP64, PATH_FUSED, 30, 1e-12),
                                           ("fused32", FP32, PATH_FUSED, 8, 1e-5)):
        print(f"small block {shape} 11-point {label}", flush=True)
        ref, t_ref, dt_ref = refs[nsteps]
        s = Solver([dict(x=x, y=y, mean=mean)], boundaries=BC_FACES, precision=prec, path=path, mode="FIX_CFL",
                   value=0.5, fd_order=11, filter_order=11, filter_strength=0.6)
        s.set_solution(ic)
        s.solve(nsteps)
        q = s.solution()
        step, time, dt = s.clock()
        s.close()
        assert step == nsteps and time == t_ref and dt == dt_ref, (label, shape)
        if tol == 0.0:
            assert np.array_equal(q, ref), (label, shape)
            continue
        vars_checked = (4,) if prec == FP32 else (0, 1, 2, 4)
        err = max(_rel(q[v], ref[v]) for v in vars_checked)
        print(f"small block {shape} 11-point {label}: {err:.2e}", flush=True)
        assert err <= tol, (label, shape, err)


def _nonuniform_mean(x, y, mean):
    """A shear profile in u, a density bump and a pressure ripple: every
    mean field and mean gradient varies across the block."""
    m = mean.copy()
    yc, h = y.mean(), max(y.max() - y.min(), 1.0)
    m[1] = 0.3 * 0.5 * (1 + np.tanh((y - yc) / (0.2 * h)))
    m[0] = 1.0 + 0.1 * np.exp(-((x - x.mean()) ** 2 + (y - yc) ** 2) / (0.1 * h * h + 1))
    m[4] = 1.0 / GAMMA + 0.01 * np.sin(x)
    return m


@pytest.mark.parametrize("shape", SHAPES)
def test_small_block_nonuniform_mean_matches_oracle(shape):
    """The non-uniform-mean stage kernels (mean fields and their gradients,
    formed at create through the closures, spatial_discretization.cpp:
    2053-2130, 2250-2271, and the stale density of state.cpp:668-736) on
    the small blocks with a BC on every face. Bounded: the oracle's
    response to a 1e-7 relative perturbation of the initial state stays
    below 2.2e-7 over 30 steps. After 30 steps plain bit for bit, fused
    FP64 to 1e-12, FP32 to 1e-5 (p')."""
    from oracle.pyoracle import Oracle
    ni, nj = shape
    x, y, mean, ic = _inputs(ni, nj, seed=ni * 1000 + nj + 5)
    mean = _nonuniform_mean(x, y, mean)
    nsteps = 30
    o = Oracle([dict(x=x, y=y, mean=mean, ic=ic)], SCHEME, BC_FACES, None, "FIX_CFL", 0.5)
    o.run(nsteps)
    ref = o.solution(0)[:, 0]
    assert np.all(np.isfinite(ref)) and np.max(np.abs(ref[4])) > 1e-5
    for label, prec, path, tol in (("plain", FP64, PATH_PLAIN, 0.0), ("fused64", FP64, PATH_FUSED, 1e-12),
                                   ("fused32", FP32, PATH_FUSED, 1e-5)):
        print(f"small block {shape} non-uniform mean {label}", flush=True)
        s = Solver([dict(x=x, y=y, mean=mean)], boundaries=BC_FACES, precision=prec, path=path, mode="FIX_CFL",
                   value=0.5)
        s.set_solution(ic)
        s.solve(nsteps)
        q = s.solution()
        step, time, dt = s.clock()
        s.close()
        assert step == nsteps and time == o.time and dt == o.dt, (label, shape)
        if tol == 0.0:
            assert np.array_equal(q, ref), (label, shape)
            continue
        vars_checked = (4,) if prec == FP32 else (0, 1, 2, 4)
        err = max(_rel(q[v], ref[v]) for v in vars_checked)
        print(f"small block {shape} non-uniform mean {label}: {err:.2e}", flush=True)
        assert err <= tol, (label, shape, err)
