"""Deterministic definitions of the small LEE cases used as golden fixtures.

Every array is built point by point with Python's `math` module (glibc libm)
so that the same inputs come out bit-for-bit on any x86-64 host; the cases
are small enough for that to take well under a second each.

A case is a dict with:
  name, ni, nj, x, y (float64 (nj, ni)), mean (nv, nj, ni) [s, u, v, w, p,
  then the passive scalars of a LEE_SCALAR case], ic (nv, nj, ni), bc
  (BC-file lines), deck overrides, nsteps, and the solver settings the deck
  carries; nv = 5 + num_scalar (num_scalar absent: 0).
"""
import math

import numpy as np

GAMMA = 1.4
LN2 = math.log(2.0)


def _grid(ni, nj, fx, fy):
    x = np.empty((nj, ni))
    y = np.empty((nj, ni))
    for j in range(nj):
        for i in range(ni):
            x[j, i] = fx(i, j)
            y[j, i] = fy(i, j)
    return x, y


def _field(x, y, f):
    nj, ni = x.shape
    out = np.empty((nj, ni))
    for j in range(nj):
        for i in range(ni):
            out[j, i] = f(x[j, i], y[j, i])
    return out


def _pulse(amp, x0, y0, halfwidth):
    k = LN2 / (halfwidth * halfwidth)
    return lambda x, y: amp * math.exp(-k * ((x - x0) ** 2 + (y - y0) ** 2))


def _uniform_mean(x, u=0.0, v=0.0):
    nj, ni = x.shape
    m = np.zeros((5, nj, ni))
    m[1] = u
    m[2] = v
    m[4] = 1.0 / GAMMA
    return m


def _sponges(width, sides=("+xi", "-xi", "+eta", "-eta")):
    w = width - 1
    spec = {
        "+xi": ("0", str(w), "0", "-0", "0", "0"),
        "-xi": (f"-{w}", "-0", "0", "-0", "0", "0"),
        "+eta": ("0", "-0", "0", str(w), "0", "0"),
        "-eta": ("0", "-0", f"-{w}", "-0", "0", "0"),
    }
    return [(0, "sponge_freund_ambient", s) + spec[s] for s in sides]


BASE_SCHEME = dict(fd_order=4, filter_order=8, filter_blend=1.0, rk=4,
                   sponge_order=2, sponge_constant=5.0)


def case_c1():
    """201x201 on [-50,50]^2, quiescent, Gaussian pulse (1e-3, half-width
    3 dx) at the origin, four 20-point sponges, RK4 CFL 0.5, 200 steps."""
    ni = nj = 201
    x, y = _grid(ni, nj, lambda i, j: -50.0 + 0.5 * i, lambda i, j: -50.0 + 0.5 * j)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, 0.0, 0.0, 1.5))
    return dict(name="c1", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x), ic=ic,
                bc=_sponges(20), mode="FIX_CFL", cfl=0.5, dt=None, nsteps=200,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def _wavy(ni, nj, lx, ly, amp_frac):
    """x = xi + a sin(2 pi eta / L), y = eta + a sin(2 pi xi / L) on a box
    [-lx/2, lx/2] x [-ly/2, ly/2] (the C2 mapping, scaled)."""
    L = max(lx, ly)
    a = amp_frac * L
    dxi = lx / (ni - 1)
    deta = ly / (nj - 1)

    def xi(i):
        return -0.5 * lx + dxi * i

    def eta(j):
        return -0.5 * ly + deta * j

    return _grid(ni, nj,
                 lambda i, j: xi(i) + a * math.sin(2.0 * math.pi * eta(j) / L),
                 lambda i, j: eta(j) + a * math.sin(2.0 * math.pi * xi(i) / L))


def case_curv():
    """97x81 wavy curvilinear grid, uniform mean flow M = 0.5 in x, off-centre
    pulse, 12-point sponges, RK4 CFL 0.5, 60 steps."""
    ni, nj = 97, 81
    x, y = _wavy(ni, nj, 48.0, 40.0, 0.05)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, -4.0, 2.0, 1.5))
    return dict(name="curv", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.5), ic=ic,
                bc=_sponges(12), mode="FIX_CFL", cfl=0.5, dt=None, nsteps=60,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_filt():
    """One step with DT = 1e-300 on a curvilinear grid: the RK update adds
    nothing representable, so the output is the filter applied to the IC."""
    ni, nj = 61, 53
    x, y = _wavy(ni, nj, 30.0, 26.0, 0.05)
    ic = np.zeros((5, nj, ni))
    ic[0] = _field(x, y, _pulse(2e-4, 3.0, -2.0, 2.0))
    ic[1] = _field(x, y, _pulse(5e-4, -2.0, 1.0, 1.0))
    ic[2] = _field(x, y, _pulse(-4e-4, 1.0, 3.0, 1.2))
    ic[4] = _field(x, y, lambda a, b: 1e-3 * math.exp(-LN2 * (a * a + b * b) / 0.8 ** 2)
                   + 2e-4 * math.sin(0.9 * a) * math.cos(1.3 * b))
    return dict(name="filt", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.2, v=-0.1), ic=ic,
                bc=[], mode="FIX_DT", cfl=None, dt=1e-300, nsteps=1,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_shear():
    """81x65 Cartesian box with a non-uniform mean (shear layer in u, entropy
    and pressure bumps) so that every mean-gradient term and the aux density
    enter the RHS; pulse plus entropy spot IC; RK4 CFL 0.5, 40 steps."""
    ni, nj = 81, 65
    x, y = _grid(ni, nj, lambda i, j: -10.0 + 0.25 * i, lambda i, j: -8.0 + 0.25 * j)
    mean = np.zeros((5, nj, ni))
    mean[0] = _field(x, y, lambda a, b: 0.1 * math.exp(-(b * b) / 4.0))
    mean[1] = _field(x, y, lambda a, b: 0.25 * (1.0 + math.tanh(b / 1.5)))
    mean[2] = _field(x, y, lambda a, b: 0.02 * math.sin(0.3 * a))
    mean[4] = _field(x, y, lambda a, b: (1.0 / GAMMA) * (1.0 + 0.05 * math.exp(-(a * a + b * b) / 16.0)))
    ic = np.zeros((5, nj, ni))
    ic[0] = _field(x, y, _pulse(1e-4, 2.0, -1.0, 1.0))
    ic[4] = _field(x, y, _pulse(1e-3, -2.0, 0.5, 1.0))
    return dict(name="shear", ni=ni, nj=nj, x=x, y=y, mean=mean, ic=ic,
                bc=_sponges(10), mode="FIX_CFL", cfl=0.5, dt=None, nsteps=40,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_cylsrc():
    """The reference's two-cylinder scattering source (a time-dependent,
    ramped pressure source at the origin) on an 81x81 box, FIX_DT 0.01,
    40 steps: pins the time at which sources are evaluated."""
    ni = nj = 81
    x, y = _grid(ni, nj, lambda i, j: -2.0 + 0.05 * i, lambda i, j: -2.0 + 0.05 * j)
    ic = np.zeros((5, nj, ni))
    return dict(name="cylsrc", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x), ic=ic,
                bc=_sponges(10), mode="FIX_DT", cfl=None, dt=0.01, nsteps=40,
                simulation="CASE_SCATTERING_TWOCYLINDER", **BASE_SCHEME)


def case_wall():
    """Wavy channel with wall_slip_kinematic on both eta faces, uniform mean
    flow 0.3, pulse IC, sponges on both xi faces; RK4 CFL 0.5, 60 steps."""
    ni, nj = 121, 41
    lx = 30.0

    def xf(i, j):
        return -15.0 + 0.25 * i

    def yf(i, j):
        return -5.0 + 0.25 * j + 0.6 * math.sin(2.0 * math.pi * xf(i, j) / lx)

    x, y = _grid(ni, nj, xf, yf)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, -3.0, 0.0, 1.0))
    bc = [(0, "wall_slip_kinematic", "+eta", "0", "-0", "0", "0", "0", "0"),
          (0, "wall_slip_kinematic", "-eta", "0", "-0", "-0", "-0", "0", "0")]
    bc += _sponges(12, sides=("+xi", "-xi"))
    return dict(name="wall", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.3), ic=ic,
                bc=bc, mode="FIX_CFL", cfl=0.5, dt=None, nsteps=60,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_variants():
    """Scheme variants on a curvilinear grid with a mean flow: RK2, 4th-order
    filter blended at 0.7, FIX_CFL 0.4, 30 steps.  (The reference's 2nd-order
    FD is not usable: its closure table is one entry long but two entries are
    written, spatial_discretization.cpp:52,70-71, and runs diverge.)"""
    ni, nj = 73, 65
    x, y = _wavy(ni, nj, 36.0, 32.0, 0.04)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, 2.0, -1.0, 1.5))
    ic[1] = _field(x, y, _pulse(3e-4, -1.0, 2.0, 1.2))
    scheme = dict(BASE_SCHEME, filter_order=4, filter_blend=0.7, rk=2)
    return dict(name="variants", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.3, v=0.1), ic=ic,
                bc=_sponges(8), mode="FIX_CFL", cfl=0.4, dt=None, nsteps=30,
                simulation="CASE_TANNA_TPN49", **scheme)


def case_jet():
    """The reference's 2-D jet harmonic source (a cos(omega t) pressure
    source) on a 97x73 box with 10-point sponges, FIX_DT 0.05, 40 steps."""
    ni, nj = 97, 73
    x, y = _grid(ni, nj, lambda i, j: -6.0 + 0.125 * i, lambda i, j: -4.5 + 0.125 * j)
    ic = np.zeros((5, nj, ni))
    return dict(name="jet", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.2), ic=ic,
                bc=_sponges(10), mode="FIX_DT", cfl=None, dt=0.05, nsteps=40,
                simulation="CASE_2DJET_WITH_A_HARMONIC_SOURCE", **BASE_SCHEME)


def case_opt11():
    """The optimized 11-point schemes (FDO11P derivative, SFO11P filter at
    strength 0.6) on a 97x81 wavy curvilinear grid with a non-uniform mean
    (shear layer, entropy and pressure bumps), a Neumann face and a
    slip wall, 12-point sponges on the other faces; RK4 CFL 0.5, 40 steps."""
    ni, nj = 97, 81
    x, y = _wavy(ni, nj, 48.0, 40.0, 0.04)
    mean = np.zeros((5, nj, ni))
    mean[0] = _field(x, y, lambda a, b: 0.1 * math.exp(-(b * b) / 16.0))
    mean[1] = _field(x, y, lambda a, b: 0.25 * (1.0 + math.tanh(b / 3.0)))
    mean[2] = _field(x, y, lambda a, b: 0.02 * math.sin(0.15 * a))
    mean[4] = _field(x, y, lambda a, b: (1.0 / GAMMA) * (1.0 + 0.05 * math.exp(-(a * a + b * b) / 64.0)))
    ic = np.zeros((5, nj, ni))
    ic[0] = _field(x, y, _pulse(1e-4, 3.0, -2.0, 2.0))
    ic[4] = _field(x, y, _pulse(1e-3, -4.0, 1.0, 2.0))
    bc = _sponges(12, sides=("+xi", "-xi")) + [
        (0, "neumann", "+eta", "0", "-0", "0", "0", "0", "0"),
        (0, "wall_slip_kinematic", "-eta", "0", "-0", "-0", "-0", "0", "0")]
    scheme = dict(BASE_SCHEME, fd_order=11, filter_order=11, filter_strength=0.6)
    return dict(name="opt11", ni=ni, nj=nj, x=x, y=y, mean=mean, ic=ic, bc=bc, mode="FIX_CFL", cfl=0.5,
                dt=None, nsteps=40, simulation="CASE_TANNA_TPN49", **scheme)


def case_axisym():
    """Axisymmetric LEE (axis x, r = y): 81x49 box x in [-10, 10], r in
    [0, 12] with the axis at j = 0 (centerline_axisym), a jet-like mean
    u(r) = 0.15 (1 - tanh(r - 3)), an off-axis pressure pulse at (0, 3),
    sponges on the x ends and the outer radius; RK4 CFL 0.5, 40 steps."""
    ni, nj = 81, 49
    x, y = _grid(ni, nj, lambda i, j: -10.0 + 0.25 * i, lambda i, j: 0.25 * j)
    mean = np.zeros((5, nj, ni))
    mean[1] = _field(x, y, lambda a, b: 0.15 * (1.0 - math.tanh(b - 3.0)))
    mean[4] = 1.0 / GAMMA
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, 0.0, 3.0, 1.0))
    bc = _sponges(10, sides=("+xi", "-xi", "-eta")) + [
        (0, "centerline_axisym", "+eta", "0", "-0", "0", "0", "0", "0")]
    return dict(name="axisym", ni=ni, nj=nj, x=x, y=y, mean=mean, ic=ic, bc=bc, mode="FIX_CFL", cfl=0.5,
                dt=None, nsteps=40, simulation="CASE_TANNA_TPN49", axisym=1, **BASE_SCHEME)


def case_dirichlet():
    """dirichlet_allzero on the +xi and -eta faces of a wavy 81x65 grid with
    a mean flow (u, v) = (0.3, 0.1), sponges on the other two faces, a pulse
    next to the +xi wall; RK4 CFL 0.5, 50 steps."""
    ni, nj = 81, 65
    x, y = _wavy(ni, nj, 20.0, 16.0, 0.03)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, -6.5, -2.0, 1.0))
    ic[0] = _field(x, y, _pulse(2e-4, -5.0, 4.0, 1.0))
    bc = [(0, "dirichlet_allzero", "+xi", "0", "0", "0", "-0", "0", "0"),
          (0, "dirichlet_allzero", "-eta", "0", "-0", "-0", "-0", "0", "0")]
    bc += _sponges(10, sides=("-xi", "+eta"))
    return dict(name="dirichlet", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.3, v=0.1), ic=ic,
                bc=bc, mode="FIX_CFL", cfl=0.5, dt=None, nsteps=50,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_centerline():
    """Quarter domain x, y in [0, 18] x [0, 15] (73x61) of a quiescent
    problem symmetric about both axes: centerline_cart_norm2x on the x = 0
    face (u' = 0, Neumann for the rest), centerline_cart_norm2y on y = 0
    (v' = 0), sponges on the far faces; pressure pulse at (2, 0) and an
    entropy spot at (0, 3); RK4 CFL 0.5, 50 steps."""
    ni, nj = 73, 61
    x, y = _grid(ni, nj, lambda i, j: 0.25 * i, lambda i, j: 0.25 * j)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, 2.0, 0.0, 1.0))
    ic[0] = _field(x, y, _pulse(2e-4, 0.0, 3.0, 1.0))
    bc = [(0, "centerline_cart_norm2x", "+xi", "0", "0", "0", "-0", "0", "0"),
          (0, "centerline_cart_norm2y", "+eta", "0", "-0", "0", "0", "0", "0")]
    bc += _sponges(10, sides=("-xi", "-eta"))
    return dict(name="centerline", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x), ic=ic,
                bc=bc, mode="FIX_CFL", cfl=0.5, dt=None, nsteps=50,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_entropy():
    """An entropy wave (HARMONIC_WAVE = WAVE_ENTROPY, WAVEFORM_HOMOGENEOUS,
    amplitude 0.01, wavelength 6) imposed by dirichlet_harmonicwave on the
    +xi face of an 81x57 box and relaxed to by a 10-point
    sponge_freund_harmonicwave zone behind it, convected by a mean flow
    u = 0.4; ambient sponges on the other three faces, a pressure pulse;
    RK4 CFL 0.5, 50 steps."""
    ni, nj = 81, 57
    x, y = _grid(ni, nj, lambda i, j: -10.0 + 0.25 * i, lambda i, j: -7.0 + 0.25 * j)
    ic = np.zeros((5, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, 2.0, 1.0, 1.0))
    bc = [(0, "dirichlet_harmonicwave", "+xi", "0", "0", "0", "-0", "0", "0"),
          (0, "sponge_freund_harmonicwave", "+xi", "0", "9", "0", "-0", "0", "0")]
    bc += _sponges(10, sides=("-xi", "+eta", "-eta"))
    harmonic = dict(wave="WAVE_ENTROPY", shape="WAVEFORM_HOMOGENEOUS", direction="X", amplitude=0.01,
                    wavelength=6.0)
    return dict(name="entropy", ni=ni, nj=nj, x=x, y=y, mean=_uniform_mean(x, u=0.4), ic=ic,
                bc=bc, mode="FIX_CFL", cfl=0.5, dt=None, nsteps=50, harmonic=harmonic,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_roundjet():
    """A round-jet entropy wave (WAVE_ENTROPY, WAVEFORM_GAUSSIAN_ROUNDJET,
    half-width 1.5, amplitude 0.02, wavelength 5, direction x) through the
    +xi face of a wavy 89x65 grid with a jet-like mean flow
    u = 0.3 (1 - tanh(|y| - 2)) / 2: dirichlet_harmonicwave on the face and a
    12-point sponge_freund_harmonicwave zone, ambient sponges elsewhere;
    RK4 CFL 0.5, 50 steps."""
    ni, nj = 89, 65
    x, y = _wavy(ni, nj, 22.0, 16.0, 0.02)
    mean = _uniform_mean(x)
    mean[1] = _field(x, y, lambda a, b: 0.15 * (1.0 - math.tanh(abs(b) - 2.0)))
    ic = np.zeros((5, nj, ni))
    bc = [(0, "sponge_freund_harmonicwave", "+xi", "0", "11", "0", "-0", "0", "0"),
          (0, "dirichlet_harmonicwave", "+xi", "0", "0", "0", "-0", "0", "0")]
    bc += _sponges(10, sides=("-xi", "+eta", "-eta"))
    harmonic = dict(wave="WAVE_ENTROPY", shape="WAVEFORM_GAUSSIAN_ROUNDJET", direction="X", amplitude=0.02,
                    wavelength=5.0, halfwidth=1.5)
    return dict(name="roundjet", ni=ni, nj=nj, x=x, y=y, mean=mean, ic=ic,
                bc=bc, mode="FIX_CFL", cfl=0.5, dt=None, nsteps=50, harmonic=harmonic,
                simulation="CASE_TANNA_TPN49", **BASE_SCHEME)


def case_scalar():
    """LEE_SCALAR with two passive scalars (NUMBER_SCALAR = 2) on a wavy 85x61
    grid: shear mean u = 0.2 + 0.1 tanh(y), mean scalar fields
    Zbar_0 = (1 + tanh(y / 1.5)) / 2 and Zbar_1 a Gaussian bump, so both the
    ubar . grad Z' and the u' . grad Zbar terms act; a scalar blob
    (amplitude 0.01, the size of CASE_LINEAR_NOZZLE's, state.cpp:630-639),
    a pressure pulse and an entropy spot; wall_slip_kinematic on +eta
    (scalars take the Neumann data), neumann on -eta, dirichlet_allzero on
    +xi, a 10-point sponge on -xi; RK4 CFL 0.5, 50 steps.  The reference
    admits LEE_SCALAR only with SIMULATION = CASE_LINEAR_NOZZLE
    (input.cpp:596-606), whose initial state the restart file replaces."""
    ni, nj = 85, 61
    x, y = _wavy(ni, nj, 21.0, 15.0, 0.02)
    mean = np.zeros((7, nj, ni))
    mean[1] = _field(x, y, lambda a, b: 0.2 + 0.1 * math.tanh(b))
    mean[4] = 1.0 / GAMMA
    mean[5] = _field(x, y, lambda a, b: 0.5 * (1.0 + math.tanh(b / 1.5)))
    mean[6] = _field(x, y, _pulse(1.0, -2.0, 0.0, 2.0))
    ic = np.zeros((7, nj, ni))
    ic[4] = _field(x, y, _pulse(1e-3, -3.0, 1.0, 1.0))
    ic[0] = _field(x, y, _pulse(2e-4, 1.0, 2.0, 1.0))
    ic[5] = _field(x, y, _pulse(1e-2, 2.0, -1.0, 1.0))
    bc = [(0, "wall_slip_kinematic", "+eta", "0", "-0", "0", "0", "0", "0"),
          (0, "neumann", "-eta", "0", "-0", "-0", "-0", "0", "0"),
          (0, "dirichlet_allzero", "+xi", "0", "0", "0", "-0", "0", "0")]
    bc += _sponges(10, sides=("-xi",))
    return dict(name="scalar", ni=ni, nj=nj, x=x, y=y, mean=mean, ic=ic, bc=bc, mode="FIX_CFL", cfl=0.5,
                dt=None, nsteps=50, num_scalar=2, simulation="CASE_LINEAR_NOZZLE", **BASE_SCHEME)


ALL = {f.__name__[5:]: f for f in (case_c1, case_curv, case_filt, case_shear, case_cylsrc, case_wall,
                                    case_variants, case_jet, case_opt11, case_axisym, case_dirichlet,
                                    case_centerline, case_entropy, case_roundjet, case_scalar)}

# cases whose filter (SFO11P) runs in the tile form only: the column-march
# forms are written for the standard filters' 4-point halo
TILE_FILTER_ONLY = {"opt11"}


def _harmonic_deck(h):
    """HARMONIC_WAVE deck value (input.cpp:219-231)."""
    if not h:
        return "NONE"
    out = f"{h['wave']}, AMPLITUDE = {h['amplitude']!r}, SHAPE = {h['shape']}, DIRECTION = {h['direction']}, " \
          f"WAVELENGTH = {h['wavelength']!r}"
    if "halfwidth" in h:
        out += f", HALFWIDTH = {h['halfwidth']!r}"
    return out


def deck_items(case, grid_file="grid.xyz", mean_file="mean.q", bc_file="bc.in"):
    """Deck lines for the reference solver (keys of src/core/input.cpp:99-284)."""
    items = [
        ("DIMENSION", "2"),
        ("AXISYMMETRY", "TRUE, AXIS_OF_SYMMETRY = X" if case.get("axisym") else "FALSE"),
        ("PHYSICAL_MODEL", f"LEE_SCALAR, NUMBER_SCALAR = {case['num_scalar']}" if case.get("num_scalar")
         else "LEE"),
        ("FLUID_MODEL", "IDEAL_GAS_CALORIC"),
        ("SIMULATION", case["simulation"]),
        ("GAMMA_SPECIFICHEAT", repr(GAMMA)),
        ("FIX_DT_OR_CFL", case["mode"]),
    ]
    if case["mode"] == "FIX_CFL":
        items.append(("CFL", repr(case["cfl"])))
    else:
        items.append(("DT", repr(case["dt"])))
    items += [
        ("TEMPORAL_SCHEME", "RUNGE_KUTTA4" if case["rk"] == 4 else "RUNGE_KUTTA2"),
        ("NUM_TOTAL_TIMESTEP", str(case["nsteps"])),
        ("REPORT_TIMESTEP", str(case["nsteps"])),
        ("SPATIAL_SCHEME", "FDO11P" if case["fd_order"] == 11
         else f"STANDARD_CENTRAL, ORDER_ACCURACY = {case['fd_order']}"),
    ]
    if case["filter_order"] == 11:
        items.append(("SPATIAL_FILTER", f"SFO11P, STRENGTH = {case['filter_strength']!r}, "
                                        f"BLEND = {case['filter_blend']!r}"))
    elif case["filter_order"]:
        items.append(("SPATIAL_FILTER",
                      f"STANDARD_CENTRAL, ORDER_ACCURACY = {case['filter_order']}, BLEND = {case['filter_blend']!r}"))
    else:
        items.append(("SPATIAL_FILTER", "NONE"))
    items += [
        ("BUFFER_ZONE", f"POLYNOMIAL_ORDER = {case['sponge_order']}, CONSTANT = {case['sponge_constant']!r}"),
        ("METRIC_SCHEME", "THOMAS_LOMBARD"),
        ("OVERSET", "NONE"),
        ("DOMAIN_DECOMPOSITION", "DECOMP_1D"),
        ("FILE_FORMAT", "PLOT3D"),
        ("GRID_FILE", grid_file),
        ("BASESTATE_FILE", mean_file),
        ("BC_FILE", bc_file),
        ("SOLUTION_WRITING_TIME", "1.0E+06"),
        ("HARMONIC_WAVE", _harmonic_deck(case.get("harmonic"))),
        ("INTERPOLATE_SOLUTION", "NONE"),
    ]
    return items
