"""Overset test problems: grids, hole cutting, donor search and Lagrange
weights for explicit ORDER-4 (5-point) interpolation.

Tables are returned both as the flattened arrays of the C ABI
(lee_overset) and as Overture-file records for the reference solver.
"""
import math

import numpy as np

import cases

GAMMA = cases.GAMMA


def lagrange_weights(t, m=5):
    """1-D Lagrange weights of nodes 0..m-1 at position t."""
    w = []
    for a in range(m):
        v = 1.0
        for b in range(m):
            if b != a:
                v *= (t - b) / (a - b)
        w.append(v)
    return w


def _stencil(fi, fj, ni, nj, m=5):
    """Lower-left corner and tensor weights (w1 fastest) of an m x m stencil
    around continuous index (fi, fj), kept inside [0, n)."""
    i0 = min(max(int(math.floor(fi)) - (m // 2), 0), ni - m)
    j0 = min(max(int(math.floor(fj)) - (m // 2), 0), nj - m)
    wi = lagrange_weights(fi - i0, m)
    wj = lagrange_weights(fj - j0, m)
    coef = [wj[b] * wi[a] for b in range(m) for a in range(m)]
    return (i0, j0), coef


def case_overset_small():
    """Background 97x81 Cartesian box [-12, 12] x [-10, 10] plus a 41x41
    square patch (spacing 0.2) rotated by 25 degrees around (1, -0.5).
    Patch: its two outer rings are receivers from the background.
    Background: points deep inside the patch are holes, a fringe around
    the hole receives from the patch; on every fifth row the fringe is
    left out on the low-x side, so there a fluid point meets the hole
    directly.  Uniform mean flow 0.2, Gaussian pulse, 8-point sponges on
    the background, RK4 CFL 0.5, 40 steps."""
    n0i, n0j, h0 = 97, 81, 0.25
    x0 = np.empty((n0j, n0i))
    y0 = np.empty((n0j, n0i))
    for j in range(n0j):
        for i in range(n0i):
            x0[j, i] = -12.0 + h0 * i
            y0[j, i] = -10.0 + h0 * j
    n1, h1, th, xc, yc = 41, 0.2, math.radians(25.0), 1.0, -0.5
    ct, st = math.cos(th), math.sin(th)
    x1 = np.empty((n1, n1))
    y1 = np.empty((n1, n1))
    for j in range(n1):
        for i in range(n1):
            xi, eta = -4.0 + h1 * i, -4.0 + h1 * j
            x1[j, i] = xc + (xi * ct - eta * st)
            y1[j, i] = yc + (xi * st + eta * ct)

    def to_patch(x, y):
        dx, dy = x - xc, y - yc
        return dx * ct + dy * st, -dx * st + dy * ct

    recv, holes = [], []
    ib0 = np.zeros((n0j, n0i), np.int32)
    for j in range(n0j):
        for i in range(n0i):
            a, b = to_patch(x0[j, i], y0[j, i])
            if abs(a) < 2.4 and abs(b) < 2.4:
                ib0[j, i] = 1
    for j in range(n0j):
        for i in range(n0i):
            if ib0[j, i] == 1:
                continue
            a, b = to_patch(x0[j, i], y0[j, i])
            if abs(a) < 3.0 and abs(b) < 3.0:
                if j % 5 == 0 and x0[j, i] < xc:
                    continue
                ib0[j, i] = -1
                ll, coef = _stencil((a + 4.0) / h1, (b + 4.0) / h1, n1, n1)
                recv.append((0, (i, j, 0), 1, (ll[0], ll[1], 0), coef))
    for j in range(n0j):
        for i in range(n0i):
            if ib0[j, i] == 1:
                holes.append((0, (i, j, 0)))
    for j in range(n1):
        for i in range(n1):
            if i < 2 or i > n1 - 3 or j < 2 or j > n1 - 3:
                ll, coef = _stencil((x1[j, i] + 12.0) / h0, (y1[j, i] + 10.0) / h0, n0i, n0j)
                recv.append((1, (i, j, 0), 0, (ll[0], ll[1], 0), coef))
    pulse = cases._pulse(1e-3, -5.0, 2.0, 1.0)

    def block(x, y):
        ic = np.zeros((5,) + x.shape)
        ic[4] = cases._field(x, y, pulse)
        return dict(x=x, y=y, mean=cases._uniform_mean(x, u=0.2), ic=ic)

    blocks = [block(x0, y0), block(x1, y1)]
    bc = cases._sponges(8)
    return dict(name="overset_small", blocks=blocks, bc=bc, recv=recv, holes=holes,
                widths=[(5, 5, 1), (5, 5, 1)], order=4, mode="FIX_CFL", cfl=0.5, dt=None, nsteps=40,
                simulation="CASE_TANNA_TPN49", **cases.BASE_SCHEME)


def abi_tables(case):
    """Flattened receiver/donor/hole arrays (lee_overset / orc_set_overset)."""
    dims = [b["x"].shape for b in case["blocks"]]  # (nj, ni)
    recv = sorted(case["recv"], key=lambda r: r[0])
    rb = [r[0] for r in recv]
    rl = [r[1][1] * dims[r[0]][1] + r[1][0] for r in recv]
    db = [r[2] for r in recv]
    dl = [r[3][1] * dims[r[2]][1] + r[3][0] for r in recv]
    coef = np.array([r[4] for r in recv], dtype=np.float64)
    hb = [h[0] for h in case["holes"]]
    hl = [h[1][1] * dims[h[0]][1] + h[1][0] for h in case["holes"]]
    return dict(recv_block=rb, recv_l=rl, donor_block=db, donor_l=dl, widths=list(case["widths"]),
                coef=coef, hole_block=hb, hole_l=hl)


def oracle_for(case, Oracle):
    scheme = {k: case[k] for k in ("fd_order", "filter_order", "filter_blend", "rk", "sponge_order",
                                   "sponge_constant")}
    value = case["cfl"] if case["mode"] == "FIX_CFL" else case["dt"]
    return Oracle(case["blocks"], scheme, case["bc"], None, case["mode"], value, overset=abi_tables(case))


def case_cylinder_small():
    """A small instance of the cylinder problem (camille_amd.overset_gen):
    background 193x97 on [-6, 6] x [-3, 3], O-grid 69x21 from r = 0.5 to
    1.5 with its seam, holes for r < 1 and a 4-cell fringe, ORDER-4
    interpolation; potential-flow mean at U = 0.3, wall-slip on the
    cylinder, 10-point sponges on the background, a Gaussian pulse at
    (-1.5, 0.2); RK4 CFL 0.5, 60 steps."""
    from camille_amd import overset_gen as og
    geo = og.cylinder(nbg=(193, 97), extent=(-6.0, 6.0, -3.0, 3.0), nog=(69, 21), r_in=0.5, r_out=1.5,
                      r_hole=1.0, fringe=4, order=4)
    pulse = cases._pulse(1e-3, -1.5, 0.2, 0.3)
    blocks = []
    for x, y in geo["blocks"]:
        ic = np.zeros((5,) + x.shape)
        ic[4] = cases._field(x, y, pulse)
        blocks.append(dict(x=x, y=y, mean=og.potential_flow(x, y, 0.3, 0.5), ic=ic))
    noi = geo["blocks"][1][0].shape[1]
    bc = cases._sponges(10) + [(1, "wall_slip_kinematic", "+eta", "0", "-0", "0", "0", "0", "0")]
    return dict(name="cylinder_small", blocks=blocks, bc=bc, geo=geo, widths=geo["widths"], order=4,
                mode="FIX_CFL", cfl=0.5, dt=None, nsteps=60, simulation="CASE_TANNA_TPN49", noi=noi,
                **cases.BASE_SCHEME)


def cylinder_abi_tables(case):
    from camille_amd import overset_gen as og
    return og.abi_tables(case["geo"])


def oracle_for_cylinder(case, Oracle):
    scheme = {k: case[k] for k in ("fd_order", "filter_order", "filter_blend", "rk", "sponge_order",
                                   "sponge_constant")}
    return Oracle(case["blocks"], scheme, case["bc"], None, case["mode"], case["cfl"],
                  overset=cylinder_abi_tables(case))
