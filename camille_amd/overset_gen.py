"""Overset connectivity for a cylinder in a Cartesian box: hole cutting,
donor search and Lagrange interpolation weights, in the explicit Overture
layout that src/io/overture.cpp:473-680 reads (receiver block and index,
donor block and lower-left corner of a width^D stencil, the stencil's
coefficients w1 fastest, hole list).

Two blocks:
  0  background Cartesian grid, uniform spacing;
  1  body-fitted O-grid around a cylinder of radius r_in at the origin,
     x = r_j cos(theta_i), y = -r_j sin(theta_i) with
     theta_i = theta_s + 2 pi (i - 2) / (ni - 5), r_j = r_in + j dr: clockwise in i
     and outwards in j, so x_xi y_eta - x_eta y_xi > 0 like the background
     (the reference's CFL evaluation multiplies by J and keeps the largest
     value, so a left-handed block would drop out of it).  Columns
     i = 2 .. ni - 3 cover one period and the two columns on each side
     repeat columns ni - 3, ni - 4 and 3, 4 exactly (seam).  The seam sits
     at theta_s = 45 degrees by default, away from the stagnation points of
     the mean flow: with the seam at the rear stagnation point the explicit
     seam interpolation next to the one-sided closures of the block edges
     feeds a fast-growing wall mode in the reference's discretization.

Connectivity:
  * background points with r < r_hole are holes; points with
    r_hole <= r < r_fringe receive from the O-grid (polar donor search);
  * O-grid points in the two outer rings receive from the background
    (Cartesian donor search);
  * seam columns 0, 1 and ni - 2, ni - 1 of the O-grid receive from the
    matching interior columns of the same block (coincident nodes, so the
    weights reduce to an exact copy).
The stencil width is order + 1 in each direction (overture.cpp:608-609),
placed around the receiver's continuous index and kept inside the block.

The mean flow is the incompressible potential flow past the cylinder at
free-stream speed U (c = 1), p from Bernoulli, s = 0.

Transcendental functions are applied element by element with Python's
math module for small grids (bit-reproducible inputs for fixtures) and
with numpy for large ones.
"""
import math

import numpy as np

GAMMA = 1.4
_EXACT_LIMIT = 200_000


def _ufunc(name, *arrays):
    a0 = np.asarray(arrays[0])
    if a0.size <= _EXACT_LIMIT:
        f = np.frompyfunc(getattr(math, name), len(arrays), 1)
        return np.asarray(f(*arrays), dtype=np.float64)
    return getattr(np, {"atan2": "arctan2"}.get(name, name))(*arrays)


def lagrange_weights(t, m):
    """1-D Lagrange weights of nodes 0..m-1 at positions t (array)."""
    t = np.asarray(t, np.float64)
    w = np.ones((m,) + t.shape)
    for a in range(m):
        for b in range(m):
            if b != a:
                w[a] = w[a] * ((t - b) / (a - b))
    return w


def _stencils(fi, fj, ni, nj, m):
    """Lower-left corners and tensor weights (w1 fastest) of m x m stencils
    around continuous indices (fi, fj), kept inside [0, n)."""
    i0 = np.clip(np.floor(fi).astype(np.int64) - m // 2, 0, ni - m)
    j0 = np.clip(np.floor(fj).astype(np.int64) - m // 2, 0, nj - m)
    wi = lagrange_weights(fi - i0, m)
    wj = lagrange_weights(fj - j0, m)
    coef = (wj[:, None, :] * wi[None, :, :]).reshape(m * m, -1).T  # [receiver][w2 * m + w1]
    return i0, j0, coef


def potential_flow(x, y, U, a):
    """Mean state (s, u, v, w, p) of potential flow past a cylinder of radius
    a at the origin; free stream inside the body."""
    r2 = x * x + y * y
    inside = r2 < a * a
    r2s = np.where(inside, 1.0, r2)
    a2 = a * a
    u = U * (1.0 - a2 * (x * x - y * y) / (r2s * r2s))
    v = -U * 2.0 * a2 * x * y / (r2s * r2s)
    u = np.where(inside, U, u)
    v = np.where(inside, 0.0, v)
    p = 1.0 / GAMMA + 0.5 * (U * U - (u * u + v * v))
    mean = np.zeros((5,) + x.shape)
    mean[1], mean[2], mean[4] = u, v, p
    return mean


def cylinder(nbg=(2049, 1025), extent=(-16.0, 16.0, -8.0, 8.0), nog=(513, 129), r_in=0.5, r_out=2.0,
             r_hole=1.5, fringe=5, order=4, seam_angle=math.pi / 4):
    """Grids and connectivity of the cylinder problem.  nbg, nog = (ni, nj)
    of the background and the O-grid; `fringe` background cells of
    receivers around the holes.  Returns dict(blocks=[(x, y), (x, y)],
    recv=arrays, holes=arrays, widths, order, dx, dr, dtheta)."""
    m = order + 1
    nbi, nbj = nbg
    x0, x1, y0, y1 = extent
    dx = (x1 - x0) / (nbi - 1)
    dy = (y1 - y0) / (nbj - 1)
    xb = x0 + dx * np.arange(nbi)[None, :] + 0.0 * np.arange(nbj)[:, None]
    yb = y0 + dy * np.arange(nbj)[:, None] + 0.0 * np.arange(nbi)[None, :]
    noi, noj = nog
    period = noi - 5
    dth = 2.0 * math.pi / period
    dr = (r_out - r_in) / (noj - 1)
    ii = np.arange(noi)
    ic = np.where(ii < 2, ii + period, np.where(ii > noi - 3, ii - period, ii))  # seam columns repeat
    th = seam_angle + dth * (ic - 2)
    rr = r_in + dr * np.arange(noj)
    TH, RR = np.meshgrid(th, rr)
    xo = RR * _ufunc("cos", TH)
    yo = -(RR * _ufunc("sin", TH))

    # background: holes and fringe receivers (polar donor search in the O-grid)
    rb = np.sqrt(xb * xb + yb * yb)
    hole = rb < r_hole
    r_fr = r_hole + fringe * dx
    recv_b = (rb >= r_hole) & (rb < r_fr)
    jb, ib = np.nonzero(recv_b)
    xr, yr = xb[jb, ib], yb[jb, ib]
    theta = np.mod(_ufunc("atan2", -yr, xr) - seam_angle, 2.0 * math.pi)
    fi = 2.0 + theta / dth
    fj = (np.sqrt(xr * xr + yr * yr) - r_in) / dr
    di, dj, cb = _stencils(fi, fj, noi, noj, m)

    # O-grid outer rings from the background (Cartesian donor search)
    jo, io = np.nonzero(np.arange(noj)[:, None] >= noj - 2 + 0 * ii[None, :])
    xq, yq = xo[jo, io], yo[jo, io]
    gi, gj = (xq - x0) / dx, (yq - y0) / dy
    di2, dj2, co = _stencils(gi, gj, nbi, nbj, m)

    # seam: columns 0, 1 <- ni-3, ni-4... coincident interior columns
    seam = [c for c in (0, 1, noi - 2, noi - 1)]
    js_list, is_list, si_list = [], [], []
    for c in seam:
        src = c + period if c < 2 else c - period
        for j in range(noj - 2):  # the outer rings already receive from the background
            js_list.append(j)
            is_list.append(c)
            si_list.append(src)
    js = np.array(js_list, np.int64)
    is_ = np.array(is_list, np.int64)
    fs_i = np.array(si_list, np.float64)
    fs_j = js.astype(np.float64)
    di3, dj3, cs = _stencils(fs_i, fs_j, noi, noj, m)

    recv = dict(
        rblock=np.concatenate([np.zeros(len(ib), np.int64), np.ones(len(io) + len(is_), np.int64)]),
        ri=np.concatenate([ib, io, is_]), rj=np.concatenate([jb, jo, js]),
        dblock=np.concatenate([np.ones(len(ib), np.int64), np.zeros(len(io), np.int64), np.ones(len(is_), np.int64)]),
        di=np.concatenate([di, di2, di3]), dj=np.concatenate([dj, dj2, dj3]),
        coef=np.concatenate([cb, co, cs]))
    hj, hi = np.nonzero(hole)
    holes = dict(block=np.zeros(len(hi), np.int64), i=hi, j=hj)
    return dict(blocks=[(xb, yb), (xo, yo)], recv=recv, holes=holes, widths=[(m, m, 1), (m, m, 1)], order=order,
                dx=dx, dr=dr, dtheta=dth, r_in=r_in)


def abi_tables(case):
    """Flattened tables of the C ABI (lee_overset), receivers grouped by block."""
    rc, hl = case["recv"], case["holes"]
    ni = [b[0].shape[1] for b in case["blocks"]]
    order = np.argsort(rc["rblock"], kind="stable")
    rb = rc["rblock"][order]
    db = rc["dblock"][order]
    rl = rc["rj"][order] * np.take(ni, rb) + rc["ri"][order]
    dl = rc["dj"][order] * np.take(ni, db) + rc["di"][order]
    return dict(recv_block=rb, recv_l=rl, donor_block=db, donor_l=dl, widths=[list(w) for w in case["widths"]],
                coef=np.ascontiguousarray(rc["coef"][order]), hole_block=hl["block"],
                hole_l=hl["j"] * np.take(ni, hl["block"]) + hl["i"])


def write_overture(path, case):
    """The Overture explicit-interpolation file of overture.cpp:473-680 (2-D)."""
    t = abi_tables(case)
    blocks = case["blocks"]
    nb = len(blocks)
    ni = [b[0].shape[1] for b in blocks]
    with open(path, "wb") as f:
        def ints(v):
            np.asarray(v, dtype=np.int32).tofile(f)
        ints([2, nb])
        ints([b[0].size for b in blocks])
        ints([int(np.sum(t["recv_block"] == b)) for b in range(nb)])
        ints(t["recv_block"])
        ints(np.stack([t["recv_l"] % np.take(ni, t["recv_block"]), t["recv_l"] // np.take(ni, t["recv_block"])], 1))
        ints(t["donor_block"])
        ints(np.stack([t["donor_l"] % np.take(ni, t["donor_block"]), t["donor_l"] // np.take(ni, t["donor_block"])], 1))
        for b in range(nb):
            ints(t["widths"][b])
            np.asarray(t["coef"][t["recv_block"] == b], dtype=np.float64).tofile(f)
        ints([int(np.sum(t["hole_block"] == b)) for b in range(nb)])
        for b in range(nb):
            sel = t["hole_block"] == b
            hl = t["hole_l"][sel]
            ints(np.stack([hl % ni[b], hl // ni[b], np.zeros_like(hl)], 1))
