"""Batched ensemble (lee_batch_create): independent solves of uneven sizes in
one persistent launch per RK stage, replayed from a HIP graph.

Every solve of a batch must give exactly what the same solve gives on its
own handle (same kernels' arithmetic, same host clock), whatever the tile
shape, persistent grid or launch mode; the single-handle fused path is in
turn held to the reference and the oracle by test_gpu.py, and one solve is
checked against the oracle here directly.
"""
import ctypes
import math

import numpy as np
import pytest

from camille_amd import FP32, FP64, GRID_RESIDENT, PATH_FUSED, PATH_PLAIN, Batch, LeeError, Solver, lee

GAMMA = 1.4


def _solve_inputs(ni, nj, mach, omega, pos, dx=0.5, sponge=8, amp_frac=0.03):
    """A curvilinear block with uniform mean flow `mach` in x, a monopole at
    `pos` (fractions of the extent) and sponges on all faces."""
    i = np.arange(ni)
    j = np.arange(nj)
    I, J = np.meshgrid(i, j)
    lx, ly = dx * (ni - 1), dx * (nj - 1)
    x = dx * I + amp_frac * lx * np.sin(2 * np.pi * J / (nj - 1))
    y = dx * J + amp_frac * ly * np.sin(2 * np.pi * I / (ni - 1))
    mean = np.zeros((5, nj, ni))
    mean[1] = mach
    mean[4] = 1.0 / GAMMA
    x0, y0 = pos[0] * lx, pos[1] * ly
    src = (1e-3, omega, x0, y0, 0.0, 4 * dx)
    w = sponge - 1
    bc = [(0, "sponge_freund_ambient", "+xi", "0", str(w), "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "-xi", f"-{w}", "-0", "0", "-0", "0", "0"),
          (0, "sponge_freund_ambient", "+eta", "0", "-0", "0", str(w), "0", "0"),
          (0, "sponge_freund_ambient", "-eta", "0", "-0", f"-{w}", "-0", "0", "0")]
    ic = np.zeros((5, nj, ni))
    return dict(x=x, y=y, mean=mean, boundaries=bc, simulation="CASE_MONOPOLE", source_params=src,
                mode="FIX_CFL", value=0.5, ic=ic)


def _ensemble(seed=7, n=5):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ni = int(rng.integers(33, 150))
        nj = int(rng.integers(33, 90))
        out.append(_solve_inputs(ni, nj, float(rng.uniform(0.0, 0.6)), float(rng.uniform(0.1, 0.5) * math.pi),
                                 rng.uniform(0.2, 0.8, 2)))
    return out


def _single(sv, prec, path=PATH_FUSED):
    s = Solver([dict(x=sv["x"], y=sv["y"], mean=sv["mean"])], boundaries=sv["boundaries"], mode=sv["mode"],
               value=sv["value"], simulation=sv["simulation"], source_params=sv["source_params"], precision=prec,
               path=path)
    s.set_solution(sv["ic"])
    return s


def _batch(solves, prec):
    b = Batch(solves, precision=prec)
    for k, sv in enumerate(solves):
        b.set_solution(sv["ic"], k)
    return b


# ---------------------------------------------------------------- CPU checks

def _cfg(**kw):
    return lee._make_config(**kw)


def _create_batch(cfgs, n=16):
    """lee_batch_create on two 16x16 blocks; these cases fail validation
    before any HIP call."""
    L = lee.library()
    xyz = np.zeros((3, n, n))
    mean = np.zeros((5, n, n))
    blk = (lee._Block * 2)()
    for k in range(2):
        blk[k].ni, blk[k].nj, blk[k].nk = n, n, 1
        blk[k].xyz = xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        blk[k].mean = mean.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    arr = (lee._Config * 2)(*cfgs)
    h = ctypes.c_void_p()
    rc = L.lee_batch_create(arr, blk, 2, ctypes.byref(h))
    return rc, L.lee_last_error(None).decode(), h


def test_batch_create_rejects_mixed_schemes():
    rc, err, h = _create_batch([_cfg(filter_order=8), _cfg(filter_order=4)])
    assert rc == 1 and "share the scheme" in err and not h.value
    rc, err, _ = _create_batch([_cfg(precision=FP32), _cfg(precision=FP64)])
    assert rc == 1 and "share the scheme" in err


def test_batch_create_rejects_plain_path_and_no_filter():
    rc, err, _ = _create_batch([_cfg(path=PATH_PLAIN), _cfg(path=PATH_PLAIN)])
    assert rc == 1 and "fused path" in err
    rc, err, _ = _create_batch([_cfg(filter_order=0), _cfg(filter_order=0)])
    assert rc == 1 and "need the filter" in err


def test_batch_create_rejects_empty_batch():
    L = lee.library()
    arr = (lee._Config * 1)(_cfg())
    assert L.lee_batch_create(arr, None, 0, ctypes.byref(ctypes.c_void_p())) == 1
    assert "empty batch" in L.lee_last_error(None).decode()


# ---------------------------------------------------------------- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("prec", [FP64, FP32])
def test_batch_equals_single_handles_bitwise(prec):
    solves = _ensemble()
    nsteps = 40
    b = _batch(solves, prec)
    b.solve(nsteps)
    print(f"batch prec {prec}: tile {b.tile_shape()} launch {b.launch()}", flush=True)
    for k, sv in enumerate(solves):
        s = _single(sv, prec)
        s.solve(nsteps)
        assert b.clock(k) == s.clock(), k
        assert np.array_equal(b.solution(k), s.solution()), k
        s.close()
    b.close()


@pytest.mark.gpu
def test_batch_launch_modes_bit_identical():
    """Graph replay or direct launches, any persistent grid, tile shape or
    column-march stack: the same bits; also two solve calls equal one."""
    solves = _ensemble(seed=11, n=4)
    ref = None
    for use_graph, gs, gf, shape, march in [
            (True, 0, 0, (1, 4), None), (False, 0, 0, (1, 4), None), (True, 256, 512, (2, 8), None),
            (True, 1024, 0, (4, 16), None), (False, 8, 8, (1, 8), None), (True, 0, 64, (2, 12), None),
            (True, 0, 0, (1, 4), (4, 8)), (False, 0, 24, (2, 8), (4, 4)), (True, 0, 512, (1, 4), (4, 16)),
            (True, GRID_RESIDENT, GRID_RESIDENT, (2, 8), None), (False, GRID_RESIDENT, GRID_RESIDENT, (4, 4), (4, 8))]:
        b = _batch(solves, FP32)
        b.set_tile_shape(*shape)
        if march:
            b.set_filter_march(*march)
            assert b.filter_march() == march
        b.set_launch(use_graph, gs, gf)
        b.solve(7)
        b.solve(13)
        sols = [b.solution(k) for k in range(len(solves))]
        print(f"graph {use_graph} grids {gs},{gf} tile {shape} march {march}: {b.launch()}", flush=True)
        if ref is None:
            ref = sols
        for k in range(len(solves)):
            assert np.array_equal(sols[k], ref[k]), (use_graph, gs, gf, shape, march, k)
        b.close()
    b = _batch(solves, FP32)
    with pytest.raises(Exception):
        b.set_filter_march(1, 8)  # one launch covers every solve: stacks of 4 waves only
    b.solve(20)
    for k in range(len(solves)):
        assert np.array_equal(b.solution(k), ref[k])
    b.close()


@pytest.mark.gpu
def test_batch_persistent_queue_small_lists():
    """The persistent stage's per-eighth work queue on short tile lists:
    two small solves give eighths with few or no entries (some workgroups
    find theirs empty at once), grids of 1 to 4 workgroups per eighth and
    the resident grid deal out and draw every tile exactly once, on every
    stage of 9 steps and again on the next solve call (the counters reset
    themselves): the same bits as one workgroup per tile."""
    solves = [_solve_inputs(33, 33, 0.3, 0.3 * math.pi, (0.5, 0.5)),
              _solve_inputs(41, 35, 0.1, 0.45 * math.pi, (0.3, 0.6))]
    ref = None
    # 18 tiles of 4 rows (eighths of 3, the last two empty) or 10 of 8 rows
    # (eighths of 2, the last three empty); grids from the list's length up
    # launch one workgroup per tile
    for gs, shape in [(0, (1, 4)), (8, (1, 4)), (16, (1, 4)), (8, (2, 8)), (GRID_RESIDENT, (1, 4))]:
        b = _batch(solves, FP32)
        b.set_tile_shape(*shape)
        b.set_launch(True, gs, 0)
        b.solve(4)
        b.solve(5)
        sols = [b.solution(k) for k in range(len(solves))]
        b.close()
        if ref is None:
            ref = sols
            assert np.max(np.abs(ref[0][4])) > 0
        for k in range(len(solves)):
            assert np.array_equal(sols[k], ref[k]), (gs, shape, k)
    print("persistent queue on short lists: bit-identical to one workgroup per tile", flush=True)


@pytest.mark.gpu
def test_batch_schedule_chunks():
    """More steps than one schedule chunk (2048): small solves, compared with
    single handles after 2100 steps."""
    solves = [_solve_inputs(33, 33, 0.3, 0.3 * math.pi, (0.5, 0.5)),
              _solve_inputs(41, 35, 0.1, 0.45 * math.pi, (0.3, 0.6))]
    b = _batch(solves, FP32)
    b.solve(2100)
    for k, sv in enumerate(solves):
        s = _single(sv, FP32)
        s.solve(2100)
        assert b.clock(k) == s.clock()
        assert np.array_equal(b.solution(k), s.solution())
        s.close()
    b.close()


@pytest.mark.gpu
def test_batch_solve_matches_oracle():
    """One solve of a batch against the CPU oracle (FP64 <= 1e-12, FP32
    bulk <= 1e-5, relative to max |p|)."""
    from oracle.pyoracle import Oracle
    solves = _ensemble(seed=3, n=3)
    sv = solves[1]
    scheme = dict(fd_order=4, filter_order=8, filter_blend=1.0, rk=4, sponge_order=2, sponge_constant=5.0)
    o = Oracle([dict(x=sv["x"], y=sv["y"], mean=sv["mean"], ic=sv["ic"])], scheme, sv["boundaries"],
               (sv["simulation"], sv["source_params"]), sv["mode"], sv["value"])
    o.run(60)
    ref = o.solution(0)[:, 0]
    for prec, tol in ((FP64, 1e-12), (FP32, 1e-5)):
        b = _batch(solves, prec)
        b.solve(60)
        assert b.clock(1)[1] == o.time
        p = b.solution(1)[4]
        err = float(np.max(np.abs(p - ref[4])) / np.max(np.abs(ref[4])))
        print(f"batch solve vs oracle prec {prec}: {err:.2e}", flush=True)
        assert err <= tol
        b.close()


@pytest.mark.gpu
def test_batch_rejects_nonzero_w():
    solves = _ensemble(n=2)
    b = Batch(solves, precision=FP32)
    ic = solves[0]["ic"].copy()
    ic[3, 5, 5] = 1e-3
    with pytest.raises(LeeError):
        b.set_solution(ic, 0)
    b.close()
