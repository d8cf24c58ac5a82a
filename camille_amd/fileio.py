"""Python view of the host file layer (include/lee_io.h): the reference's
input deck, PLOT3D files, BC file and Overture file, read and written by
the C++ code of liblee_amd.so, and lee_run_deck, the reference's step loop
over a case folder.

Every function raises LeeError with the library's message on failure.
"""
import ctypes

import numpy as np

from .lee import LeeError, _Boundary, _Config, _Overset, library

GRID, SOLUTION, FUNCTION = 0, 1, 2


def SOLUTION_VARS(nv):
    """Kind of a solution file whose blocks hold nv variables after tau (the
    restart file of a LEE_SCALAR run)."""
    return 100 + int(nv)


class _Settings(ctypes.Structure):
    _fields_ = [("cfg", _Config), ("num_time_steps", ctypes.c_int), ("report_freq", ctypes.c_int),
                ("time_writing_solutions", ctypes.c_double), ("overset", ctypes.c_int),
                ("overset_order", ctypes.c_int), ("grid_file", ctypes.c_char * 512),
                ("mean_file", ctypes.c_char * 512), ("bc_file", ctypes.c_char * 512),
                ("overset_file", ctypes.c_char * 512)]


_ready = False


def _lib():
    global _ready
    L = library()
    if not _ready:
        vp, i, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.lee_io_last_error.restype = cp
        L.lee_deck_open.argtypes = [cp, ctypes.POINTER(vp)]
        L.lee_deck_close.argtypes = [vp]
        L.lee_deck_close.restype = None
        L.lee_deck_get.argtypes = [vp, cp, cp, i, cp, i]
        L.lee_deck_settings.argtypes = [vp, ctypes.POINTER(_Settings)]
        L.lee_plot3d_dims.argtypes = [cp, i, ip, ip, i]
        L.lee_plot3d_read.argtypes = [cp, i, i, dp, dp]
        L.lee_plot3d_write.argtypes = [cp, i, i, ip, ctypes.POINTER(dp), dp]
        L.lee_bc_file_read.argtypes = [cp, i, ip, ctypes.POINTER(_Boundary), ip, i, ip]
        L.lee_overture_open.argtypes = [cp, i, i, ip, i, ctypes.POINTER(vp)]
        L.lee_overture_tables.argtypes = [vp]
        L.lee_overture_tables.restype = ctypes.POINTER(_Overset)
        L.lee_overture_close.argtypes = [vp]
        L.lee_overture_close.restype = None
        L.lee_run_deck.argtypes = [cp, cp, i, i, cp, i]
        _ready = True
    return L


def _check(rc):
    if rc != 0:
        raise LeeError(_lib().lee_io_last_error().decode())


def _b(s):
    return None if s is None else str(s).encode()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Deck:
    """An input deck read by lee_deck_open (input.cpp:870-999 rules)."""

    def __init__(self, path):
        self._h = ctypes.c_void_p()
        _check(_lib().lee_deck_open(_b(path), ctypes.byref(self._h)))

    def get(self, name, keyword=None, index=0):
        buf = ctypes.create_string_buffer(512)
        _check(_lib().lee_deck_get(self._h, _b(name), _b(keyword), index, buf, 512))
        return buf.value.decode()

    def settings(self):
        s = _Settings()
        _check(_lib().lee_deck_settings(self._h, ctypes.byref(s)))
        c = s.cfg
        cfg = {f: getattr(c, f) for f, _ in _Config._fields_ if f != "source_params"}
        return dict(cfg=cfg, num_time_steps=s.num_time_steps, report_freq=s.report_freq,
                    time_writing_solutions=s.time_writing_solutions, overset=s.overset,
                    overset_order=s.overset_order, grid_file=s.grid_file.decode(), mean_file=s.mean_file.decode(),
                    bc_file=s.bc_file.decode(), overset_file=s.overset_file.decode())

    def close(self):
        if self._h:
            _lib().lee_deck_close(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plot3d_dims(path, kind):
    """[(ni, nj, nk, nvar)] per block."""
    nb = ctypes.c_int()
    _check(_lib().lee_plot3d_dims(_b(path), kind, ctypes.byref(nb), None, 0))
    dims = np.zeros(4 * nb.value, dtype=np.int32)
    _check(_lib().lee_plot3d_dims(_b(path), kind, ctypes.byref(nb), _ip(dims), nb.value))
    return [tuple(int(v) for v in dims[4 * b:4 * b + 4]) for b in range(nb.value)]


def plot3d_read(path, kind, iblock):
    """(nvar, nk, nj, ni) array of block iblock, and tau for a solution file
    (SOLUTION or SOLUTION_VARS(nv))."""
    dims = plot3d_dims(path, kind)
    if not 0 <= iblock < len(dims):
        raise LeeError(f"block {iblock} is not in {path}")
    ni, nj, nk, nv = dims[iblock]
    out = np.empty((nv, nk, nj, ni))
    tau = np.zeros(4)
    _check(_lib().lee_plot3d_read(_b(path), kind, iblock, _dp(out), _dp(tau)))
    return (out, tau) if kind == SOLUTION or kind > 100 else out


def plot3d_write(path, kind, blocks, tau=None):
    """blocks: list of (nvar, nk, nj, ni) arrays."""
    arrs = [np.ascontiguousarray(b, dtype=np.float64) for b in blocks]
    dims = np.array([[a.shape[3], a.shape[2], a.shape[1], a.shape[0]] for a in arrs], dtype=np.int32).reshape(-1)
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dp(a) for a in arrs])
    t = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64)
    _check(_lib().lee_plot3d_write(_b(path), kind, len(arrs), _ip(dims), ptrs, None if t is None else _dp(t)))


def bc_file_read(path, dims):
    """dims: [(ni, nj, nk)] per block -> [(iblock, type, dir, end, is[3], ie[3])]."""
    d = np.array(dims, dtype=np.int32).reshape(-1)
    cap = 4096
    regs = (_Boundary * cap)()
    blk = np.zeros(cap, dtype=np.int32)
    cnt = ctypes.c_int()
    _check(_lib().lee_bc_file_read(_b(path), len(dims), _ip(d), regs, _ip(blk), cap, ctypes.byref(cnt)))
    return [(int(blk[r]), regs[r].type, regs[r].dir, regs[r].end, list(regs[r].is_), list(regs[r].ie))
            for r in range(cnt.value)]


def overture_read(path, num_dim, dims, order):
    """The Overture file as the flattened lee_overset tables (dict of arrays)."""
    d = np.array(dims, dtype=np.int32).reshape(-1)
    h = ctypes.c_void_p()
    _check(_lib().lee_overture_open(_b(path), num_dim, len(dims), _ip(d), order, ctypes.byref(h)))
    try:
        t = _lib().lee_overture_tables(h).contents
        nr, nh, st = t.nrecv, t.nhole, t.coef_stride

        def arr(p, n, dt=np.int32):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy() if n else np.zeros(0, dt)
        return dict(recv_block=arr(t.recv_block, nr), recv_l=arr(t.recv_l, nr), donor_block=arr(t.donor_block, nr),
                    donor_l=arr(t.donor_l, nr), widths=arr(t.widths, 3 * len(dims)).reshape(-1, 3),
                    coef=arr(t.coef, nr * st, np.float64).reshape(nr, st), hole_block=arr(t.hole_block, nh),
                    hole_l=arr(t.hole_l, nh))
    finally:
        _lib().lee_overture_close(h)


def run_deck(deck, restart=None, precision=0, path=0, out_dir=".", quiet=True):
    """lee_run_deck: the reference's run of a case folder through the GPU
    advance; writes out_dir/camille.NNNNNN.q and the report files."""
    _check(_lib().lee_run_deck(_b(deck), _b(restart), precision, path, _b(out_dir), 1 if quiet else 0))
