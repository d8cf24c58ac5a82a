"""ctypes front end of the CPU oracle (oracle/liblee_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and
by the CPU-baseline leg of bench.py.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblee_oracle.so")

BC_TYPES = {
    "dirichlet_allzero": 1,
    "dirichlet_harmonicwave": 2,
    "neumann": 100,
    "wall_slip_kinematic": 201,
    "centerline_cart_norm2x": 900,
    "centerline_cart_norm2y": 901,
    "centerline_cart_norm2z": 902,
    "centerline_axisym": 903,
    "sponge_freund_ambient": 1000,
    "sponge_freund_harmonicwave": 1002,
}
WAVES = {"NONE": 0, "WAVE_ACOUSTIC": 1, "WAVE_ENTROPY": 2}
WAVEFORMS = {"WAVEFORM_PLANE": 1, "WAVEFORM_HOMOGENEOUS": 2, "WAVEFORM_GAUSSIAN_ROUNDJET": 3}
# physical source per SIMULATION (apply_physicalSource, source.cpp:22-107);
# the cases without a source term (TPN49's is empty, the nozzle has none)
SRC = {"CASE_TANNA_TPN49": 0, "CASE_LINEAR_NOZZLE": 0, "CASE_SCATTERING_TWOCYLINDER": 1,
       "CASE_2DJET_WITH_A_HARMONIC_SOURCE": 2, "CASE_MONOPOLE": 3}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: run `make -C {HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp, i, d = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, d, i, i, d, d, i, i, d]
        L.orc_set_block.argtypes = [vp, i, i, i, i, dp, dp]
        L.orc_add_boundary.argtypes = [vp, i, i, i, i, ip, ip]
        L.orc_set_source.argtypes = [vp, i, dp]
        L.orc_set_axisym.argtypes = [vp, i]
        L.orc_set_scalars.argtypes = [vp, i]
        L.orc_set_timestep.argtypes = [vp, i, d]
        L.orc_set_harmonic.argtypes = [vp, i, i, i, d, d, d]
        L.orc_set_overset.argtypes = [vp, i, ip, ip, ip, ip, ip, dp, i, i, ip, ip]
        L.orc_init.argtypes = [vp, dp, i, d]
        L.orc_step.argtypes = [vp]
        L.orc_run.argtypes = [vp, i]
        for f in ("orc_time", "orc_dt", "orc_cfl"):
            getattr(L, f).restype = d
            getattr(L, f).argtypes = [vp]
        L.orc_time_step.argtypes = [vp]
        L.orc_get_solution.argtypes = [vp, i, dp]
        L.orc_set_solution.argtypes = [vp, i, dp]
        L.orc_set_time.argtypes = [vp, i, d, d]
        L.orc_get_metrics.argtypes = [vp, i, dp, dp, dp]
        L.orc_get_mean_gradient.argtypes = [vp, i, dp]
        L.orc_get_aux.argtypes = [vp, i, dp]
        L.orc_get_iblank.argtypes = [vp, i, ip]
        L.orc_get_sigma.argtypes = [vp, i, dp]
        L.orc_rhs.argtypes = [vp, i, dp, dp, d, i]
        L.orc_filter.argtypes = [vp, i, dp]
        L.orc_apply_dir.argtypes = [vp, i, i, dp, dp, i]
        L.orc_deriv_xyz.argtypes = [vp, i, dp, dp, i]
        L.orc_set_iblank.argtypes = [vp, i, ip]
        L.orc_destroy.argtypes = [vp]
        L.orc_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def parse_bc_line(tokens, dims):
    """One BC-file line (io.cpp:1158-1268) -> (block, type, dir, end, is[3], ie[3]).
    dims = (ni, nj, nk) of that block."""
    blk, typ, side = int(tokens[0]), str(tokens[1]).lower(), str(tokens[2])
    end = 0 if side[0] == "+" else 1
    dname = side[1:].lower()
    d = {"xi": 0, "eta": 1, "zeta": 2}[dname]
    idx = []
    for t, q in enumerate(tokens[3:9]):
        s = str(q)
        v = int(s.lstrip("+-"))
        if s.startswith("-"):
            v = (dims[t // 2] - 1) - v
        idx.append(v)
    is3 = [idx[0], idx[2], idx[4]]
    ie3 = [idx[1], idx[3], idx[5]]
    return blk, BC_TYPES[typ], d, end, is3, ie3


class Oracle:
    """One multi-block LEE problem on the CPU, FP64, reference semantics."""

    def __init__(self, blocks, scheme, bc_lines=(), source=None, mode="FIX_CFL", value=0.5,
                 overset=None, ndim=2, gamma=1.4, harmonic=None, num_scalar=0):
        """blocks: list of dict(x, y[, z], mean (nv,...), ic (nv,...)) arrays shaped (nk?, nj, ni),
        nv = 5 + num_scalar (LEE_SCALAR: passive scalars after p')
        scheme: dict(fd_order, filter_order, filter_blend, rk, sponge_order, sponge_constant
        [, filter_strength, axisym]); order 11 selects FDO11P / SFO11P
        source: None or (kind_name, params[8])."""
        L = lib()
        self.L = L
        self.h = L.orc_create(len(blocks), ndim, gamma, scheme["fd_order"], scheme["filter_order"],
                              scheme.get("filter_strength", 1.0),
                              scheme.get("filter_blend", 1.0), scheme["rk"], scheme["sponge_order"],
                              scheme["sponge_constant"])
        if not self.h:
            raise RuntimeError(L.orc_error().decode())
        self.nv = 5 + int(num_scalar)
        if num_scalar and L.orc_set_scalars(self.h, int(num_scalar)):
            raise RuntimeError(L.orc_error().decode())
        if scheme.get("axisym") and L.orc_set_axisym(self.h, 1):
            raise RuntimeError(L.orc_error().decode())
        self.shapes = []
        for ib, b in enumerate(blocks):
            x = np.asarray(b["x"], dtype=np.float64)
            shp = x.shape if x.ndim == 3 else (1,) + x.shape
            nk, nj, ni = shp
            z = b.get("z")
            z = np.zeros(shp) if z is None else np.asarray(z, np.float64).reshape(shp)
            xyz = np.ascontiguousarray(np.stack([x.reshape(shp), np.asarray(b["y"], np.float64).reshape(shp), z]))
            mean = np.ascontiguousarray(np.asarray(b["mean"], np.float64).reshape((self.nv,) + shp))
            L.orc_set_block(self.h, ib, ni, nj, nk, _dp(xyz), _dp(mean))
            self.shapes.append(shp)
        for ln in bc_lines:
            blk = int(ln[0])
            nk, nj, ni = self.shapes[blk]
            _, typ, d, end, is3, ie3 = parse_bc_line(ln, (ni, nj, nk))
            a, bb = np.array(is3, np.int32), np.array(ie3, np.int32)
            if L.orc_add_boundary(self.h, blk, typ, d, end, _ip(a), _ip(bb)):
                raise RuntimeError(L.orc_error().decode())
        if source is not None:
            kind, params = source
            p = np.zeros(8)
            p[:len(params)] = params
            L.orc_set_source(self.h, SRC[kind], _dp(p))
        L.orc_set_timestep(self.h, 1 if mode == "FIX_CFL" else 0, float(value))
        if harmonic:
            L.orc_set_harmonic(self.h, WAVES[harmonic["wave"]], WAVEFORMS.get(harmonic.get("shape", ""), 0),
                               "XYZ".index(str(harmonic.get("direction", "X")).upper()),
                               float(harmonic.get("amplitude", 0.0)), float(harmonic.get("wavelength", 1.0)),
                               float(harmonic.get("halfwidth", 1.0)))
        if overset is not None:
            self._set_overset(overset)
        ic = np.concatenate([np.asarray(b["ic"], np.float64).reshape(-1) for b in blocks])
        ic = np.ascontiguousarray(ic)
        if L.orc_init(self.h, _dp(ic), 0, 0.0):
            raise RuntimeError(L.orc_error().decode())

    def _set_overset(self, ov):
        rb = np.ascontiguousarray(ov["recv_block"], np.int32)
        rl = np.ascontiguousarray(ov["recv_l"], np.int32)
        db = np.ascontiguousarray(ov["donor_block"], np.int32)
        dl = np.ascontiguousarray(ov["donor_l"], np.int32)
        w = np.ascontiguousarray(ov["widths"], np.int32).reshape(-1)
        c = np.ascontiguousarray(ov["coef"], np.float64)
        hb = np.ascontiguousarray(ov.get("hole_block", np.zeros(0)), np.int32)
        hl = np.ascontiguousarray(ov.get("hole_l", np.zeros(0)), np.int32)
        self.L.orc_set_overset(self.h, len(rb), _ip(rb), _ip(rl), _ip(db), _ip(dl), _ip(w), _dp(c),
                               c.shape[1], len(hb), _ip(hb), _ip(hl))

    @classmethod
    def from_case(cls, case, **kw):
        src = None
        if SRC[case["simulation"]] != 0:
            src = (case["simulation"], case.get("source_params", ()))
        value = case["cfl"] if case["mode"] == "FIX_CFL" else case["dt"]
        blocks = [dict(x=case["x"], y=case["y"], mean=case["mean"], ic=case["ic"])]
        scheme = {k: case[k] for k in ("fd_order", "filter_order", "filter_blend", "rk", "sponge_order",
                                       "sponge_constant")}
        scheme["filter_strength"] = case.get("filter_strength", 1.0)
        scheme["axisym"] = case.get("axisym", 0)
        return cls(blocks, scheme, case["bc"], src, case["mode"], value, harmonic=case.get("harmonic"),
                   num_scalar=case.get("num_scalar", 0), **kw)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def n(self, ib=0):
        nk, nj, ni = self.shapes[ib]
        return ni * nj * nk

    def run(self, nsteps):
        if self.L.orc_run(self.h, int(nsteps)):
            raise RuntimeError(self.L.orc_error().decode())

    def step(self):
        if self.L.orc_step(self.h):
            raise RuntimeError(self.L.orc_error().decode())

    @property
    def time(self):
        return self.L.orc_time(self.h)

    @property
    def dt(self):
        return self.L.orc_dt(self.h)

    @property
    def time_step(self):
        return self.L.orc_time_step(self.h)

    def cfl(self):
        return self.L.orc_cfl(self.h)

    def solution(self, ib=0):
        out = np.empty((self.nv,) + self.shapes[ib])
        self.L.orc_get_solution(self.h, ib, _dp(out))
        return out

    def set_solution(self, sol, ib=0):
        a = np.ascontiguousarray(sol, np.float64)
        self.L.orc_set_solution(self.h, ib, _dp(a))

    def set_time(self, step, time, dt):
        self.L.orc_set_time(self.h, int(step), float(time), float(dt))

    def metrics(self, ib=0):
        n = self.n(ib)
        met, minv, jac = np.empty(9 * n), np.empty(9 * n), np.empty(n)
        self.L.orc_get_metrics(self.h, ib, _dp(met), _dp(minv), _dp(jac))
        shp = self.shapes[ib]
        return met.reshape((3, 3) + shp), minv.reshape((3, 3) + shp), jac.reshape(shp)

    def mean_gradient(self, ib=0, ndim=2):
        out = np.empty((self.nv * ndim,) + self.shapes[ib])
        self.L.orc_get_mean_gradient(self.h, ib, _dp(out))
        return out

    def aux(self, ib=0):
        out = np.empty((4,) + self.shapes[ib])
        self.L.orc_get_aux(self.h, ib, _dp(out))
        return out

    def iblank(self, ib=0):
        out = np.empty(self.shapes[ib], np.int32)
        self.L.orc_get_iblank(self.h, ib, _ip(out))
        return out

    def sigma(self, ib=0):
        out = np.empty(self.shapes[ib])
        self.L.orc_get_sigma(self.h, ib, _dp(out))
        return out

    def rhs(self, y, time, ib=0, aux_from_y=True):
        y = np.ascontiguousarray(y, np.float64)
        out = np.empty_like(y)
        if self.L.orc_rhs(self.h, ib, _dp(y), _dp(out), float(time), 1 if aux_from_y else 0):
            raise RuntimeError(self.L.orc_error().decode())
        return out

    def filter(self, y, ib=0):
        y = np.array(y, np.float64, copy=True, order="C")
        if self.L.orc_filter(self.h, ib, _dp(y)):
            raise RuntimeError(self.L.orc_error().decode())
        return y

    def apply_dir(self, f, which, direction, ib=0):
        f = np.ascontiguousarray(f, np.float64)
        out = np.empty_like(f)
        if self.L.orc_apply_dir(self.h, ib, which, _dp(f), _dp(out), direction):
            raise RuntimeError(self.L.orc_error().decode())
        return out

    def deriv_xyz(self, f, direction, ib=0):
        f = np.ascontiguousarray(f, np.float64)
        out = np.empty_like(f)
        if self.L.orc_deriv_xyz(self.h, ib, _dp(f), _dp(out), direction):
            raise RuntimeError(self.L.orc_error().decode())
        return out

    def set_iblank(self, ibk, ib=0):
        a = np.ascontiguousarray(ibk, np.int32)
        self.L.orc_set_iblank(self.h, ib, _ip(a))
