"""ctypes binding of the C ABI in include/lee.h, plus a thin Python mirror of
the reference's solver call sequence.

The shared library camille_amd/liblee_amd.so is built ahead of time
(python -m camille_amd.build); importing this module never compiles and
loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblee_amd.so")

FP64, FP32 = 0, 1
PATH_FUSED, PATH_PLAIN = 0, 1
GRID_RESIDENT = -1  # batch launch grid: the workgroups the device holds at once (include/lee.h)

# BC-file type names (src/io/io.cpp:1273-1320) -> src/core/param.h codes
BC_TYPES = {
    "dirichlet": 0,
    "dirichlet_allzero": 1,
    "dirichlet_harmonicwave": 2,
    "neumann": 100,
    "wall_slip_kinematic": 201,
    "centerline_cart_norm2x": 900,
    "centerline_cart_norm2y": 901,
    "centerline_cart_norm2z": 902,
    "centerline_axisym": 903,
    "sponge_freund_ambient": 1000,
    "sponge_freund_dirichlet": 1001,
    "sponge_freund_harmonicwave": 1002,
}
# HARMONIC_WAVE type and SHAPE names (input.cpp:219-231)
WAVES = {"NONE": 0, "WAVE_ACOUSTIC": 1, "WAVE_ENTROPY": 2}
WAVEFORMS = {"WAVEFORM_PLANE": 1, "WAVEFORM_HOMOGENEOUS": 2, "WAVEFORM_GAUSSIAN_ROUNDJET": 3}
# SIMULATION key -> physical source
SOURCES = {
    "SOURCE_FIELD": 4,
    "CASE_TANNA_TPN49": 0,
    "CASE_KBK_COMBUSTOR": 0,
    "CASE_LINEAR_NOZZLE": 0,
    "CASE_SCATTERING_TWOCYLINDER": 1,
    "CASE_2DJET_WITH_A_HARMONIC_SOURCE": 2,
    "CASE_MONOPOLE": 3,
}

NV = 5


class LeeError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [
        ("num_dim", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("fd_order", ctypes.c_int),
        ("filter_order", ctypes.c_int),
        ("filter_blend", ctypes.c_double),
        ("rk_stages", ctypes.c_int),
        ("fix_cfl", ctypes.c_int),
        ("cfl_or_dt", ctypes.c_double),
        ("sponge_order", ctypes.c_int),
        ("sponge_constant", ctypes.c_double),
        ("source_kind", ctypes.c_int),
        ("source_params", ctypes.c_double * 8),
        ("precision", ctypes.c_int),
        ("path", ctypes.c_int),
        ("filter_strength", ctypes.c_double),
        ("axisym", ctypes.c_int),
        ("harmonic_wave", ctypes.c_int),
        ("harmonic_shape", ctypes.c_int),
        ("harmonic_direction", ctypes.c_int),
        ("harmonic_amplitude", ctypes.c_double),
        ("harmonic_wavelength", ctypes.c_double),
        ("harmonic_halfwidth", ctypes.c_double),
        ("num_scalar", ctypes.c_int),
    ]


class _Boundary(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("dir", ctypes.c_int), ("end", ctypes.c_int),
                ("is_", ctypes.c_int * 3), ("ie", ctypes.c_int * 3)]


class _Block(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int),
                ("xyz", ctypes.POINTER(ctypes.c_double)), ("mean", ctypes.POINTER(ctypes.c_double)),
                ("nboundary", ctypes.c_int), ("boundary", ctypes.POINTER(_Boundary))]


class _Overset(ctypes.Structure):
    _fields_ = [("nrecv", ctypes.c_int),
                ("recv_block", ctypes.POINTER(ctypes.c_int)), ("recv_l", ctypes.POINTER(ctypes.c_int)),
                ("donor_block", ctypes.POINTER(ctypes.c_int)), ("donor_l", ctypes.POINTER(ctypes.c_int)),
                ("widths", ctypes.POINTER(ctypes.c_int)), ("coef", ctypes.POINTER(ctypes.c_double)),
                ("coef_stride", ctypes.c_int), ("nhole", ctypes.c_int),
                ("hole_block", ctypes.POINTER(ctypes.c_int)), ("hole_l", ctypes.POINTER(ctypes.c_int))]


EXPORTS = [
    "lee_create", "lee_destroy", "lee_last_error", "lee_set_solution", "lee_get_solution", "lee_set_time",
    "lee_get_time", "lee_compute_cfl", "lee_time_advance", "lee_solve", "lee_synchronize", "lee_compute_rhs",
    "lee_apply_filter", "lee_take_derivative_xyz", "lee_update_boundary", "lee_get_metrics", "lee_get_iblank",
    "lee_get_mean_gradient", "lee_get_sponge", "lee_last_solve_ms", "lee_version",
    "lee_set_kernel_timing", "lee_get_kernel_timing", "lee_autotune", "lee_set_tile_shape", "lee_get_tile_shape",
    "lee_set_filter_grid", "lee_get_filter_grid", "lee_set_filter_march", "lee_get_filter_march",
    "lee_batch_create", "lee_get_solve_time", "lee_set_batch_launch", "lee_get_batch_launch",
    "lee_set_source_field", "lee_set_observer", "lee_objective_forward", "lee_objective_backward",
    "lee_get_source_gradient", "lee_get_state_gradient", "lee_compute_rhs_adjoint", "lee_apply_filter_adjoint",
    "lee_sample_observers", "lee_update_boundary_noalloc", "lee_update_boundary_adjoint",
]

_lib = None


def library():
    """Loads liblee_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LeeError(f"{LIB_PATH} is missing: build it with `python -m camille_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.lee_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(_Block), i, ctypes.POINTER(_Overset),
                             ctypes.POINTER(vp)]
    L.lee_destroy.argtypes = [vp]
    L.lee_destroy.restype = None
    L.lee_last_error.argtypes = [vp]
    L.lee_last_error.restype = ctypes.c_char_p
    L.lee_version.restype = ctypes.c_char_p
    L.lee_set_solution.argtypes = [vp, i, dp]
    L.lee_get_solution.argtypes = [vp, i, dp]
    L.lee_set_time.argtypes = [vp, i, d, d]
    L.lee_get_time.argtypes = [vp, ip, dp, dp]
    L.lee_compute_cfl.argtypes = [vp, dp]
    L.lee_time_advance.argtypes = [vp]
    L.lee_solve.argtypes = [vp, i]
    L.lee_synchronize.argtypes = [vp]
    L.lee_compute_rhs.argtypes = [vp, i, dp, d, dp]
    L.lee_apply_filter.argtypes = [vp, i, dp]
    L.lee_take_derivative_xyz.argtypes = [vp, i, dp, i, dp]
    L.lee_update_boundary.argtypes = [vp, ctypes.POINTER(dp)]
    L.lee_update_boundary_noalloc.argtypes = [vp, ctypes.POINTER(vp)]
    L.lee_update_boundary_adjoint.argtypes = [vp, ctypes.POINTER(dp)]
    L.lee_get_metrics.argtypes = [vp, i, dp, dp]
    L.lee_get_iblank.argtypes = [vp, i, ip]
    L.lee_get_mean_gradient.argtypes = [vp, i, dp]
    L.lee_get_sponge.argtypes = [vp, i, dp]
    L.lee_last_solve_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.lee_set_kernel_timing.argtypes = [vp, i]
    L.lee_get_kernel_timing.argtypes = [vp, dp, ip]
    L.lee_autotune.argtypes = [vp]
    L.lee_set_tile_shape.argtypes = [vp, i, i, i]
    L.lee_get_tile_shape.argtypes = [vp, i, ip, ip]
    L.lee_set_filter_grid.argtypes = [vp, i, i]
    L.lee_get_filter_grid.argtypes = [vp, i, ip]
    L.lee_set_filter_march.argtypes = [vp, i, i, i]
    L.lee_get_filter_march.argtypes = [vp, i, ip, ip]
    L.lee_batch_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(_Block), i, ctypes.POINTER(vp)]
    L.lee_get_solve_time.argtypes = [vp, i, ip, dp, dp]
    L.lee_set_batch_launch.argtypes = [vp, i, i, i]
    L.lee_get_batch_launch.argtypes = [vp, ip, ip, ip]
    L.lee_set_source_field.argtypes = [vp, i, dp]
    L.lee_set_observer.argtypes = [vp, i, ip, ip, dp]
    L.lee_objective_forward.argtypes = [vp, i, i, i, dp]
    L.lee_objective_backward.argtypes = [vp, d]
    L.lee_get_source_gradient.argtypes = [vp, i, dp]
    L.lee_get_state_gradient.argtypes = [vp, i, dp]
    L.lee_compute_rhs_adjoint.argtypes = [vp, i, dp, dp]
    L.lee_apply_filter_adjoint.argtypes = [vp, i, dp]
    L.lee_sample_observers.argtypes = [vp, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def parse_bc_line(tokens, dims):
    """One line of the BC file (src/io/io.cpp:1158-1268): `iblock type +dir
    is ie js je ks ke`; '+' is the low face, '-' the high one; an index
    written '-k' means n-1-k.  dims = (ni, nj, nk) of the block.
    Returns (iblock, type code, dir, end, is[3], ie[3])."""
    if len(tokens) != 9:
        raise LeeError("Number of entries in the boundary condition file is incorrect.")
    blk = int(tokens[0])
    name = str(tokens[1]).lower()
    if name not in BC_TYPES:
        raise LeeError(f"unsupported boundary type {tokens[1]}")
    side = str(tokens[2])
    if side[0] not in "+-":
        raise LeeError("Unknown side for a boundary in the boundary condition file.")
    end = 0 if side[0] == "+" else 1
    dname = side[1:].lower()
    if dname not in ("xi", "eta", "zeta"):
        raise LeeError("Unknown direction for a boundary in the boundary condition file.")
    direction = ("xi", "eta", "zeta").index(dname)
    vals = []
    for t, tok in enumerate(tokens[3:9]):
        s = str(tok)
        v = int(s[1:]) if s[0] in "+-" else int(s)
        if s[0] == "-":
            v = (dims[t // 2] - 1) - v
        vals.append(v)
    return blk, BC_TYPES[name], direction, end, [vals[0], vals[2], vals[4]], [vals[1], vals[3], vals[5]]


def read_bc_file(path, dims_per_block):
    """Reads a BC file: skips empty / whitespace / '#' lines (io.cpp:522-563)."""
    out = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            toks = line.split()
            blk = int(toks[0])
            out.append(parse_bc_line(toks, dims_per_block[blk]))
    return out


def _make_config(*, gamma=1.4, fd_order=4, filter_order=8, filter_blend=1.0, filter_strength=1.0, rk=4,
                 mode="FIX_CFL", value=0.5,
                 sponge_order=2, sponge_constant=5.0, simulation="CASE_TANNA_TPN49", source_params=(),
                 precision=FP64, path=PATH_FUSED, num_dim=2, axisym=0, harmonic=None, num_scalar=0):
    """lee_config from the deck settings (src/core/input.cpp:99-284 keys).
    harmonic: None or dict(wave, shape, direction, amplitude, wavelength[,
    halfwidth]) with the HARMONIC_WAVE names (e.g. wave="WAVE_ENTROPY",
    shape="WAVEFORM_HOMOGENEOUS", direction="X").  num_scalar: the passive
    scalars of PHYSICAL_MODEL = LEE_SCALAR (0: LEE)."""
    cfg = _Config()
    cfg.num_scalar = int(num_scalar)
    if harmonic:
        cfg.harmonic_wave = WAVES[harmonic["wave"]]
        cfg.harmonic_shape = WAVEFORMS.get(harmonic.get("shape", ""), 0)
        cfg.harmonic_direction = "XYZ".index(str(harmonic.get("direction", "X")).upper())
        cfg.harmonic_amplitude = float(harmonic.get("amplitude", 0.0))
        cfg.harmonic_wavelength = float(harmonic.get("wavelength", 1.0))
        cfg.harmonic_halfwidth = float(harmonic.get("halfwidth", 1.0))
    cfg.num_dim = num_dim
    cfg.gamma = gamma
    cfg.fd_order = fd_order
    cfg.filter_order = filter_order
    cfg.filter_blend = filter_blend
    cfg.filter_strength = filter_strength
    cfg.axisym = 1 if axisym else 0
    cfg.rk_stages = rk
    cfg.fix_cfl = 1 if mode == "FIX_CFL" else 0
    cfg.cfl_or_dt = value
    cfg.sponge_order = sponge_order
    cfg.sponge_constant = sponge_constant
    cfg.source_kind = SOURCES[simulation]
    for k, v in enumerate(source_params):
        cfg.source_params[k] = v
    cfg.precision = precision
    cfg.path = path
    return cfg


class Solver:
    """One lee_handle: a set of blocks advanced on the current HIP device.

    blocks: list of dict(x, y, [z], mean (nv, nj, ni)) with x, y shaped (nj, ni)
    and nv = 5 + num_scalar (the passive scalars of LEE_SCALAR after p')
    boundaries: list of parsed BC entries (parse_bc_line) or raw BC-file
    token tuples.  overset: dict of receiver/donor/hole tables or None.
    """

    nv = NV  # variables per point; a LEE_SCALAR handle adds its scalars

    def __init__(self, blocks, *, gamma=1.4, fd_order=4, filter_order=8, filter_blend=1.0, rk=4,
                 mode="FIX_CFL", value=0.5, sponge_order=2, sponge_constant=5.0, simulation="CASE_TANNA_TPN49",
                 source_params=(), precision=FP64, path=PATH_FUSED, boundaries=(), overset=None, num_dim=2,
                 filter_strength=1.0, axisym=0, harmonic=None, num_scalar=0):
        L = library()
        self.L = L
        self.num_dim = num_dim
        self.nv = NV + int(num_scalar)
        self.cfg = _make_config(gamma=gamma, fd_order=fd_order, filter_order=filter_order,
                                filter_blend=filter_blend, filter_strength=filter_strength, rk=rk, mode=mode,
                                value=value,
                                sponge_order=sponge_order, sponge_constant=sponge_constant,
                                simulation=simulation, source_params=source_params, precision=precision, path=path,
                                num_dim=num_dim, axisym=axisym, harmonic=harmonic, num_scalar=num_scalar)
        self._keep = []
        barr = self._blocks(blocks, boundaries)
        ovp = None
        if overset is not None:
            ovp = self._overset_struct(overset)
        h = ctypes.c_void_p()
        rc = L.lee_create(ctypes.byref(self.cfg), barr, len(blocks), ovp, ctypes.byref(h))
        if rc != 0:
            raise LeeError(f"lee_create failed ({rc}): {L.lee_last_error(None).decode()}")
        self.h = h

    def _blocks(self, blocks, boundaries):
        """lee_block array of `blocks`; boundaries are BC-file lines of any block."""
        nb = len(blocks)
        barr = (_Block * nb)()
        self.shapes = []
        dims = []
        for b in blocks:
            sh = np.asarray(b["x"]).shape
            nk, nj, ni = (1,) + sh if len(sh) == 2 else sh
            dims.append((ni, nj, nk))
            self.shapes.append(sh)
        parsed = [bc if isinstance(bc[1], int) else parse_bc_line(bc, dims[int(bc[0])]) for bc in boundaries]
        for ib, b in enumerate(blocks):
            x = np.asarray(b["x"], np.float64)
            ni, nj, nk = dims[ib]
            z = np.zeros_like(x) if b.get("z") is None else np.asarray(b["z"], np.float64)
            xyz = np.ascontiguousarray(np.stack([x, np.asarray(b["y"], np.float64), z]))
            mean = np.ascontiguousarray(np.asarray(b["mean"], np.float64).reshape((self.nv,) + x.shape))
            mine = [p for p in parsed if p[0] == ib]
            bar = (_Boundary * max(1, len(mine)))()
            for k, (_, typ, dr, end, is3, ie3) in enumerate(mine):
                bar[k].type, bar[k].dir, bar[k].end = typ, dr, end
                for q in range(3):
                    bar[k].is_[q] = is3[q]
                    bar[k].ie[q] = ie3[q]
            self._keep += [xyz, mean, bar]
            barr[ib].ni, barr[ib].nj, barr[ib].nk = ni, nj, nk
            barr[ib].xyz = _dp(xyz)
            barr[ib].mean = _dp(mean)
            barr[ib].nboundary = len(mine)
            barr[ib].boundary = bar
        self._keep.append(barr)
        return barr

    def _overset_struct(self, ov):
        o = _Overset()
        arrs = {}
        for k in ("recv_block", "recv_l", "donor_block", "donor_l", "widths", "hole_block", "hole_l"):
            arrs[k] = np.ascontiguousarray(np.asarray(ov.get(k, []), np.int32).reshape(-1))
        coef = np.ascontiguousarray(np.asarray(ov["coef"], np.float64))
        self._keep += list(arrs.values()) + [coef]
        o.nrecv = len(arrs["recv_block"])
        o.recv_block, o.recv_l = _ip(arrs["recv_block"]), _ip(arrs["recv_l"])
        o.donor_block, o.donor_l = _ip(arrs["donor_block"]), _ip(arrs["donor_l"])
        o.widths = _ip(arrs["widths"])
        o.coef = _dp(coef)
        o.coef_stride = coef.shape[1]
        o.nhole = len(arrs["hole_block"])
        o.hole_block, o.hole_l = _ip(arrs["hole_block"]), _ip(arrs["hole_l"])
        self._keep.append(o)
        return ctypes.byref(o)

    @classmethod
    def from_case(cls, case, precision=FP64, path=PATH_FUSED):
        value = case["cfl"] if case["mode"] == "FIX_CFL" else case["dt"]
        s = cls([dict(x=case["x"], y=case["y"], mean=case["mean"])], fd_order=case["fd_order"],
                filter_order=case["filter_order"], filter_blend=case["filter_blend"],
                filter_strength=case.get("filter_strength", 1.0), axisym=case.get("axisym", 0), rk=case["rk"],
                mode=case["mode"], value=value, sponge_order=case["sponge_order"],
                sponge_constant=case["sponge_constant"], simulation=case["simulation"],
                source_params=case.get("source_params", ()), precision=precision, path=path,
                boundaries=case["bc"], harmonic=case.get("harmonic"), num_scalar=case.get("num_scalar", 0))
        s.set_solution(case["ic"])
        return s

    def _check(self, rc):
        if rc != 0:
            raise LeeError(f"camille_amd error {rc}: {self.L.lee_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.lee_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def n(self, ib=0):
        return int(np.prod(self.shapes[ib]))

    # state ---------------------------------------------------------------
    def set_solution(self, sol, ib=0):
        a = np.ascontiguousarray(np.asarray(sol, np.float64).reshape(-1))
        if a.size != self.nv * self.n(ib):
            raise LeeError("solution size mismatch")
        self._check(self.L.lee_set_solution(self.h, ib, _dp(a)))

    def solution(self, ib=0):
        out = np.empty((self.nv,) + self.shapes[ib])
        self._check(self.L.lee_get_solution(self.h, ib, _dp(out)))
        return out

    def set_time(self, step, time, dt):
        self._check(self.L.lee_set_time(self.h, int(step), float(time), float(dt)))

    def clock(self):
        st, t, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._check(self.L.lee_get_time(self.h, ctypes.byref(st), ctypes.byref(t), ctypes.byref(dt)))
        return st.value, t.value, dt.value

    @property
    def time(self):
        return self.clock()[1]

    @property
    def dt(self):
        return self.clock()[2]

    @property
    def time_step(self):
        return self.clock()[0]

    # advance -------------------------------------------------------------
    def compute_cfl(self):
        c = ctypes.c_double()
        self._check(self.L.lee_compute_cfl(self.h, ctypes.byref(c)))
        return c.value

    def time_advance(self):
        self._check(self.L.lee_time_advance(self.h))

    def solve(self, nsteps, sync=True):
        self._check(self.L.lee_solve(self.h, int(nsteps)))
        if sync:
            self.synchronize()

    def synchronize(self):
        self._check(self.L.lee_synchronize(self.h))

    def last_solve_ms(self):
        ms = ctypes.c_float()
        self._check(self.L.lee_last_solve_ms(self.h, ctypes.byref(ms)))
        return ms.value

    KERNEL_CLASSES = ("stage_first", "stage_mid", "stage_last", "filter", "boundary")

    def autotune(self):
        self._check(self.L.lee_autotune(self.h))

    def set_tile_shape(self, stage_rows, filter_rows, ib=0):
        self._check(self.L.lee_set_tile_shape(self.h, ib, int(stage_rows), int(filter_rows)))

    def set_filter_grid(self, workgroups, ib=0):
        """0: one filter workgroup per tile; >0: that many persistent
        pipelined workgroups (rounded up to a multiple of 8)."""
        self._check(self.L.lee_set_filter_grid(self.h, ib, int(workgroups)))

    def filter_grid(self, ib=0):
        g = ctypes.c_int()
        self._check(self.L.lee_get_filter_grid(self.h, ib, ctypes.byref(g)))
        return g.value

    def set_filter_march(self, waves, rows, ib=0):
        """Column-march filter forms: waves 1 with rows 8 or 16 (blocks
        without holes), stacks of 4 waves with rows 4, 8 or 16 or of 8 waves
        with rows 8; (0, 0) returns to the tile forms."""
        self._check(self.L.lee_set_filter_march(self.h, ib, int(waves), int(rows)))

    def filter_march(self, ib=0):
        c, r = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.lee_get_filter_march(self.h, ib, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def tile_shape(self, ib=0):
        r, f = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.lee_get_tile_shape(self.h, ib, ctypes.byref(r), ctypes.byref(f)))
        return r.value, f.value

    def set_kernel_timing(self, on):
        self._check(self.L.lee_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        """{class: (total ms, launches)} since the last call."""
        ms = np.zeros(5)
        cnt = np.zeros(5, np.int32)
        self._check(self.L.lee_get_kernel_timing(self.h, _dp(ms), _ip(cnt)))
        return {k: (float(ms[c]), int(cnt[c])) for c, k in enumerate(self.KERNEL_CLASSES)}

    # unit entry points ---------------------------------------------------
    def compute_rhs(self, y, time, ib=0):
        y = np.ascontiguousarray(np.asarray(y, np.float64))
        out = np.empty_like(y)
        self._check(self.L.lee_compute_rhs(self.h, ib, _dp(y), float(time), _dp(out)))
        return out

    def apply_filter(self, y, ib=0):
        y = np.array(y, np.float64, copy=True, order="C")
        self._check(self.L.lee_apply_filter(self.h, ib, _dp(y)))
        return y

    def take_derivative_xyz(self, f, direction, ib=0):
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        out = np.empty_like(f)
        self._check(self.L.lee_take_derivative_xyz(self.h, ib, _dp(f), int(direction), _dp(out)))
        return out

    def update_boundary(self, ys):
        arrs = [np.array(y, np.float64, copy=True, order="C") for y in ys]
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dp(a) for a in arrs])
        self._check(self.L.lee_update_boundary(self.h, ptrs))
        return arrs

    # discrete adjoint ----------------------------------------------------
    def set_source_field(self, q, ib=0):
        """Spatial field of a SOURCE_FIELD block: p-source q(x) A sin(omega t)."""
        q = np.ascontiguousarray(np.asarray(q, np.float64).reshape(-1))
        if q.size != self.n(ib):
            raise LeeError("source field size mismatch")
        self._check(self.L.lee_set_source_field(self.h, ib, _dp(q)))

    def set_observer(self, points):
        """points: iterable of (block, i, j, weight) in 2-D or (block, i, j, k,
        weight) in 3-D; J sums weight * p'^2."""
        pts = list(points)
        blk = np.array([p[0] for p in pts], np.int32)

        def lin(p):
            sh = self.shapes[int(p[0])]
            ni, nj = sh[-1], sh[-2]
            k = int(p[3]) if len(p) == 5 else 0
            return (k * nj + int(p[2])) * ni + int(p[1])
        idx = np.array([lin(p) for p in pts], np.int32)
        w = np.array([p[-1] for p in pts], np.float64)
        self.set_observer_points(blk, idx, w)

    def sample_observers(self):
        """p' at the observer points of set_observer, in its order."""
        out = np.empty(getattr(self, "_nobs", 0))
        self._check(self.L.lee_sample_observers(self.h, _dp(out)))
        return out

    def set_observer_points(self, block, index, weight):
        """Observers as arrays: block, 1-D point index within the block, weight."""
        blk = np.ascontiguousarray(np.asarray(block, np.int32))
        idx = np.ascontiguousarray(np.asarray(index, np.int32))
        w = np.ascontiguousarray(np.asarray(weight, np.float64))
        self._keep += [blk, idx, w]
        self._check(self.L.lee_set_observer(self.h, len(blk), _ip(blk), _ip(idx), _dp(w)))
        self._nobs = len(blk)
        self.observers = (blk, idx, w)
        self.objective_generation = getattr(self, "objective_generation", 0) + 1

    def objective_forward(self, nsteps, window_start, ckpt_every):
        """J of a run of nsteps from the current state and clock.  Each call
        (and each set_observer) advances `objective_generation`: the reverse
        pass always belongs to the latest forward run."""
        J = ctypes.c_double()
        self._check(self.L.lee_objective_forward(self.h, int(nsteps), int(window_start), int(ckpt_every),
                                                 ctypes.byref(J)))
        self.objective_generation = getattr(self, "objective_generation", 0) + 1
        return J.value

    def objective_backward(self, scale=1.0):
        self._check(self.L.lee_objective_backward(self.h, float(scale)))

    def source_gradient(self, ib=0):
        out = np.empty(self.shapes[ib])
        self._check(self.L.lee_get_source_gradient(self.h, ib, _dp(out)))
        return out

    def state_gradient(self, ib=0):
        out = np.empty((NV,) + self.shapes[ib])
        self._check(self.L.lee_get_state_gradient(self.h, ib, _dp(out)))
        return out

    def update_boundary_adjoint(self, ys):
        """B^T of update_boundary on every block (ys: one 5-field array per block)."""
        arrs = [np.array(y, np.float64, copy=True, order="C") for y in ys]
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dp(a) for a in arrs])
        self._check(self.L.lee_update_boundary_adjoint(self.h, ptrs))
        return arrs

    def compute_rhs_adjoint(self, kappa, ib=0):
        k = np.ascontiguousarray(np.asarray(kappa, np.float64))
        out = np.empty_like(k)
        self._check(self.L.lee_compute_rhs_adjoint(self.h, ib, _dp(k), _dp(out)))
        return out

    def apply_filter_adjoint(self, y, ib=0):
        y = np.array(y, np.float64, copy=True, order="C")
        self._check(self.L.lee_apply_filter_adjoint(self.h, ib, _dp(y)))
        return y

    # init products -------------------------------------------------------
    def metrics(self, ib=0):
        n = self.n(ib)
        met, jac = np.empty(9 * n), np.empty(n)
        self._check(self.L.lee_get_metrics(self.h, ib, _dp(met), _dp(jac)))
        return met.reshape((3, 3) + self.shapes[ib]), jac.reshape(self.shapes[ib])

    def iblank(self, ib=0):
        out = np.empty(self.shapes[ib], np.int32)
        self._check(self.L.lee_get_iblank(self.h, ib, _ip(out)))
        return out

    def mean_gradient(self, ib=0):
        out = np.empty((self.nv * getattr(self, "num_dim", 2),) + self.shapes[ib])
        self._check(self.L.lee_get_mean_gradient(self.h, ib, _dp(out)))
        return out

    def sponge(self, ib=0):
        out = np.empty(self.shapes[ib])
        self._check(self.L.lee_get_sponge(self.h, ib, _dp(out)))
        return out


class Batch(Solver):
    """A batched ensemble (lee_batch_create): independent single-block solves
    advanced together, one persistent launch per RK stage over all of them,
    replayed from a HIP graph.

    solves: list of dict(x, y, mean, [boundaries], [simulation],
    [source_params], [mode], [value]) -- grid, base state, BC-file lines,
    source and CFL/dt of each solve.  The scheme settings are shared.
    The per-block methods take the solve index as `ib`.
    """

    def __init__(self, solves, *, gamma=1.4, fd_order=4, filter_order=8, filter_blend=1.0, rk=4,
                 sponge_order=2, sponge_constant=5.0, precision=FP64):
        L = library()
        self.L = L
        self._keep = []
        ns = len(solves)
        cfgs = (_Config * ns)()
        for k, sv in enumerate(solves):
            cfgs[k] = _make_config(gamma=gamma, fd_order=fd_order, filter_order=filter_order,
                                   filter_blend=filter_blend, rk=rk, mode=sv.get("mode", "FIX_CFL"),
                                   value=sv.get("value", 0.5), sponge_order=sponge_order,
                                   sponge_constant=sponge_constant,
                                   simulation=sv.get("simulation", "CASE_TANNA_TPN49"),
                                   source_params=sv.get("source_params", ()), precision=precision, path=PATH_FUSED)
        self.cfg = cfgs[0]
        self._keep.append(cfgs)
        bcs = []
        for k, sv in enumerate(solves):
            for ln in sv.get("boundaries", ()):
                bcs.append((k,) + tuple(ln[1:]))
        barr = self._blocks(solves, bcs)
        h = ctypes.c_void_p()
        rc = L.lee_batch_create(cfgs, barr, ns, ctypes.byref(h))
        if rc != 0:
            raise LeeError(f"lee_batch_create failed ({rc}): {L.lee_last_error(None).decode()}")
        self.h = h

    def clock(self, ib=0):
        st, t, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._check(self.L.lee_get_solve_time(self.h, int(ib), ctypes.byref(st), ctypes.byref(t),
                                              ctypes.byref(dt)))
        return st.value, t.value, dt.value

    def set_launch(self, use_graph=True, grid_stage=0, grid_filter=0):
        """Graph replay on/off and the persistent grids of the stage and
        filter launches: 0 one workgroup per tile, GRID_RESIDENT one
        resident wave of workgroups, else that many workgroups."""
        self._check(self.L.lee_set_batch_launch(self.h, 1 if use_graph else 0, int(grid_stage), int(grid_filter)))

    def launch(self):
        g, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.L.lee_get_batch_launch(self.h, ctypes.byref(g), ctypes.byref(a), ctypes.byref(b)))
        return bool(g.value), a.value, b.value
