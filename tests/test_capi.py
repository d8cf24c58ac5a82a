"""CPU checks of the C-ABI library and its host-side logic (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import camille_amd
from camille_amd import lee

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = "".join(open(os.path.join(ROOT, "include", f)).read() for f in sorted(os.listdir(os.path.join(ROOT, "include")))
                  if f.endswith(".h"))
    return sorted(set(re.findall(r"\b(lee_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports_every_declared_symbol():
    L = camille_amd.library()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(lee.EXPORTS) <= set(names)
    assert L.lee_version().startswith(b"camille_amd")


def _macros():
    txt = open(os.path.join(ROOT, "include", "lee.h")).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(LEE_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", txt)}


def test_binding_constants_match_the_header():
    """The ctypes binding's precision, path, launch-grid and BC-type codes
    are the header's macros."""
    m = _macros()
    assert (lee.FP64, lee.FP32) == (m["LEE_FP64"], m["LEE_FP32"])
    assert (lee.PATH_FUSED, lee.PATH_PLAIN) == (m["LEE_PATH_FUSED"], m["LEE_PATH_PLAIN"])
    assert lee.GRID_RESIDENT == m["LEE_GRID_RESIDENT"] == camille_amd.GRID_RESIDENT
    for name, code in lee.BC_TYPES.items():
        key = "LEE_BC_" + name.upper() if not name.startswith("sponge") else "LEE_" + name.upper()
        assert m[key] == code, name
    for name, code in lee.WAVES.items():
        assert m["LEE_WAVE_" + name.replace("WAVE_", "")] == code, name
    for name, code in lee.WAVEFORMS.items():
        assert m["LEE_" + name] == code, name


def _cfg(**kw):
    c = lee._Config()
    c.num_dim = 2
    c.gamma = 1.4
    c.fd_order = 4
    c.filter_order = 8
    c.filter_blend = 1.0
    c.rk_stages = 4
    c.fix_cfl = 1
    c.cfl_or_dt = 0.5
    c.sponge_order = 2
    c.sponge_constant = 5.0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _create(cfg, ni=16, nj=16, boundary=None):
    """Calls lee_create; every case here fails validation before any HIP call."""
    L = camille_amd.library()
    xyz = np.zeros((3, nj, ni))
    mean = np.zeros((5, nj, ni))
    blk = (lee._Block * 1)()
    blk[0].ni, blk[0].nj, blk[0].nk = ni, nj, 1
    blk[0].xyz = xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    blk[0].mean = mean.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if boundary is not None:
        arr = (lee._Boundary * 1)(boundary)
        blk[0].nboundary = 1
        blk[0].boundary = arr
    h = ctypes.c_void_p()
    rc = L.lee_create(ctypes.byref(cfg), blk, 1, None, ctypes.byref(h))
    return rc, L.lee_last_error(None).decode(), h


@pytest.mark.parametrize("field,value,msg", [
    ("num_dim", 4, "DIMENSION must be 2 or 3"),
    ("fd_order", 2, "ORDER_ACCURACY 4 or FDO11P"),
    ("filter_order", 6, "filter order"),
    ("filter_blend", 1.5, "out-of-bound"),
    ("rk_stages", 3, "temporal scheme"),
    ("precision", 7, "precision"),
    ("source_kind", 9, "source"),
])
def test_create_rejects_bad_config(field, value, msg):
    rc, err, h = _create(_cfg(**{field: value}))
    assert rc == 1 and msg in err and not h.value


def test_create_rejects_plain_fp32_and_small_blocks():
    rc, err, _ = _create(_cfg(precision=1, path=1))
    assert rc == 1 and "FP64 only" in err
    rc, err, _ = _create(_cfg(fd_order=11, num_dim=3))
    assert rc == 1 and "run in 2-D" in err
    rc, err, _ = _create(_cfg(filter_order=11, filter_strength=1.5, path=1))
    assert rc == 1 and "STRENGTH" in err
    rc, err, _ = _create(_cfg(fd_order=11, path=1), ni=10, nj=20)
    assert rc == 1 and "too small" in err
    rc, err, _ = _create(_cfg(num_dim=3, axisym=1))
    assert rc == 1 and "requires 2-D" in err
    rc, err, _ = _create(_cfg(num_dim=3, path=1))
    assert rc == 1 and "fused path" in err
    rc, err, _ = _create(_cfg(), ni=6, nj=20)
    assert rc == 1 and "too small" in err
    rc, err, _ = _create(_cfg(num_scalar=5))
    assert rc == 1 and "NUMBER_SCALAR" in err
    rc, err, _ = _create(_cfg(num_scalar=1, num_dim=3))
    assert rc == 1 and "LEE_SCALAR" in err and "2-D" in err


def test_create_rejects_bad_boundaries():
    b = lee._Boundary()
    b.type, b.dir, b.end = 1000, 0, 0
    for q in range(3):
        b.is_[q], b.ie[q] = 0, 0
    b.ie[0], b.ie[1] = 30, 15  # beyond ni = 16
    rc, err, _ = _create(_cfg(), boundary=b)
    assert rc == 1 and "outside" in err
    b.ie[0] = 5
    for t in (0, 1001):  # Dirichlet face and sponge: the reference exits for LEE (state.cpp:746-749)
        b.type = t
        rc, err, _ = _create(_cfg(), boundary=b)
        assert rc == 1 and "not implemented yet" in err
        rc, err, _ = _create(_cfg(num_scalar=1), boundary=b)  # LEE_SCALAR's own message (state.cpp:751-754)
        assert rc == 1 and "with a scalar is not implemented yet" in err
    b.type = 7
    rc, err, _ = _create(_cfg(), boundary=b)
    assert rc == 1 and "unsupported boundary" in err


def test_create_rejects_harmonic_waves_like_the_reference():
    """The reference's deck check (input.cpp:630-655) ends every acoustic
    wave in "Unknown wave." (its else binds to the inner if) and checks an
    entropy wave's amplitude; bc_dirichlet_harmonicwave refuses an entropy
    plane wave and an unset shape."""
    b = lee._Boundary()
    b.type, b.dir, b.end = 2, 0, 0
    for q in range(3):
        b.is_[q], b.ie[q] = 0, 0
    b.ie[1] = 15
    cases_ = [
        (dict(harmonic_wave=1, harmonic_shape=1, harmonic_wavelength=4.0), "Unknown wave."),
        (dict(harmonic_wave=1, harmonic_shape=3, harmonic_wavelength=4.0), "Unknown spatial distribution"),
        (dict(harmonic_wave=2, harmonic_shape=2, harmonic_amplitude=1.5, harmonic_wavelength=4.0),
         "between 0 and 1"),
        (dict(harmonic_wave=2, harmonic_shape=1, harmonic_amplitude=0.1, harmonic_wavelength=4.0),
         "Unknown type of wave for this wave form."),
        (dict(harmonic_wave=0), "Unknown type of wave form."),
    ]
    for kw, msg in cases_:
        rc, err, _ = _create(_cfg(**kw), boundary=b)
        assert rc == 1 and msg in err, (kw, err)


def test_bc_line_parsing_matches_reference_convention():
    dims = (201, 101, 1)
    blk, typ, d, end, is3, ie3 = camille_amd.parse_bc_line(
        ("0", "sponge_freund_ambient", "-xi", "-19", "-0", "0", "-0", "0", "0"), dims)
    assert (blk, typ, d, end) == (0, 1000, 0, 1)
    assert is3 == [181, 0, 0] and ie3 == [200, 100, 0]
    _, typ, d, end, is3, ie3 = camille_amd.parse_bc_line(
        ("0", "WALL_SLIP_KINEMATIC", "+eta", "0", "-0", "0", "0", "0", "0"), dims)
    assert (typ, d, end) == (201, 1, 0) and is3 == [0, 0, 0] and ie3 == [200, 0, 0]
    with pytest.raises(camille_amd.LeeError):
        camille_amd.parse_bc_line(("0", "sponge_freund_ambient", "*xi", "0", "1", "0", "0", "0", "0"), dims)
    with pytest.raises(camille_amd.LeeError):
        camille_amd.parse_bc_line(("0", "sponge_freund_ambient", "+xi", "0", "1"), dims)


def test_bc_file_reader_skips_comments(tmp_path):
    p = tmp_path / "bc.in"
    p.write_text("# header\n\n   \n0 sponge_freund_ambient +xi 0 19 0 -0 0 0\n")
    out = camille_amd.read_bc_file(str(p), [(201, 201, 1)])
    assert len(out) == 1 and out[0][4] == [0, 0, 0] and out[0][5] == [19, 200, 0]
