"""ESRI ASCII grids (include/photic_io.h), the reference's real-scene grid format
(model/io.c:143-269): the header, the north-first row order against the south-first grid,
the value format, and errors returned instead of exit()."""
import ctypes as C
import os

import numpy as np
import pytest

import photic_amd as P
from photic_amd.samodel_types import RowGrid

L = C.CDLL(P.LIB_PATH)
FP = C.POINTER(C.POINTER(C.c_float))
L.photic_read_asc.argtypes = [C.c_char_p, C.POINTER(FP), C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
L.photic_write_asc.argtypes = [C.c_char_p, FP, C.c_int, C.c_int, C.c_double, C.c_double,
                               C.c_double, C.c_double]
L.photic_free_grid.argtypes = [FP, C.c_int]

# A 3 x 4 grid as a file lists it: the first data line is the NORTHERN row.
TEXT = """ncols        4
nrows        3
xllcorner    115.250000
yllcorner    -21.500000
cellsize     0.000250
NODATA_value -9999.000000
1.5 2.5 3.5 4.5
5 6 7 -9999
-0.25 0.125 9 10
"""


def _read(path):
    g = FP()
    nc, nr = C.c_int(), C.c_int()
    x, y, cs, nd = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    rc = L.photic_read_asc(path.encode(), C.byref(g), C.byref(nc), C.byref(nr), C.byref(x),
                           C.byref(y), C.byref(cs), C.byref(nd))
    if rc:
        return rc, None, None
    a = np.array([[g[i][j] for j in range(nc.value)] for i in range(nr.value)], np.float32)
    L.photic_free_grid(g, nr.value)
    return rc, a, (x.value, y.value, cs.value, nd.value)


def test_read_flips_rows_south_first(tmp_path):
    p = str(tmp_path / "g.asc")
    open(p, "w").write(TEXT)
    rc, a, hdr = _read(p)
    assert rc == 0
    # row 0 is the southern (last-listed) row, model/io.c:224-232
    assert np.array_equal(a, np.array([[-0.25, 0.125, 9, 10], [5, 6, 7, -9999],
                                       [1.5, 2.5, 3.5, 4.5]], np.float32))
    assert hdr == (115.25, -21.5, 0.00025, -9999.0)


def test_write_layout_and_round_trip(tmp_path):
    g = np.arange(12, dtype=np.float32).reshape(3, 4) * 1.25 - 2.0
    rg = RowGrid(g)
    p = str(tmp_path / "w.asc")
    assert L.photic_write_asc(p.encode(), C.cast(rg.pp, FP), 4, 3, 115.25, -21.5, 0.00025,
                              -9999.0) == 0
    lines = open(p).read().split("\n")
    assert lines[0] == "ncols        4" and lines[1] == "nrows        3"
    assert lines[2] == "xllcorner    115.250000" and lines[5] == "NODATA_value -9999.000000"
    # north row first, "%.2f " each (model/io.c:258-264)
    assert lines[6] == "".join("%.2f " % v for v in g[2])
    assert lines[8] == "".join("%.2f " % v for v in g[0])
    rc, a, _ = _read(p)
    assert rc == 0 and np.array_equal(a, g)


@pytest.mark.parametrize("text", [TEXT.replace("nrows", "rows"),
                                  TEXT.replace("NODATA_value", "nodata"),
                                  TEXT.rsplit("\n", 2)[0] + "\n"])
def test_malformed_is_an_error_not_an_exit(tmp_path, text):
    p = str(tmp_path / "bad.asc")
    open(p, "w").write(text)
    rc, _, _ = _read(p)
    assert rc == P.EINVAL
    assert b"photic_read_asc" in P.lib.photic_last_error()


# ---- the int16 packing of the netCDF output grids (model/nc.c:244-300) ------------------------

L.photic_pack_int16.argtypes = [C.c_void_p, C.c_long, C.c_double, C.c_void_p,
                                C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_short)]
L.photic_unpack_int16.argtypes = [C.c_void_p, C.c_long, C.c_float, C.c_float, C.c_short,
                                  C.c_double, C.c_void_p]


def _pack(v, spval=-9999.0):
    v = np.ascontiguousarray(v, np.float32)
    c = np.zeros(v.size, np.int16)
    off, sc, miss = C.c_float(), C.c_float(), C.c_short()
    assert L.photic_pack_int16(v.ctypes.data, v.size, spval, c.ctypes.data, C.byref(off),
                               C.byref(sc), C.byref(miss)) == 0
    return c, off.value, sc.value, miss.value


def _unpack(c, off, sc, miss, spval=-9999.0):
    v = np.zeros(c.size, np.float32)
    assert L.photic_unpack_int16(np.ascontiguousarray(c, np.int16).ctypes.data, c.size, off, sc,
                                 miss, spval, v.ctypes.data) == 0
    return v


def test_pack_round_trip_within_half_a_step():
    """Every value comes back within half a quantisation step (plus float rounding), missing
    values as SHRT_MIN and back as the missing value, the smallest value exactly; the
    quantised values span 0..32767."""
    rng = np.random.default_rng(1)
    v = rng.uniform(0.5, 40.0, 5000).astype(np.float32)
    v[0] = 0.2                        # the smallest, not the largest, first
    v[rng.integers(0, v.size, 300)] = -9999.0
    c, off, sc, miss = _pack(v)
    assert miss == -32768 and off == np.float32(0.2)
    keep = v != np.float32(-9999.0)
    assert np.all(c[~keep] == -32768) and c[keep].min() == 0 and c[keep].max() == 32767
    back = _unpack(c, off, sc, miss)
    assert np.all(back[~keep] == np.float32(-9999.0))
    assert np.all(np.abs(back[keep] - v[keep]) <= 0.5 * sc + 4e-6 * np.abs(v[keep]))
    assert back[0] == np.float32(0.2)


def test_pack_scale_is_range_over_32767_in_float():
    v = np.array([1.0, 3.0, 2.0, 5.5], np.float32)
    c, off, sc, miss = _pack(v)
    assert off == 1.0
    assert sc == np.float32(np.float32(4.5) / np.float32(32767.0))
    assert c.tolist() == [0, int(np.rint(np.float32(2.0) / np.float32(sc))),
                          int(np.rint(np.float32(1.0) / np.float32(sc))), 32767]


def test_pack_keeps_the_reference_scan():
    """The reference's scan starts the largest value at FLT_MIN and tests a value for the
    largest only when it is not a new smallest.  So a grid of negative depths keeps FLT_MIN
    as its largest (they still round-trip), and a first value that is the grid's maximum is
    missed: its quotient leaves the short range and keeps the low 16 bits of its 32-bit
    conversion, as the reference's build stores it."""
    neg = np.array([-3.0, -12.5, -0.25, -7.0], np.float32)
    c, off, sc, _ = _pack(neg)
    assert off == np.float32(-12.5)
    assert sc == np.float32((np.float32(np.finfo(np.float32).tiny) - np.float32(-12.5)) /
                            np.float32(32767.0))
    assert np.allclose(_unpack(c, off, sc, -32768), neg, atol=0.5 * sc + 1e-6)
    c, off, sc, _ = _pack(np.array([10.0, 1.0, 2.0, 3.0], np.float32))
    assert off == 1.0 and sc == np.float32(np.float32(2.0) / np.float32(32767.0))
    q = int(np.rint(np.float32(9.0) / np.float32(sc)))     # 147452: out of short range
    assert q > 32767 and c[0] == np.int16(np.uint16(q & 0xFFFF))
    c, off, sc, _ = _pack(np.array([5.0, 5.0, 5.0], np.float32))   # flat: 0 / 0
    assert sc == 0.0 and np.all(c == 0)
    assert np.all(_unpack(c, off, sc, -32768) == 5.0)
