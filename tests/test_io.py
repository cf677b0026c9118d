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
