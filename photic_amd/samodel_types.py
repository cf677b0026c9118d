"""ctypes mirrors of the drop-in's structures (include/photic_samodel.h, which carry the
layouts of the reference's geogrid and scene, model/common.h:69-218) and a helper that builds
row-pointer grids, for Python callers of samodel() / photic_samodel()."""
import ctypes as C

import numpy as np

MAX_GRIDS = 265
MAX_STRING_LEN = 2048


class Geogrid(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("cellsize", C.c_float),
                ("wlon", C.c_float), ("slat", C.c_float), ("elon", C.c_float),
                ("nlat", C.c_float), ("nodata_value", C.c_float), ("lambda_", C.c_float),
                ("theta_v", C.c_float), ("theta_w", C.c_float),
                ("array", C.POINTER(C.POINTER(C.c_float)))]


class Scene(C.Structure):
    _fields_ = [("scene_name", C.c_char * MAX_STRING_LEN), ("n_bands", C.c_int),
                ("nrows", C.c_int), ("ncols", C.c_int),
                ("band_indexes", C.c_int * MAX_GRIDS), ("wavelengths", C.c_int * MAX_GRIDS),
                ("theta_w", C.c_double), ("theta_v", C.c_double), ("H_tide", C.c_double),
                ("R_inf", C.c_double * MAX_GRIDS), ("R_sigma", C.c_double * MAX_GRIDS),
                ("K", C.c_double * MAX_GRIDS), ("K_sigma", C.c_double * MAX_GRIDS),
                ("ratio_min", C.c_double), ("ratio_max", C.c_double),
                ("slope_min", C.c_double), ("slope_max", C.c_double),
                ("pgx_present", C.c_int), ("ps_p", C.c_int), ("ps_g", C.c_int),
                ("ps_x", C.c_int)]


class RowGrid:
    """A float grid held as separately allocated rows behind a float** (the reference's
    allocate_float_array_2d layout)."""

    def __init__(self, a):
        a = np.asarray(a, np.float32)
        self.rows = [np.ascontiguousarray(r) for r in a]
        self.ptrs = (C.POINTER(C.c_float) * len(self.rows))(
            *[r.ctypes.data_as(C.POINTER(C.c_float)) for r in self.rows])

    @property
    def pp(self):
        return C.cast(self.ptrs, C.POINTER(C.POINTER(C.c_float)))

    def array(self):
        return np.stack(self.rows)
