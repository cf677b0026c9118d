"""The C-ABI library: it loads without a GPU, exports every function include/*.h declares, and
carries nothing of the oracle."""
import ctypes as C
import os
import re
import subprocess

import photic_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b(\w+)\s*\(", text, flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "#")):
                names.add(m.group(1))
    return names


def test_every_declared_function_is_exported():
    names = _declared()
    assert {"samodel", "photic_samodel", "photic_invert", "photic_invert_h", "photic_forward",
            "photic_objective", "photic_mlp_train", "photic_band_coefficients",
            "photic_invert_regions", "photic_read_asc", "photic_write_asc"} <= names
    lib = C.CDLL(P.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_library_holds_no_oracle_code():
    out = subprocess.run(["nm", "-D", "--defined-only", P.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    assert " orc_" not in out
    deps = subprocess.run(["ldd", P.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in deps


def test_entry_points_fail_cleanly_without_a_gpu():
    """With no GPU every computing entry point returns PHOTIC_EDEVICE (no CPU fallback); on a
    machine with a GPU this only checks that the status is a defined code."""
    cfg = P.cfg_default()
    rc = P.lib.photic_invert(C.byref(cfg), 0, None, None, None, None, None, None)
    assert rc in (P.OK, P.EINVAL, P.EDEVICE)
    if P.device_count() == 0:
        out = P.Outputs()
        rc = P.lib.photic_invert(C.byref(cfg), 16, C.c_void_p(16), None, None, None,
                                 C.byref(out), None)
        assert rc == P.EDEVICE
        assert b"GPU" in P.lib.photic_last_error()


def test_structure_layouts_match_the_reference_header():
    """include/photic_samodel.h's geogrid and scene mirrors have model/common.h's size and
    field offsets (static assertions compiled against the reference's own header by
    oracle/Makefile's layout target); skipped where the reference tree is not mounted."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "layout"],
                       capture_output=True, text=True)
    if "not present" in r.stdout + r.stderr:
        import pytest
        pytest.skip("reference tree not mounted")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "layouts match" in r.stdout
