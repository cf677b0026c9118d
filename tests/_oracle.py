"""The oracle's ctypes bindings (oracle/oracle_bindings.py), for the test modules."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
from oracle_bindings import *  # noqa: E402,F401,F403
from oracle_bindings import lib, ref, ref_nelmin_ptr  # noqa: E402,F401
