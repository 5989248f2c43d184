"""fricc -- MI355X-native FRI-CCSD hot path.

Drop-in for jamesETsmith/fri-cc's ``fricc`` package on the path it
accelerates: ``fricc.FRICCSD`` / ``update_amps`` (fricc/friccsd.py),
``fricc.FRILCCSD`` (fricc/frilccsd.py) and the
``fricc.py_rccsd`` binding (src/py_rccsd.cpp), backed by hand-written gfx950
HIP kernels in ``libfricc_hip.so`` (C ABI: include/fricc_capi.h).
"""
from .friccsd import FRICCSD, default_fri_settings, kernel, update_amps  # noqa: F401
from .frilccsd import FRILCCSD  # noqa: F401
from . import py_rccsd  # noqa: F401

__all__ = ["FRICCSD", "FRILCCSD", "default_fri_settings", "kernel", "update_amps", "py_rccsd"]
