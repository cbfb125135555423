"""Fixed-state evaluation of the relaxation's source terms over the C ABI.

The reference's gates for the charm first and second moments (the 1-D limit,
rotation covariance, tracelessness, the Bjorken limit; test/test_consistent_fm2d.jl,
test/test_consistent_m22d.jl) call its source functions at chosen states. These
wrappers hand such states, in the reference functions' own argument order, to
a library's fivo_probe_* / orc_probe_* entry points, which run the same code the
relaxation runs.
"""
import ctypes as C

import numpy as np

from . import _abi

FM_ARGS = ("ux", "uy", "ut", "T", "dxT", "dyT", "dtT", "nux", "nuy", "n", "dn_dT", "tau_n", "Ds", "h", "hp",
           "theta", "ax", "ay", "dxux", "dxuy", "dyux", "dyuy")
M2_ARGS = ("ux", "uy", "ut", "tau", "T", "dxT", "dyT", "dtT", "nux", "nuy", "pxx", "pxy", "pyy", "peta", "PiQ",
           "dta", "dxa", "dya", "theta_nu", "anux", "anuy", "dtnux", "dtnuy", "dxnux", "dxnuy", "dynux", "dynuy",
           "theta", "ax", "ay", "dtux", "dtuy", "dxux", "dxuy", "dyux", "dyuy", "n", "tau_n", "Ds", "h", "hp",
           "tauM", "etaM", "m")
_DP = C.POINTER(C.c_double)


def _rows(rows, names):
    """rows: an (n, k) array in argument order, or a list of dicts keyed by name."""
    if isinstance(rows, dict):
        rows = [rows]
    if len(rows) and isinstance(rows[0], dict):
        rows = [[float(r[k]) for k in names] for r in rows]
    a = np.ascontiguousarray(rows, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != len(names):
        raise ValueError(f"expected rows of {len(names)} arguments ({', '.join(names)})")
    return a


def fm_source(lib, rows, terms=None):
    """consistent_fm_source_2d (src2d/hq_consistent_firstmoment2d.jl:218-282) on each
    row -> (n, 2) array of (s^x, s^y)."""
    a = _rows(rows, FM_ARGS)
    out = np.zeros((a.shape[0], 2))
    t = _abi.TERMS_DEFAULT if terms is None else int(terms)
    rc = lib.probe_fm_source(t, a.shape[0], a.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    if rc != 0:
        raise RuntimeError(f"probe_fm_source failed ({rc}): {lib.last_error(None)}")
    return out


def m2_source(lib, rows, terms=None):
    """consistent_m2_source_2d (src2d/hq_consistent_m2_2d.jl:247-427) on each row ->
    (n, 4) array of the rates of (pixx_Q, pixy_Q, piyy_Q, Pi_Q)."""
    a = _rows(rows, M2_ARGS)
    out = np.zeros((a.shape[0], 4))
    t = _abi.TERMS_DEFAULT if terms is None else int(terms)
    rc = lib.probe_m2_source(t, a.shape[0], a.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    if rc != 0:
        raise RuntimeError(f"probe_m2_source failed ({rc}): {lib.last_error(None)}")
    return out
