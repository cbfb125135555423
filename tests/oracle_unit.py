"""ctypes bindings of the oracle's single-cell exports (test infrastructure)."""
import ctypes as C

import numpy as np

import fivo2d as F
from fivo2d import _abi

_DP = C.POINTER(C.c_double)


class OracleUnit:
    def __init__(self, lib):
        d = lib.dll
        self.d = d
        d.orc_besselkx.restype = C.c_double
        d.orc_besselkx.argtypes = [C.c_int, C.c_double]
        d.orc_besselk2x_fast.restype = C.c_double
        d.orc_besselk2x_fast.argtypes = [C.c_double]
        for n in ("orc_eos_Pne", "orc_eos_Pne_2d", "orc_light"):
            getattr(d, n).restype = None
            getattr(d, n).argtypes = [C.POINTER(_abi.EosDesc), C.c_double] + \
                ([C.c_double] if n != "orc_light" else []) + [_DP]
        d.orc_eos_cs2.restype = C.c_double
        d.orc_eos_cs2.argtypes = [C.POINTER(_abi.EosDesc), C.c_double]
        d.orc_prim_to_cons.restype = C.c_int
        d.orc_prim_to_cons.argtypes = [C.POINTER(_abi.EosDesc), C.c_double, C.c_double, C.c_double,
                                       C.c_double, _DP, C.c_double, _DP]
        d.orc_cons_to_prim.restype = None
        d.orc_cons_to_prim.argtypes = [C.POINTER(_abi.EosDesc), _DP, _DP, C.c_double, _DP, _DP]

    def besselkx(self, nu, x):
        return self.d.orc_besselkx(nu, x)

    def besselk2x_fast(self, x):
        return self.d.orc_besselk2x_fast(x)

    def eos_Pne(self, eos, T, mu, two_d=False):
        out = (C.c_double * 3)()
        e = eos.desc()
        (self.d.orc_eos_Pne_2d if two_d else self.d.orc_eos_Pne)(C.byref(e), T, mu, out)
        return tuple(out)

    def light(self, eos, T):
        out = (C.c_double * 3)()
        e = eos.desc()
        self.d.orc_light(C.byref(e), T, out)
        return tuple(out)

    def cs2(self, eos, T):
        e = eos.desc()
        return self.d.orc_eos_cs2(C.byref(e), T)

    def prim_to_cons(self, eos, T, mu, ux, uy, diss7, tau):
        e = eos.desc()
        d = (C.c_double * 7)(*diss7)
        out = (C.c_double * 4)()
        ok = self.d.orc_prim_to_cons(C.byref(e), T, mu, ux, uy, d, tau, out)
        return bool(ok), tuple(out)

    def cons_to_prim(self, eos, D, Sx, Sy, E, diss7, tau, warm=(np.nan,) * 4):
        """D is J^tau (the conserved tau*J^tau divided by tau), as cons_to_prim_2d! takes it."""
        e = eos.desc()
        c = (C.c_double * 4)(D, Sx, Sy, E)
        d = (C.c_double * 7)(*diss7)
        w = (C.c_double * 4)(*warm)
        out = (C.c_double * 11)()
        self.d.orc_cons_to_prim(C.byref(e), c, d, tau, w, out)
        keys = ("T", "mu", "ux", "uy", "n", "e", "P", "ok", "reason", "iters", "resnorm")
        return dict(zip(keys, out))
