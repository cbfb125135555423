"""Test helpers: build the same problem on two libraries and compare."""
import numpy as np

import fivo2d as F
from fivo2d import ic


def gubser_problem(N=128, L=10.0, tau0=1.0, Tc0=0.6, q=1.0, g_hq=0.0, alpha=-20.0):
    """Config C1 by default: conformal e = 3P (g_hq = 0), no charge, ideal."""
    g = F.make_grid2d(N, N, xmax=L, ymax=L)
    eos = F.ConformalHQEOS(g_hq=g_hq)
    m = F.build_model_2d(eos=eos, with_charge=(g_hq > 0))
    prims, Ts = ic.gubser_ic(g, tau0, q, Tc0, alpha)
    return g, m, prims, Ts


def make_solver(lib, g, m, prims, tau0):
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(tau0, **prims)
    return s


def rel_diff(a, b):
    """max |a-b| / max |b| per field (columns), as an array."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return np.max(np.abs(a - b), axis=0) / scale
