"""MOOD's mask dilation and a run that retries stages (src2d/timestepper2d.jl:113-140,
src2d/floors2d.jl:350-383).

expand_bad_2d! grows the inadmissible-cell mask by the four neighbours of every
marked cell, `radius` times (radius 1 + attempt), updating interior cells only:
checked against a brute-force numpy dilation on random masks and against its
closed form for one marked cell (the L1 diamond |dx| + |dy| <= radius, clipped to
the interior). The run: a sharp lumpy 64^2 event (sigma 0.25 fm, peak T 0.5 GeV) at
CFL 0.6 with SSPRK2 makes inadmissible stage outputs, so stages are retried with
first-order faces on the dilated mask; the GPU test requires the same retries and
the same final state as the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic


def _expand(oracle, g, mask, radius):
    d = oracle.dll
    d.orc_debug_expand.restype = None
    d.orc_debug_expand.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_int]
    s = F.Solver(g, F.build_model_2d(eos=F.LatticeHRGEOS()), lib=oracle)
    a = np.array(mask, dtype=np.uint8).ravel()          # a copy: the probe works in place
    d.orc_debug_expand(s.ctx, a.ctypes.data_as(C.POINTER(C.c_uint8)), int(radius))
    s.close()
    return a.reshape(g.Nxtot, g.Nytot)


def _numpy_expand(g, mask, radius):
    b = mask.astype(bool)
    inner = np.zeros_like(b)
    inner[g.nghost:g.nghost + g.Nx, g.nghost:g.nghost + g.Ny] = True
    for _ in range(radius):
        nb = np.zeros_like(b)
        nb[1:, :] |= b[:-1, :]
        nb[:-1, :] |= b[1:, :]
        nb[:, 1:] |= b[:, :-1]
        nb[:, :-1] |= b[:, 1:]
        b = b | (inner & nb)
    return b.astype(np.uint8)


def test_expand_matches_brute_force(oracle):
    g = F.make_grid2d(20, 14, xmax=5.0, ymax=3.5)
    rng = np.random.default_rng(5)
    for radius in (1, 2, 3):
        for p in (0.02, 0.1):
            m = np.zeros((g.Nxtot, g.Nytot), np.uint8)
            inner = rng.random((g.Nx, g.Ny)) < p
            m[g.nghost:g.nghost + g.Nx, g.nghost:g.nghost + g.Ny] = inner
            assert np.array_equal(_expand(oracle, g, m, radius), _numpy_expand(g, m, radius)), (radius, p)


def test_expand_of_one_cell_is_the_l1_diamond(oracle):
    g = F.make_grid2d(16, 16, xmax=4.0, ymax=4.0)
    for (cx, cy) in ((8, 8), (3, 4), (17, 18)):           # centre, near a corner, the last interior cell
        for radius in (1, 2, 3):
            m = np.zeros((g.Nxtot, g.Nytot), np.uint8)
            m[cx, cy] = 1
            out = _expand(oracle, g, m, radius)
            X, Y = np.meshgrid(np.arange(g.Nxtot), np.arange(g.Nytot), indexing="ij")
            inner = (X >= g.nghost) & (X < g.nghost + g.Nx) & (Y >= g.nghost) & (Y < g.nghost + g.Ny)
            want = ((np.abs(X - cx) + np.abs(Y - cy) <= radius) & inner) | (m == 1)
            assert np.array_equal(out.astype(bool), want), (cx, cy, radius)


def mood_run(lib):
    g = F.make_grid2d(64, 64, xmax=8.0, ymax=8.0)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                         enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                         kappa_coeff=0.1163, pi_clip_factor=1.0)
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(0.6, **ic.hotspot_event(g, seed=3, nspots=60, sigma=0.25, Tpeak=0.5))
    r = s.run_sim_2d(tau0=0.6, taufinal=1.6, CFL=0.6, CFLtau=0.05, integrator="ssprk2")
    f = s.fields_2d()
    s.close()
    return r, f


def test_mood_retries_on_the_oracle(oracle):
    r, f = mood_run(oracle)
    print(f"steps {r['nsteps']} MOOD retries {r['nmood']} halvings {r['nhalvings']} failures {r['nprimfail']}")
    assert r["ok"] and r["nmood"] >= 1 and np.all(np.isfinite(f["T"]))
