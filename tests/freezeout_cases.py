"""Cases for the freeze-out tracker and its Cooper-Frye integral
(freezeout2d.jl:68-192), built the same way for any library with the C ABI (the
oracle on the CPU, the HIP library on the GPU), and their closed forms.

Test infrastructure shared by test_oracle_freezeout.py and test_gpu_freezeout.py.

  crossing_histories   prescribed temperature histories through the tracker:
                       linear cooling (the interpolated crossing is exact),
                       never hot, always hot, re-heating, cold-hot-cold, a step
  static_yield         closed form of a uniform surface at rest: the eta integral
                       is 2 m_T K1(m_T/T) (Boltzmann) and the Bose distribution
                       its series over K1(n m_T/T)
  tilted_flow_yield    closed form of tau_fo = tau0 + c.x with uniform flow u: the
                       phi integral gives I0 and I1 of |u| p_T/T, the eta integral
                       K1 and K0 of u^tau m_T/T
"""
import numpy as np
from scipy.special import iv, kv

import fivo2d as F
from fivo2d import freezeout as FO

HBARC = 0.1973269804
PREF = FO.FYIELD / (2 * np.pi * HBARC) ** 3

# a fine grid on which the rectangle sums equal the integrals to round-off
# (eta: 121 points on [-6, 6] resolves cosh(eta) exp(-z cosh(eta)) for z up to
# u^tau m_T/T of p_T = 3 GeV; phi: 24 points are exact for the periodic
# integrand to 1e-10 at |u| p_T/T <= 3.4)
FINE = dict(n_eta=121, n_pt=30, n_phi=24, eta_max=6.0, pt_lo=0.1, pt_hi=3.0)


def medium():
    return F.build_model_2d(eos=F.ConformalHQEOS())


def tracker(lib, n=12, box=6.0, T_fo=FO.T_FO):
    """A tracker on an n x n context (the state is never evolved)."""
    g = F.make_grid2d(n, n, xmax=box, ymax=box)
    s = F.Solver(g, medium(), lib=lib)
    return s, FO.FreezeOut(s, T_fo)


def set_prims(s, T, ux=None, uy=None):
    """Put interior temperatures T (nx, ny) and flow into the context's work."""
    g = s.g
    sx, sy = g.interior()
    yT = np.full((g.Nxtot, g.Nytot), np.log(1e-3))
    yT[sx, sy] = np.log(T)
    w = {"yT": yT.ravel()}
    for k, v in (("ux", ux), ("uy", uy)):
        a = np.zeros((g.Nxtot, g.Nytot))
        if v is not None:
            a[sx, sy] = v
        w[k] = a.ravel()
    s.set_work(w)


def crossing_histories(lib):
    """Run six kinds of history on a 12 x 12 tracker, samples every 0.1 fm from
    0.4 (each switch of a history falls half-way between two samples). Returns
    (surface, expected) with the expected tau_fo, u^x and counts."""
    s, fo = tracker(lib)
    n = 12
    taus = 0.4 + 0.1 * np.arange(30)
    kind = np.arange(n * n).reshape(n, n) % 6
    rate = 0.021 + 0.004 * (np.arange(n * n).reshape(n, n) % 7)      # GeV / fm, no crossing on a sample
    T1 = 0.2
    for j, tau in enumerate(taus):
        T = np.empty((n, n))
        lin = T1 - rate * (tau - 0.4)
        T[kind == 0] = lin[kind == 0]                                  # linear cooling
        T[kind == 1] = 0.1                                             # never hot
        T[kind == 2] = 0.3                                             # hot throughout
        # re-heating: cools through T_fo near 0.95, re-heats at 1.6, cools for good at 2.4
        T[kind == 3] = 0.2 - 0.08 * (tau - 0.4) if tau < 1.55 else (0.17 if tau < 2.35 else 0.15)
        T[kind == 4] = 0.2 if 0.75 <= tau < 1.25 else 0.1              # cold first, hot, cold
        T[kind == 5] = 0.2 if tau < 0.95 else 0.1                      # a step down
        ux = np.full((n, n), 0.01 * j)
        set_prims(s, T, ux=ux, uy=-ux)
        fo.record(tau)
    exp_tau = np.full((n, n), np.nan)
    exp_ux = np.zeros((n, n))
    # linear cooling: the interpolated crossing is the exact one; the flow is that
    # of the first sample at or below T_fo
    c0 = kind == 0
    tstar = 0.4 + (T1 - FO.T_FO) / rate
    exp_tau[c0] = tstar[c0]
    exp_ux[c0] = 0.01 * np.searchsorted(taus, tstar[c0], side="left")
    # the others cross between two samples (tau_a, T_a) -> (tau_a + 0.1, T_b)
    for k, tau_a, Ta, Tb, j in ((3, 2.3, 0.17, 0.15, 20), (4, 1.2, 0.2, 0.1, 9), (5, 0.9, 0.2, 0.1, 6)):
        c = kind == k
        exp_tau[c] = tau_a + 0.1 * (Ta - FO.T_FO) / (Ta - Tb)
        exp_ux[c] = 0.01 * j
    expected = dict(tau_fo=exp_tau, uxf=exp_ux, nhot=int(np.sum(kind != 1)),
                    nfroze=int(np.sum((kind == 0) | (kind >= 3))), refroze=int(np.sum(kind == 3)))
    sf = fo.surface()
    s.close()
    return sf, expected


def static_yield(tau0, ncells, dx, dy, T=FO.T_FO, m=FO.M_PI, bose=False, nbose=80, **grid):
    """Closed form of a uniform surface at rest over ncells cells."""
    o = dict(FINE, **grid)
    pt = o["pt_lo"] + np.arange(o["n_pt"]) * (o["pt_hi"] - o["pt_lo"]) / (o["n_pt"] - 1)
    dpt = (o["pt_hi"] - o["pt_lo"]) / (o["n_pt"] - 1)
    mT = np.sqrt(pt * pt + m * m)
    ns = np.arange(1, nbose + 1) if bose else np.array([1])
    s_eta = np.array([np.sum(2 * mt * kv(1, ns * mt / T)) for mt in mT])
    return PREF * tau0 * ncells * dx * dy * 2 * np.pi * np.sum(pt * dpt * s_eta)


def tilted_flow_yield(taus, gx, gy, ux, uy, dx, dy, T=FO.T_FO, m=FO.M_PI, bose=False, nbose=80, **grid):
    """Closed form of cells at tau = taus (array) on a plane of gradient (gx, gy)
    with uniform flow (ux, uy): per cell and p_T,
      2 pi tau [m_T 2 K1(n b) I0(n a) - p_T (g.u/|u|) 2 K0(n b) I1(n a)],
    a = |u| p_T/T, b = u^tau m_T/T, summed over n >= 1 for Bose."""
    o = dict(FINE, **grid)
    pt = o["pt_lo"] + np.arange(o["n_pt"]) * (o["pt_hi"] - o["pt_lo"]) / (o["n_pt"] - 1)
    dpt = (o["pt_hi"] - o["pt_lo"]) / (o["n_pt"] - 1)
    mT = np.sqrt(pt * pt + m * m)
    u = np.hypot(ux, uy)
    ut = np.sqrt(1 + u * u)
    gu = (gx * ux + gy * uy) / u if u > 0 else 0.0
    ns = np.arange(1, nbose + 1) if bose else np.array([1])
    per_pt = []
    for p, mt in zip(pt, mT):
        a, b = u * p / T, ut * mt / T
        per_pt.append(np.sum(mt * 2 * kv(1, ns * b) * iv(0, ns * a) - p * gu * 2 * kv(0, ns * b) * iv(1, ns * a)))
    return PREF * np.sum(taus) * dx * dy * 2 * np.pi * np.sum(pt * dpt * np.array(per_pt))


def plane_surface(fo, tau0, gx, gy, mask=None):
    """tau_fo = tau0 + gx x + gy y on the tracker's interior (NaN off `mask`)."""
    X, Y = np.meshgrid(fo.xs, fo.ys, indexing="ij")
    t = tau0 + gx * X + gy * Y
    if mask is not None:
        t = np.where(mask, t, np.nan)
    return t


def counted(mask):
    """Cells the integral takes: 1..n-2 in each direction, frozen, with a
    frozen neighbour on at least one side in x and in y."""
    m = np.asarray(mask, dtype=bool)
    ok = np.zeros_like(m)
    ok[1:-1, 1:-1] = m[1:-1, 1:-1] & (m[:-2, 1:-1] | m[2:, 1:-1]) & (m[1:-1, :-2] | m[1:-1, 2:])
    return ok
