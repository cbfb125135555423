"""The reference's fixed-state gates of the charm first- and second-moment
sources, run on a library's probe entry points (the oracle's orc_probe_* on the
CPU, the HIP library's fivo_probe_* on the GPU). Test infrastructure shared by
test_oracle_source_gates.py and test_gpu_source_gates.py.

  Gc1-Gc4  test/test_consistent_fm2d.jl:73-221   first moment: 1-D limit, rotation,
                                                  Bjorken limit, zero source
  Gm1-Gm4  test/test_consistent_m22d.jl:56-303   second moment: 1-D limit, rotation
                                                  covariance, tracelessness, Bjorken
  Gd3      test/test_dnmr2d.jl:107-139           brute-force index algebra, here for
                                                  every term of the second moment
"""
import numpy as np

from fivo2d import probes
from fivo2d import _abi
from referees import Milne, fm_extras_1d, m2_coeffs_1d, m2_rhs_1d

# term bits (the order of the reference's Terms register, src/terms.jl:82-134)
BIT = {name: k for k, name in enumerate(("nu_gradalpha", "fm_gradT", "fm_inertial", "fm_nu_gradu", "fm_expansion",
                                          "fm_dlnh", "m2_nu_gradient", "m2_bg_gradu", "m2_bg_DlnT", "m2_bg_Dalpha",
                                          "m2_expansion", "m2_pi_sigma", "m2_PiQ_sigma", "m2_vorticity",
                                          "m2_projector", "m2_accel_nu", "m2_nu_gradTh", "shear_vorticity"))}
M2_TERMS = [k for k in BIT if k.startswith("m2_")]
ALL_TERMS = (1 << len(BIT)) - 1


def fm_row(ux, uy, ut, T, dxT, dyT, dtT, nux, nuy, n, dn_dT, taun, Ds, h, hp, th, ax, ay, dxux, dxuy, dyux, dyuy):
    return [ux, uy, ut, T, dxT, dyT, dtT, nux, nuy, n, dn_dT, taun, Ds, h, hp, th, ax, ay, dxux, dxuy, dyux, dyuy]


# ---- first moment --------------------------------------------------------------------
GC1_POINTS = [   # test/test_consistent_fm2d.jl:109-119
    dict(tau=2.3, r=3.1, ur=0.42, T=0.31, dtT=-0.055, drT=-0.021, drur=0.13, dtur=0.07, n=0.9, dn_dT=1.7, nu=0.031,
         Ds=0.44, h=2.20, hp=-0.9),
    dict(tau=1.1, r=0.7, ur=-0.30, T=0.45, dtT=-0.120, drT=0.033, drur=-0.21, dtur=0.15, n=1.4, dn_dT=2.9, nu=-0.070,
         Ds=0.30, h=1.90, hp=-1.4),
    dict(tau=5.0, r=8.0, ur=0.90, T=0.20, dtT=-0.010, drT=-0.005, drur=0.02, dtur=0.01, n=0.3, dn_dT=0.8, nu=0.010,
         Ds=0.60, h=2.60, hp=-0.5),
    dict(tau=0.6, r=0.05, ur=-1.30, T=0.55, dtT=0.300, drT=-0.440, drur=0.55, dtur=-0.62, n=2.2, dn_dT=4.1, nu=0.190,
         Ds=0.12, h=1.75, hp=-2.2),
    dict(tau=3.7, r=12.0, ur=0.05, T=0.16, dtT=-0.002, drT=-0.0007, drur=0.004, dtur=0.002, n=0.05, dn_dT=0.3,
         nu=1e-4, Ds=0.80, h=3.10, hp=-0.2),
]


def gate_Gc1(lib):
    """1-D limit (test/test_consistent_fm2d.jl:73-142): the 2-D source on the +x axis
    with u^y = 0, d_y = 0 except d_y u^y = u^r / r reproduces the radial source.
    Returns (worst relative deviation, worst |s^y| / max(|s^x|, 1))."""
    rows, refs = [], []
    for p in GC1_POINTS:
        taun = p["Ds"] * p["h"] / p["T"]
        refs.append(fm_extras_1d(p["tau"], p["r"], p["ur"], p["T"], p["dtT"], p["drT"], p["drur"], p["dtur"],
                                 p["n"], p["dn_dT"], p["nu"], taun, p["Ds"], p["h"], p["hp"]))
        ux, uy = p["ur"], 0.0
        ut = np.sqrt(1 + ux * ux)
        ax, ay = ut * p["dtur"] + p["ur"] * p["drur"], 0.0
        dxux, dxuy, dyux, dyuy = p["drur"], 0.0, 0.0, p["ur"] / p["r"]
        th = p["ur"] * p["dtur"] / ut + dxux + dyuy + ut / p["tau"]
        rows.append(fm_row(ux, uy, ut, p["T"], p["drT"], 0.0, p["dtT"], p["nu"], 0.0, p["n"], p["dn_dT"], taun,
                           p["Ds"], p["h"], p["hp"], th, ax, ay, dxux, dxuy, dyux, dyuy))
    s = probes.fm_source(lib, rows)
    refs = np.array(refs)
    rel = np.abs(s[:, 0] - refs) / np.abs(refs)
    sy = np.abs(s[:, 1]) / np.maximum(np.abs(s[:, 0]), 1.0)
    return float(rel.max()), float(sy.max())


def gate_Gc2(lib, terms=None):
    """Rotation (test/test_consistent_fm2d.jl:147-184): the rotated state's source
    is the rotated source. Returns the worst relative deviation over four angles."""
    ux, uy = 0.31, -0.22
    ut = np.sqrt(1 + ux * ux + uy * uy)
    nu = np.array([0.017, 0.009])
    gT = np.array([-0.03, 0.012])
    G = np.array([[0.11, -0.07], [0.05, 0.09]])
    tau, T, dtT, n, dn_dT, taun, Ds, h, hp = 2.0, 0.28, -0.04, 1.1, 2.0, 0.55, 0.35, 2.1, -1.0
    dtut = 0.02
    a = np.array([ut * 0.03 + ux * G[0, 0] + uy * G[1, 0], ut * 0.01 + ux * G[0, 1] + uy * G[1, 1]])

    def row(u, v, gt, g, av):
        th = dtut + g[0, 0] + g[1, 1] + ut / tau
        return fm_row(u[0], u[1], ut, T, gt[0], gt[1], dtT, v[0], v[1], n, dn_dT, taun, Ds, h, hp, th, av[0], av[1],
                      g[0, 0], g[0, 1], g[1, 0], g[1, 1])

    rows = [row(np.array([ux, uy]), nu, gT, G, a)]
    Rs = []
    for phi in (0.3, 1.0, 2.4, -0.8):
        R = np.array([[np.cos(phi), -np.sin(phi)], [np.sin(phi), np.cos(phi)]])
        Rs.append(R)
        rows.append(row(R @ [ux, uy], R @ nu, R @ gT, R @ G @ R.T, R @ a))
    s = probes.fm_source(lib, rows, terms)
    worst = 0.0
    for k, R in enumerate(Rs):
        e = R @ s[0]
        worst = max(worst, float(np.max(np.abs(s[k + 1] - e)) / np.hypot(*s[0])))
    return worst


def gate_Gc3_Gc4(lib):
    """Bjorken limit and zero source (test/test_consistent_fm2d.jl:189-221): at rest
    with no gradients the source is tau_n nu^i / tau; with no current and no
    gradients it is exactly zero. Returns (relative deviation, (s^x, s^y) of Gc4)."""
    tau, T, n, taun, Ds, h, hp = 3.0, 0.30, 1.0, 0.7, 0.4, 2.2, -1.0
    nux, nuy = 0.02, -0.013
    r3 = fm_row(0, 0, 1.0, T, 0, 0, 0, nux, nuy, n, 0.0, taun, Ds, h, hp, 1 / tau, 0, 0, 0, 0, 0, 0)
    r4 = fm_row(0, 0, 1.0, 0.30, 0, 0, 0, 0, 0, 1.0, 2.0, 0.7, 0.4, 2.2, -1.0, 1 / 3, 0, 0, 0, 0, 0, 0)
    s = probes.fm_source(lib, [r3, r4])
    ex = np.array([taun * nux / tau, taun * nuy / tau])
    return float(np.max(np.abs(s[0] - ex) / np.abs(ex))), s[1]


# ---- second moment -----------------------------------------------------------------------
def m2_row(ux, uy, ut, tau, T, dxT, dyT, dtT, nux, nuy, pxx, pxy, pyy, peta, PiQ, dta, dxa, dya, thnu, anux, anuy,
           dtnux, dtnuy, dxnux, dxnuy, dynux, dynuy, th, ax, ay, dtux, dtuy, dxux, dxuy, dyux, dyuy,
           n, taun, Ds, h, hp, tauM, etaM, m):
    return [ux, uy, ut, tau, T, dxT, dyT, dtT, nux, nuy, pxx, pxy, pyy, peta, PiQ, dta, dxa, dya, thnu, anux, anuy,
            dtnux, dtnuy, dxnux, dxnuy, dynux, dynuy, th, ax, ay, dtux, dtuy, dxux, dxuy, dyux, dyuy,
            n, taun, Ds, h, hp, tauM, etaM, m]


GM1_POINTS = [   # test/test_consistent_m22d.jl:129-141
    dict(tau=2.3, r=3.1, ur=0.42, T=0.31, dtT=-0.055, drT=-0.021, drur=0.13, dtur=0.07, n=0.9, dn_dT=1.7, nu=0.031,
         Ds=0.44, h=2.20, hp=-0.9, p_l=1.0e-3, p_phi=-4.0e-4, PiQ=2.0e-4, dta=0.031, dra=-0.017, dtnu=2.3e-4,
         drnu=-1.1e-4),
    dict(tau=1.1, r=0.7, ur=-0.30, T=0.45, dtT=-0.120, drT=0.033, drur=-0.21, dtur=0.15, n=1.4, dn_dT=2.9, nu=-0.070,
         Ds=0.30, h=1.90, hp=-1.4, p_l=-7.0e-4, p_phi=3.0e-4, PiQ=-1.5e-4, dta=-0.02, dra=0.04, dtnu=-1.0e-4,
         drnu=3.0e-4),
    dict(tau=5.0, r=8.0, ur=0.90, T=0.20, dtT=-0.010, drT=-0.005, drur=0.02, dtur=0.01, n=0.3, dn_dT=0.8, nu=0.010,
         Ds=0.60, h=2.60, hp=-0.5, p_l=5.0e-4, p_phi=6.0e-4, PiQ=9.0e-5, dta=0.011, dra=-0.003, dtnu=5.0e-5,
         drnu=-2.0e-5),
    dict(tau=0.6, r=1.2, ur=-1.30, T=0.55, dtT=0.300, drT=-0.440, drur=0.55, dtur=-0.62, n=2.2, dn_dT=4.1, nu=0.190,
         Ds=0.12, h=1.75, hp=-2.2, p_l=9.0e-4, p_phi=1.0e-4, PiQ=-5.0e-4, dta=0.09, dra=-0.06, dtnu=8.0e-4,
         drnu=-4.0e-4),
]
M_CHARM = 1.5


def gate_Gm1(lib):
    """1-D limit (test/test_consistent_m22d.jl:56-200): on the +x axis with u^y = 0
    the 2-D rates reproduce the radial (p_l, p_phi, Pi_Q) rates. The l-channel rate
    is the rate of the projection p_l = pi^{xx} / (u^tau)^2 along the trajectory,
    with the advection of the state's d_x pi^{xx} added back (the 2-D source leaves
    it to the solver's upwinding). Returns (worst relative deviation, worst |rate
    of pi^{xy}|)."""
    rows, refs, meta = [], [], []
    for p in GM1_POINTS:
        taun = p["Ds"] * p["h"] / p["T"]
        z = M_CHARM / p["T"]
        from scipy.special import kve
        K2, K3, K4 = kve(2, z), kve(3, z), kve(4, z)
        tauM = taun * K4 * K2 / (2 * K3 ** 2)
        etaM = p["T"] * taun / 2
        refs.append(m2_rhs_1d(p["tau"], p["r"], p["ur"], p["T"], p["dtT"], p["drT"], p["drur"], p["dtur"],
                              p["p_l"], p["p_phi"], p["PiQ"], p["nu"], p["dta"], p["dra"], p["dtnu"], p["drnu"],
                              (0.0, 0.0, 0.0), p["n"], taun, p["Ds"], p["h"], p["hp"], tauM, etaM, M_CHARM))
        ur, r, tau = p["ur"], p["r"], p["tau"]
        ux, uy = ur, 0.0
        ut = np.sqrt(1 + ux * ux)
        ax, ay = ut * p["dtur"] + ur * p["drur"], 0.0
        dxux, dxuy, dyux, dyuy = p["drur"], 0.0, 0.0, ur / r
        th = ur * p["dtur"] / ut + dxux + dyuy + ut / tau
        nu = p["nu"]
        dxnx, dxny, dynx, dyny = p["drnu"], 0.0, 0.0, nu / r
        anx, any_ = ut * p["dtnu"] + ur * p["drnu"], 0.0
        nut = ur * nu / ut
        dtnut = (ur / ut) * p["dtnu"] + nu * p["dtur"] / ut ** 3
        thnu = dtnut + dxnx + dyny + nut / tau
        pxx = p["p_l"] * ut ** 2
        peta = Milne(tau, ux, uy, np.zeros((3, 2))).implied_peta(pxx, 0.0, p["p_phi"])
        rows.append(m2_row(ux, uy, ut, tau, p["T"], p["drT"], 0.0, p["dtT"], nu, 0.0, pxx, 0.0, p["p_phi"], peta,
                           p["PiQ"], p["dta"], p["dra"], 0.0, thnu, anx, any_, p["dtnu"], 0.0, dxnx, dxny, dynx,
                           dyny, th, ax, ay, p["dtur"], 0.0, dxux, dxuy, dyux, dyuy, p["n"], taun, p["Ds"], p["h"],
                           p["hp"], tauM, etaM, M_CHARM))
        meta.append((ut, pxx))
    got = probes.m2_source(lib, rows)
    worst, wxy = 0.0, 0.0
    for p, g, ref, (ut, pxx) in zip(GM1_POINTS, got, refs, meta):
        dxpxx = 2 * pxx * p["ur"] * p["drur"] / ut ** 2
        rate_xx = g[0] - (p["ur"] / ut) * dxpxx
        pl = rate_xx / ut ** 2 - 2 * pxx * p["ur"] * p["dtur"] / ut ** 4
        for a, b in ((pl, ref[0]), (g[2], ref[1]), (g[3], ref[2])):
            worst = max(worst, abs(a - b) / abs(b))
        wxy = max(wxy, abs(g[1]) / max(abs(g[0]), 1.0))
    return worst, wxy


def gm2_state():
    """The state of gate Gm2 (test/test_consistent_m22d.jl:207-225)."""
    return dict(ux=0.31, uy=-0.22, tau=2.0, nu=np.array([0.017, 0.009]), T=0.28, dtT=-0.04,
                gT=np.array([-0.03, 0.012]), G=np.array([[0.11, -0.07], [0.05, 0.09]]),
                Gnu=np.array([[4.0e-4, -2.0e-4], [1.0e-4, 3.0e-4]]), P=np.array([[7.0e-4, 2.0e-4], [2.0e-4, -3.0e-4]]),
                a=np.array([0.05, -0.03]), anu=np.array([1.2e-4, -0.8e-4]), dtut=0.02, thnu=3.0e-4, PiQ=1.5e-4,
                da=np.array([0.031, -0.017]), dta=0.05, n=1.1, taun=0.55, Ds=0.35, h=2.1, hp=-1.0, tauM=0.30,
                etaM=0.29, m=1.5, dtu=np.array([0.03, 0.01]), dtnu=np.array([1.0e-4, -0.6e-4]))


def _gm2_row(s, R):
    ut = np.sqrt(1 + s["ux"] ** 2 + s["uy"] ** 2)
    u = R @ [s["ux"], s["uy"]]
    nu, gT, a, anu, da, dtu, dtnu = (R @ s[k] for k in ("nu", "gT", "a", "anu", "da", "dtu", "dtnu"))
    G, Gnu, P = (R @ s[k] @ R.T for k in ("G", "Gnu", "P"))
    th = s["dtut"] + G[0, 0] + G[1, 1] + ut / s["tau"]
    peta = Milne(s["tau"], u[0], u[1], np.zeros((3, 2))).implied_peta(P[0, 0], P[0, 1], P[1, 1])
    return m2_row(u[0], u[1], ut, s["tau"], s["T"], gT[0], gT[1], s["dtT"], nu[0], nu[1], P[0, 0], P[0, 1], P[1, 1],
                  peta, s["PiQ"], s["dta"], da[0], da[1], s["thnu"], anu[0], anu[1], dtnu[0], dtnu[1], Gnu[0, 0],
                  Gnu[0, 1], Gnu[1, 0], Gnu[1, 1], th, a[0], a[1], dtu[0], dtu[1], G[0, 0], G[0, 1], G[1, 0], G[1, 1],
                  s["n"], s["taun"], s["Ds"], s["h"], s["hp"], s["tauM"], s["etaM"], s["m"])


def gate_Gm2(lib, terms=None):
    """Rotation covariance (test/test_consistent_m22d.jl:205-248): the rotated
    state's traceless block is R S R^T and its trace channel is unchanged. Returns
    the worst relative deviation over four angles."""
    s = gm2_state()
    rows = [_gm2_row(s, np.eye(2))]
    Rs = []
    for phi in (0.3, 1.0, 2.4, -0.8):
        R = np.array([[np.cos(phi), -np.sin(phi)], [np.sin(phi), np.cos(phi)]])
        Rs.append(R)
        rows.append(_gm2_row(s, R))
    out = probes.m2_source(lib, rows, terms)
    S0 = np.array([[out[0, 0], out[0, 1]], [out[0, 1], out[0, 2]]])
    worst = 0.0
    for k, R in enumerate(Rs):
        S2 = np.array([[out[k + 1, 0], out[k + 1, 1]], [out[k + 1, 1], out[k + 1, 2]]])
        d = np.max(np.abs(S2 - R @ S0 @ R.T)) / np.max(np.abs(S0))
        dtr = abs(out[k + 1, 3] - out[0, 3]) / abs(out[0, 3])
        worst = max(worst, d, dtr)
    return float(worst)


def gate_Gm4(lib):
    """Bjorken limit (test/test_consistent_m22d.jl:255-282): at rest with no
    transverse gradients dpxx == dpyy and dpxy == 0. Returns (|dpxx - dpyy| /
    |dpxx|, |dpxy| / max(|dpxx|, 1), dPiQ)."""
    tau = 3.0
    row = m2_row(0, 0, 1.0, tau, 0.30, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                 1.0 / 3.0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0, 0.7, 0.4, 2.2, -1.0, 0.25, 0.24, 1.5)
    s = probes.m2_source(lib, [row])[0]
    return abs(s[0] - s[2]) / abs(s[0]), abs(s[1]) / max(abs(s[0]), 1.0), s[3]


# ---- every second-moment term against brute-force index algebra --------------------------
def random_m2_state(rng):
    """A random flowing state: velocities, current, their gradients, moments and
    coefficients in physical ranges; consistent kinematics (theta, a, theta_nu
    from the same gradients)."""
    ux, uy = 1.2 * rng.uniform(-0.5, 0.5, 2)
    tau = 0.6 + 2 * rng.uniform()
    du = 0.3 * rng.uniform(-0.5, 0.5, (3, 2))
    nux, nuy = 0.05 * rng.uniform(-0.5, 0.5, 2)
    dnu = 0.01 * rng.uniform(-0.5, 0.5, (3, 2))
    M = Milne(tau, ux, uy, du, nux, nuy, dnu)
    T = 0.15 + 0.35 * rng.uniform()
    st = dict(M=M, T=T, dtT=0.1 * rng.uniform(-0.5, 0.5), dxT=0.1 * rng.uniform(-0.5, 0.5),
              dyT=0.1 * rng.uniform(-0.5, 0.5), p=2e-3 * rng.uniform(-0.5, 0.5, 3), PiQ=1e-3 * rng.uniform(-0.5, 0.5),
              dta=0.1 * rng.uniform(-0.5, 0.5), dxa=0.1 * rng.uniform(-0.5, 0.5), dya=0.1 * rng.uniform(-0.5, 0.5),
              n=0.2 + rng.uniform(), taun=0.2 + rng.uniform(), Ds=0.1 + 0.5 * rng.uniform(), h=1.7 + rng.uniform(),
              hp=-2 * rng.uniform(), tauM=0.1 + 0.5 * rng.uniform(), etaM=0.1 + 0.3 * rng.uniform(), m=1.5,
              du=du, dnu=dnu)
    return st


def m2_state_row(st):
    M = st["M"]
    ut, ux, uy = M.u[0], M.u[1], M.u[2]
    p = st["p"]
    peta = M.implied_peta(p[0], p[1], p[2])
    du, dnu = st["du"], st["dnu"]
    thnu = np.trace(M.Dnu)
    anu = M.u @ M.Dnu                         # a_nu^i = D nu^i (an argument the source does not use)
    return m2_row(ux, uy, ut, M.tau, st["T"], st["dxT"], st["dyT"], st["dtT"], M.nu[1], M.nu[2], p[0], p[1], p[2],
                  peta, st["PiQ"], st["dta"], st["dxa"], st["dya"], thnu, anu[1], anu[2], dnu[0, 0], dnu[0, 1],
                  dnu[1, 0], dnu[1, 1], dnu[2, 0], dnu[2, 1], M.theta, M.a[1], M.a[2], du[0, 0], du[0, 1],
                  du[1, 0], du[1, 1], du[2, 0], du[2, 1], st["n"], st["taun"], st["Ds"], st["h"], st["hp"],
                  st["tauM"], st["etaM"], st["m"])


def brute_force_terms(st):
    """Each second-moment term's contribution to the left-hand source, from its
    covariant definition (consistent_m2_source_2d, src2d/hq_consistent_m2_2d.jl:
    247-427, and the equations in src/hq_consistent_m2.jl:44-60): {term: (4-block
    (xx, xy, yy, tau^2 eta-eta), trace-channel scalar)}."""
    M = st["M"]
    T, tauM = st["T"], st["tauM"]
    p = st["p"]
    Pi = M.ortho_tensor(p[0], p[1], p[2], M.implied_peta(p[0], p[1], p[2]))
    cf = m2_coeffs_1d(T, st["n"], st["taun"], st["Ds"], st["h"], st["m"])
    sig = M.sigma()
    snu, thnu = M.sigma_nu()
    etaQ = st["etaM"]
    ut = M.u[0]
    dT = np.array([st["dtT"], st["dxT"], st["dyT"], 0.0])
    DT = M.u @ dT
    DlnT = DT / T
    gradT = M.Delta @ dT                                   # nabla^<mu> T
    gTh = (st["h"] + T * st["hp"]) * gradT
    da = np.array([st["dta"], st["dxa"], st["dya"], 0.0])
    Da = M.u @ da
    Z = np.zeros((4, 4))
    out = {}
    out["m2_nu_gradient"] = (M.block(2 * etaQ * snu), (5.0 / 3.0) * etaQ * thnu)
    out["m2_bg_gradu"] = (M.block(2 * cf["etabar"] * sig), cf["etabar"] * (5.0 / 3.0) * M.theta)
    out["m2_bg_DlnT"] = (M.block(Z), cf["etabar"] * cf["ABr"] * DlnT)
    out["m2_bg_Dalpha"] = (M.block(Z), cf["etabar"] * Da)
    geo = tauM * ((5.0 / 3.0) * M.theta + cf["dlnC"] * DlnT)
    out["m2_expansion"] = (M.block(geo * Pi), geo * st["PiQ"])
    ps = M.contract(Pi, sig)
    out["m2_pi_sigma"] = (M.block(2 * tauM * M.traceless_sym(Pi @ M.g @ sig)), (2.0 / 3.0) * tauM * ps)
    out["m2_PiQ_sigma"] = (M.block(2 * tauM * st["PiQ"] * sig), 0.0)
    out["m2_vorticity"] = (M.block(tauM * 2 * M.traceless_sym(Pi @ M.g @ M.omega())), 0.0)
    c = Pi @ M.g @ M.a                                     # c^j = pi^{j beta} a_beta
    out["m2_projector"] = (M.block(-tauM * (np.outer(M.u, c) + np.outer(c, M.u))), 0.0)
    out["m2_accel_nu"] = (M.block(2 * cf["lam_a"] * M.rank1(M.a, M.nu)),
                          (st["Ds"] * cf["Aco"] / (3 * T)) * M.dot(M.a, M.nu))
    out["m2_nu_gradTh"] = (M.block((st["Ds"] / T) * M.rank1(M.nu, gTh)),
                           (5.0 / 6.0) * (st["Ds"] / T) * M.dot(M.nu, gTh))
    return out


def gate_Gd3_m2(lib, nstates=300, seed=7):
    """Every second-moment term, isolated by switching it alone on (the source is a
    sum of switched terms, so a term's contribution is the rate with it on minus
    the rate with every term off, times -tau_M u^tau), against its brute-force
    covariant form at random flowing states; also the eta-eta component the
    solver's block implies by tracelessness against the brute-force one (gate
    Gm3). Returns {term: (worst block deviation, worst trace deviation, worst
    eta-eta deviation)}, relative to the term's own scale."""
    rng = np.random.default_rng(seed)
    states = [random_m2_state(rng) for _ in range(nstates)]
    rows = [m2_state_row(st) for st in states]
    base = probes.m2_source(lib, rows, 0)
    res = {}
    for name in M2_TERMS:
        got = probes.m2_source(lib, rows, 1 << BIT[name])
        wb = wt = we = 0.0
        for st, r, g, b in zip(states, rows, got, base):
            M = st["M"]
            f = -st["tauM"] * M.u[0]
            S = (g - b) * f                                   # the term's left-hand contribution
            blk, tr = brute_force_terms(st)[name]
            scale = max(np.max(np.abs(blk[:3])), 1e-3 * max(np.max(np.abs(st["p"])), 1e-12))
            wb = max(wb, float(np.max(np.abs(S[:3] - blk[:3])) / scale))
            tscale = max(abs(tr), 1e-3 * abs(st["PiQ"]), 1e-300)
            wt = max(wt, abs(S[3] - tr) / tscale)
            if name != "m2_projector":
                # the projector term -(u^i c^j + u^j c^i) is not u-orthogonal by itself:
                # it is what makes tau_M Delta^{ij}_{ab} D pi^{ab} transverse, so no
                # eta-eta component follows from its block
                eta_implied = M.implied_peta(S[0], S[1], S[2])
                we = max(we, float(abs(eta_implied - blk[3]) / max(np.max(np.abs(blk)), 1e-300)))
        res[name] = (wb, wt, we)
    return res


FM_TERMS = ("fm_gradT", "fm_inertial", "fm_nu_gradu", "fm_expansion", "fm_dlnh")


def gate_Gd3_fm(lib, nstates=300, seed=5):
    """Every first-moment term alone against its covariant form at random flowing
    states (the source with every term off is zero):
      fm_gradT     (D_s/T)(n + T dn/dT) nabla^<i> T
      fm_inertial  tau_n n a^i
      fm_nu_gradu  tau_n nu^m nabla_m u^i
      fm_expansion tau_n theta nu^i
      fm_dlnh      (D_s/T) h' DT nu^i
    (src2d/hq_consistent_firstmoment2d.jl:218-282). Returns {term: worst relative
    deviation}."""
    rng = np.random.default_rng(seed)
    states = [random_m2_state(rng) for _ in range(nstates)]
    rows = []
    for st in states:
        M = st["M"]
        du = st["du"]
        st["dn_dT"] = 3 * st["n"] * rng.uniform()
        rows.append(fm_row(M.u[1], M.u[2], M.u[0], st["T"], st["dxT"], st["dyT"], st["dtT"], M.nu[1], M.nu[2],
                           st["n"], st["dn_dT"], st["taun"], st["Ds"], st["h"], st["hp"], M.theta, M.a[1], M.a[2],
                           du[1, 0], du[1, 1], du[2, 0], du[2, 1]))
    res = {}
    for name in FM_TERMS:
        got = probes.fm_source(lib, rows, 1 << BIT[name])
        w = 0.0
        for st, g in zip(states, got):
            M, T = st["M"], st["T"]
            dT = np.array([st["dtT"], st["dxT"], st["dyT"], 0.0])
            if name == "fm_gradT":
                v = (st["Ds"] / T) * (st["n"] + T * st["dn_dT"]) * (M.Delta @ dT)
            elif name == "fm_inertial":
                v = st["taun"] * st["n"] * M.a
            elif name == "fm_nu_gradu":
                v = st["taun"] * (M.nu @ M.Du)
            elif name == "fm_expansion":
                v = st["taun"] * M.theta * M.nu
            else:
                v = (st["Ds"] / T) * st["hp"] * (M.u @ dT) * M.nu
            w = max(w, float(np.max(np.abs(g - v[1:3])) / np.max(np.abs(v[1:3]))))
        res[name] = w
    return res


def m2_random_rows(nstates=200, seed=11):
    rng = np.random.default_rng(seed)
    return np.array([m2_state_row(random_m2_state(rng)) for _ in range(nstates)])


def fm_random_rows(nstates=200, seed=13):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(nstates):
        st = random_m2_state(rng)
        M = st["M"]
        du = st["du"]
        rows.append(fm_row(M.u[1], M.u[2], M.u[0], st["T"], st["dxT"], st["dyT"], st["dtT"], M.nu[1], M.nu[2],
                           st["n"], 2 * st["n"], st["taun"], st["Ds"], st["h"], st["hp"], M.theta, M.a[1], M.a[2],
                           du[1, 0], du[1, 1], du[2, 0], du[2, 1]))
    return np.array(rows)


def default_terms():
    return _abi.TERMS_DEFAULT
