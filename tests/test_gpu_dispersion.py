"""Gate Gk on the HIP path (protocol in test_oracle_dispersion.py,
test/test_charge_dispersion2d.jl:375-589): the overdamped branch (Gk-a), the
propagating branch with its k-independent damping (Gk-b), the collision bracket
around the published k* = 0.733 1/fm (Gk-c), the tau_n -> 0 control (Gk-d), the
first-order numerical-diffusion floor on the propagating branch only (Gk-e), and
one mode's measured w against the oracle's run."""
import math

import pytest

from test_oracle_dispersion import (C_NUM, KFUND, LBOX, fick_excess, gk_a, kstar, run_mode, tauD_for_kstar,
                                    D_s, tau_n)

pytestmark = pytest.mark.gpu

TAUD_OVER = None
TAUD_PROP = None


def _taus():
    global TAUD_OVER, TAUD_PROP
    if TAUD_OVER is None:
        TAUD_OVER = tauD_for_kstar(2.5 * KFUND)
        TAUD_PROP = tauD_for_kstar(1.4 * KFUND)
    return TAUD_OVER, TAUD_PROP


def test_gk_a_overdamped_branch(hip):
    gk_a(hip)


def test_gk_b_propagating_branch(hip):
    _, prop = _taus()
    rs = []
    for harm in (2, 3):
        r = run_mode(hip, harmonic=harm, DsT=0.75, tauD=prop, trun=18.0, dts=0.2)
        Om, Omr = r["w_meas"].real, r["w_ref"].real
        g, g0 = -r["w_meas"].imag, 1 / (2 * r["tn"])
        print(f"    m={harm} k={r['k']:.4f} (k/k*={r['k'] / r['kstar']:.3f}) | Omega meas {Om:.5f} ref {Omr:.5f} "
              f"({100 * (Om / Omr - 1):.2f}%) | gamma {g:.5f} floor {g0:.5f} excess {(g - g0) / (r['dx'] * r['k'] ** 2):.4f} "
              f"dx k^2", flush=True)
        assert r["k"] > r["kstar"] and abs(r["w_ref"].imag + 1 / (2 * r["tn"])) < 1e-10
        assert 0.97 < Om / Omr < 1.03
        assert g0 <= g <= g0 + C_NUM * r["dx"] * r["k"] ** 2
        rs.append((Om, g, r))
    (O1, g1, r1), (O2, g2, r2) = rs
    print(f"    k {r1['k']:.3f} -> {r2['k']:.3f}: Omega x{O2 / O1:.3f} gamma x{g2 / g1:.3f} "
          f"(diffusive: x{(r2['k'] / r1['k']) ** 2:.3f})", flush=True)
    assert O2 / O1 > 1.2 and abs(g2 / g1 - 1) < 0.10


def test_gk_c_collision_brackets_kstar(hip):
    _, prop = _taus()
    ks = kstar(D_s(0.75), tau_n(0.75, prop))
    lastflat, firstosc = 0.0, math.inf
    for harm in range(1, 7):
        r = run_mode(hip, harmonic=harm, DsT=0.75, tauD=prop, trun=10.0, dts=0.2)
        osc = abs(r["w_meas"].real) > 0.05 * abs(r["w_meas"].imag)
        print(f"    m={harm} k={r['k']:.4f} k/k*={r['k'] / ks:.3f} | Re w {r['w_meas'].real:.5f} "
              f"Im w {r['w_meas'].imag:.5f} | {'PROPAGATING' if osc else 'overdamped'}", flush=True)
        if osc:
            firstosc = min(firstosc, r["k"])
        else:
            lastflat = max(lastflat, r["k"])
    print(f"  measured transition in ({lastflat:.4f}, {firstosc:.4f}); predicted k* = {ks:.4f}", flush=True)
    assert lastflat < ks < firstosc and f"{ks:.3f}" == "0.733"


def test_gk_d_control_tau_n_to_zero(hip):
    _, prop = _taus()
    rows = []
    for fac, trun, dts in ((1.0, 20.0, 0.2), (0.25, 6.0, 0.1), (0.0025, 1.2, 0.04)):
        r = run_mode(hip, harmonic=6, DsT=0.75, tauD=prop * fac, trun=trun, dts=dts)
        g = -r["w_meas"].imag
        print(f"    tau_n = {r['tn']:.4f} fm k/k* = {r['k'] / r['kstar']:.3f} | Re w {r['w_meas'].real:.5f} | "
              f"gamma {g:.5f} | gamma/(D k^2) {g / (r['D'] * r['k'] ** 2):.4f} | split {r['split']:.4f}", flush=True)
        rows.append(r)
    assert abs(rows[0]["w_meas"].real) > 0.2 * abs(rows[0]["w_meas"].imag)
    assert abs(rows[1]["w_meas"].real) > 0.2 * abs(rows[1]["w_meas"].imag)
    assert rows[2]["k"] < rows[2]["kstar"]
    assert abs(rows[2]["w_meas"].real) < 1e-9 * abs(rows[2]["w_meas"].imag)
    xk3 = (rows[2]["k"] / rows[2]["kstar"]) ** 2
    rat = -rows[2]["w_meas"].imag / (rows[2]["D"] * rows[2]["k"] ** 2) / fick_excess(xk3)
    print(f"    tau_n -> 0 recovers Fick to {100 * (rat - 1):.2f}%", flush=True)
    assert 0.98 < rat < 1.06
    assert rows[1]["w_meas"].real > rows[0]["w_meas"].real


def test_gk_e_numerical_diffusion_first_order_above_kstar_only(hip):
    over, prop = _taus()
    dn, dxs = [], []
    for N in (96, 192, 384):
        r = run_mode(hip, harmonic=4, DsT=0.75, tauD=prop, trun=12.0, dts=0.2, N=N)
        g, g0 = -r["w_meas"].imag, 1 / (2 * r["tn"])
        dn.append((g - g0) / r["k"] ** 2)
        dxs.append(r["dx"])
        print(f"      N={N:4d} dx={r['dx']:.5f} | gamma {g:.6f} excess {g - g0:.6f} | D_num {dn[-1]:.4e} = "
              f"{dn[-1] / r['dx']:.4f} dx ({100 * dn[-1] / r['D']:.2f}% of D_s)", flush=True)
    order = [math.log2(dn[i] / dn[i + 1]) for i in range(2)]
    print(f"      observed order in dx: {order[0]:.2f}, {order[1]:.2f}", flush=True)
    assert all(d > 0 for d in dn) and dn[1] < dn[0] and dn[2] < dn[1]
    assert all(0.80 < o < 1.35 for o in order)
    devs = []
    for N in (96, 192, 384):
        r = run_mode(hip, harmonic=1, DsT=0.75, tauD=over, trun=12.0, N=N)
        devs.append(-r["w_meas"].imag / -r["w_ref"].imag - 1)
        print(f"      N={N:4d} overdamped m=1 | dev {100 * devs[-1]:+.4f}%", flush=True)
    assert max(abs(d) for d in devs) < 3e-3 and min(devs) < 0 < max(devs)
    gs = []
    for CFL in (0.30, 0.15):
        r = run_mode(hip, harmonic=4, DsT=0.75, tauD=prop, trun=12.0, dts=0.2, CFL=CFL)
        gs.append(-r["w_meas"].imag)
    dtau_effect = abs(gs[1] - gs[0])
    dx_effect = abs(dn[1] - dn[0]) * (2 * math.pi * 4 / LBOX) ** 2
    print(f"      halving dtau moves gamma by {dtau_effect:.3e}; halving dx moved it by {dx_effect:.3e}", flush=True)
    assert dtau_effect < 0.25 * dx_effect


def test_gk_hip_equals_oracle(hip, oracle):
    over, prop = _taus()
    a = run_mode(hip, harmonic=3, DsT=0.75, tauD=prop, trun=6.0, dts=0.2)
    b = run_mode(oracle, harmonic=3, DsT=0.75, tauD=prop, trun=6.0, dts=0.2)
    d = max(abs(x - y) for x, y in zip(a["la"], b["la"]))
    p = max(abs(x - y) for x, y in zip(a["phs"], b["phs"]))
    print(f"    HIP w {a['w_meas']:.12f}  oracle w {b['w_meas']:.12f}  max d ln|c| {d:.1e}  max d arg c {p:.1e}",
          flush=True)
    assert d < 1e-9 and p < 1e-9
