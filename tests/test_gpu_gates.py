"""The reference's validation gates run on the HIP path.

Each gate is run twice, by the HIP library and by the oracle, on the same
inputs: the HIP run must pass the gate's own bounds (the published numbers
and thresholds of test/test_gubser2d.jl, test/test_bjorken2d.jl,
test/test_bjorken_bulk2d.jl, test/test_dissipation2d.jl, test/test_charge2d.jl)
and land on the oracle's final state within the final-time bound (1e-9 relative
in T, u and the dissipative fields' scale).
"""
import numpy as np
import pytest

import fivo2d as F
import test_oracle_dissipative_gates as D
import test_oracle_gates as G
from oracle_unit import OracleUnit

pytestmark = pytest.mark.gpu


def test_gubser_gate_g1_on_gpu(oracle, hip):
    for N, (eT, eU) in ((100, (3.29e-3, 7.73e-3)), (200, (8.81e-4, 1.92e-3))):
        h = G.run_gubser(hip, N)
        o = G.run_gubser(oracle, N)
        print(f"N {N}: L2(T) {h[0]:.4e} (oracle {o[0]:.4e}), L2(u) {h[1]:.4e} (oracle {o[1]:.4e}), "
              f"x<->y {h[2]:.1e}", flush=True)
        assert abs(h[0] - eT) < 0.005 * 10 ** np.floor(np.log10(eT))
        assert abs(h[1] - eU) < 0.006 * 10 ** np.floor(np.log10(eU))
        assert abs(h[0] - o[0]) <= 1e-9 * o[0] and abs(h[1] - o[1]) <= 1e-9 * o[1]
        assert h[2] < 1e-12


def test_gubser_gate_g1_three_resolutions_on_gpu(hip):
    """G1 at N = 100 / 200 / 400 (test/test_gubser2d.jl:203-245): the published
    L2(T) 3.29e-3 -> 8.81e-4 -> 2.26e-4, L2(u) 7.73e-3 -> 1.92e-3 -> 4.57e-4, orders
    T 1.90, 1.96 and u 2.01, 2.07, the entropy outflow -0.116 and its error against
    the analytic disc entropy 1.28e-3 -> 9.56e-5 (README2D.md:275); x<->y < 1e-12,
    ring spread converging and < 5e-3."""
    rs = [G.run_gubser_full(hip, N) for N in (100, 200, 400)]
    for N, r in zip((100, 200, 400), rs):
        print(f"N={N}: L2(T) {r['L2T']:.3e} L2(u) {r['L2U']:.3e} x<->y {r['asym']:.1e} ring {r['ring']:.3e} "
              f"S outflow {r['dS']:+.4f} S err {r['eS']:+.3e}", flush=True)
        assert r["ncmp"] > 100 and r["asym"] < 1e-12
    oT = [np.log2(rs[k]["L2T"] / rs[k + 1]["L2T"]) for k in range(2)]
    oU = [np.log2(rs[k]["L2U"] / rs[k + 1]["L2U"]) for k in range(2)]
    print(f"orders T {oT[0]:.2f}, {oT[1]:.2f}  u {oU[0]:.2f}, {oU[1]:.2f}", flush=True)
    # each published figure to within 0.6 of its last printed digit (the reference
    # rounds, e.g. L2(u) 1.915e-3 is printed 1.92e-3)
    near = lambda x, p, unit: abs(x - p) <= 0.6 * unit
    for r, pT, pU in zip(rs, (3.29e-3, 8.81e-4, 2.26e-4), (7.73e-3, 1.92e-3, 4.57e-4)):
        assert near(r["L2T"], pT, 10 ** (np.floor(np.log10(pT)) - 2))
        assert near(r["L2U"], pU, 10 ** (np.floor(np.log10(pU)) - 2))
    assert all(near(o, p, 0.01) for o, p in zip(oT + oU, (1.90, 1.96, 2.01, 2.07)))
    assert near(rs[2]["dS"], -0.116, 0.001)
    assert near(abs(rs[0]["eS"]), 1.28e-3, 1e-5) and near(abs(rs[2]["eS"]), 9.56e-5, 1e-7)
    assert rs[2]["ring"] < rs[0]["ring"] and rs[2]["ring"] < 5e-3


def test_bjorken_gate_g0_on_gpu(oracle, hip):
    ou = OracleUnit(oracle)
    eos = F.LatticeHRGEOS()
    r = G.run_bjorken(hip, 0.02)
    assert r["nprimfail"] == 0
    assert r["maxDerr"] < 1e-13 and r["maxU"] == 0.0 and r["Espread"] < 1e-13
    r = G.run_bjorken(hip, 0.005)
    T, n, e = G._bjorken_exact(ou, eos, 0.5, -4.2 * 0.5, 0.4, r["tau"])
    assert abs(r["E"] - e) / e < 5e-4 and abs(r["T"] - T) / T < 5e-4 and abs(r["n"] - n) / n < 5e-4
    o = G.run_bjorken(oracle, 0.005)
    assert abs(r["T"] - o["T"]) <= 1e-9 * o["T"] and abs(r["E"] - o["E"]) <= 1e-9 * o["E"]
    # the published temporal orders (README2D.md:273)
    p2, p3 = G.bjorken_order(hip, ou, eos, "ssprk2"), G.bjorken_order(hip, ou, eos, "ssprk3")
    print(f"orders SSPRK2 {p2:.3f} SSPRK3 {p3:.3f}", flush=True)
    assert f"{p2:.2f}" == "2.00" and f"{p3:.2f}" == "2.99"


@pytest.mark.parametrize("zs", [0.05, 0.15, 0.30])
def test_bulk_bjorken_gate_g0b_on_gpu(oracle, hip, zs):
    tauf = 6.0
    h = D.run_bulk(hip, zs, tauf)
    o = D.run_bulk(oracle, zs, tauf)
    ee, Pie, Te = D.bulk_referee(None, None, zs, 15.0, tauf)
    dT = abs(h["T"] - Te) / Te
    dPi = abs(h["Pi"] - Pie) / abs(Pie)
    print(f"zeta/s {zs}: dT {dT:.2e} dPi {dPi:.2e}; vs oracle T {abs(h['T'] - o['T']) / o['T']:.1e} "
          f"Pi {abs(h['Pi'] - o['Pi']) / abs(o['Pi']):.1e}", flush=True)
    assert h["Tspread"] < 1e-12
    assert dT < 1e-3 and dPi < 1e-2 and abs(h["Pi"]) / h["P"] > 0.05
    assert round(100 * dPi, 2) == D.G0B_PUBLISHED_DPI[zs]
    assert abs(h["T"] - o["T"]) <= 1e-9 * o["T"] and abs(h["Pi"] - o["Pi"]) <= 1e-9 * abs(o["Pi"])


def test_viscous_bjorken_gate_g2c_on_gpu(oracle, hip):
    h = D.run_visc_bjorken(hip)
    o = D.run_visc_bjorken(oracle, referee=False)
    print(f"e rel {h['eerr']:.2e} pi^xx rel {h['perr']:.2e}; vs oracle e {abs(h['e'] - o['e']) / o['e']:.1e}",
          flush=True)
    assert h["primfail"] == 0 and h["maxU"] == 0.0 and h["spread"] < 1e-12 and h["shear_res"] < 1e-12
    assert h["eerr"] < 2e-3 and h["perr"] < 2e-2
    assert abs(h["e"] - o["e"]) <= 1e-9 * o["e"] and abs(h["pixx"] - o["pixx"]) <= 1e-9 * abs(o["pixx"])
    ideal = D.run_visc_bjorken(hip, eta_over_s=0.0, referee=False)
    assert h["pieta"] < 0.0 and h["e"] > ideal["e"]
    # G2e: the operator split is first order, the published 1.05 (README2D.md:277)
    errs = [D.run_visc_bjorken(hip, CFLtau=c)["eerr"] for c in (0.016, 0.008, 0.004)]
    p = np.log2(errs[1] / errs[2])
    print(f"G2e order {p:.3f}", flush=True)
    assert f"{p:.2f}" == "1.05"


def test_charge_gates_g3_on_gpu(oracle, hip):
    per = D.run_bump(hip, bc=F.FIVO_BC_PERIODIC)
    dper = abs(per["Q1"] - per["Q0"]) / abs(per["Q0"])
    print(f"periodic charge drift {dper:.2e}", flush=True)
    assert per["r"]["nprimfail"] == 0 and dper < 1e-12
    on = D.run_bump(hip)
    ref = D.run_bump(oracle)
    nsame, ntot = D.upgradient_cells(on)
    print(f"cells with nu^x up the gradient: {nsame} / {ntot} (published 0 of 2104)", flush=True)
    assert nsame == 0 and ntot == 2104
    g, m = on["g"], on["m"]
    sx, sy = g.interior()
    inside = np.zeros((g.Nxtot, g.Nytot), bool)
    inside[sx, sy] = True
    Ui = on["U"][inside.ravel()]
    Vi = D.transpose_rows(on["U"], g, m.layout)[inside.ravel()]
    worst = np.max(np.abs(Ui - Vi) / np.maximum(np.maximum(np.abs(Ui), np.abs(Vi)), 1e-30))
    assert worst < 1e-12
    Ro = ref["U"][inside.ravel()]
    # per-row scale; S^x, S^y on the scale of the energy row (same units): they
    # are zero up to round-off here (uniform T, a radial charm bump exerts no
    # pressure), ~1e-17 on one side and exactly 0 on the other
    scale = np.max(np.abs(Ro), axis=0)
    for k in ("Sx", "Sy"):
        scale[m.layout.idx(k)] = max(scale[m.layout.idx(k)], np.max(np.abs(Ro[:, m.layout.idx("E")])))
    assert np.max(np.abs(Ui - Ro) / scale) <= 1e-9


def test_full_sectors_mirror_symmetry_on_gpu(hip):
    g = F.make_grid2d(32, 32, xmax=8.0, ymax=8.0)
    m = D.full_model()
    s = F.Solver(g, m, lib=hip)
    s.init_from_prims(0.6, **D.symmetric_event(g))
    r = s.run_sim_2d(tau0=0.6, taufinal=1.2, integrator="ssprk3")
    assert r["ok"]
    U = s.get_state()
    sx, sy = g.interior()
    inside = np.zeros((g.Nxtot, g.Nytot), bool)
    inside[sx, sy] = True
    Ui = U[inside.ravel()]
    Vi = D.transpose_rows(U, g, m.layout)[inside.ravel()]
    asym = np.max(np.abs(Ui - Vi), axis=0) / np.maximum(np.max(np.abs(Ui), axis=0), 1e-300)
    print("x<->y asymmetry", float(np.max(asym)), flush=True)
    assert np.max(asym) < 1e-9
