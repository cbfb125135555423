"""GPU checks at the workloads' full sizes.

- per-RHS parity (<= 1e-12 per field) at 512^2 full sectors (the benchmark's
  event) on the initial state and on a state the GPU evolved for 3 steps
  (dissipative fields and warm starts copied into the oracle); on 768^2 and
  1450^2 events, where the cooperative kernel groups fewer lanes per cell,
  primitives to the Newton tolerance and dU to 1e-12 / 1e-9 (the finer grid's
  flux divergence amplifies last-bit EOS differences);
- the 256^2 viscous + charm event (SSPRK2 with MOOD): 3 steps against the oracle,
  same step/halving/MOOD counts, final T and u within 1e-9;
- size-independent properties at 512^2 with every sector on: on a periodic box
  the interior sum of tau J^tau is conserved to round-off; an event symmetric
  under x <-> y stays symmetric (every state row < 1e-9, the all-sector gate's
  bound, test/test_production_allsectors2d.jl:123-125);
- the 64-event ensemble (grids 128^2 ... 384^2): batched runs are bit-identical
  to the events run one by one.
"""
import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic
from helpers import rel_diff
import test_oracle_dissipative_gates as D

pytestmark = pytest.mark.gpu


def full_model(m2=True):
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                            enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                            kappa_coeff=0.1163, pi_clip_factor=1.0, consistent_m2=m2)


def c3_event():
    g = F.make_grid2d(512, 512, xmax=15.0, ymax=15.0)
    return g, ic.hotspot_event(g, seed=12345, nspots=300, sigma=0.4, Tpeak=0.45)


def _rhs_compare(so, sh, tau):
    dUo, co = so.rhs_2d(tau)
    dUh, ch = sh.rhs_2d(tau)
    r = rel_diff(dUh, dUo)
    print("per-field dU rel diff", np.max(r), flush=True)
    assert np.all(np.isfinite(dUh)) and np.max(r) <= 1e-12, r
    assert co["nprimfail"] == ch["nprimfail"] and co["nvacuum"] == ch["nvacuum"]
    assert abs(co["amax"] - ch["amax"]) <= 1e-13 * co["amax"]
    wo, wh = so.get_work(("ok",)), sh.get_work(("ok",))
    assert np.array_equal(wo["ok"], wh["ok"])


def test_rhs_parity_512_full_sectors(oracle, hip):
    g, prims = c3_event()
    m = full_model()
    sh = F.Solver(g, m, lib=hip)
    sh.init_from_prims(0.4, **prims)
    so = F.Solver(g, m, lib=oracle)
    so.set_state(sh.get_state())
    _rhs_compare(so, sh, 0.4)
    # an evolved state: three GPU steps, then the same state, warm starts and
    # histories in both
    r = sh.run_sim_2d(tau0=0.4, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk3", max_steps=3)
    assert r["ok"] and r["nsteps"] == 3
    so.set_state(sh.get_state())
    so.set_work(sh.get_work())
    _rhs_compare(so, sh, r["tau"])


@pytest.mark.parametrize("N,du_bound", [(768, 1e-12), (1450, 1e-9)])
def test_rhs_parity_on_grids_past_2_19_and_2_21_cells(oracle, hip, N, du_bound):
    """The cooperative primitive kernel of a single event takes 32 lanes per queued
    cell up to 2^19 cells, 16 up to 2^21 and 8 beyond; every other test grid is
    below 2^19. On a state the GPU evolved one step (warm starts copied), the
    recovered primitives match the oracle's to the Newton tolerance (T within 1e-12
    relative, u and the charge density n within 1e-11 of their scale; phi itself is
    ill-determined, to ~1e-9, in the near charge-free cells, where it moves n by
    nothing), with the same recovery flags and counts, and
    dU per field within `du_bound`: 1e-12 at 768^2 (16 lanes); at 1450^2 (8 lanes,
    dx = 0.02 fm) the divergence of the face fluxes turns the 2e-13 by which the
    hottest cell's T differs (the device and host math libraries round the EOS
    differently in the last bit) into 1e-10 of the charge row's maximum. The same
    1450^2 numbers come out with 32 lanes, and the 8-lane path is bit-identical to
    the 32-lane one in the 64-event ensemble test below."""
    g = F.make_grid2d(N, N, xmax=15.0, ymax=15.0)
    assert g.Ntot > (1 << 19) and (g.Ntot > (1 << 21)) == (N > 1000)
    m = full_model()
    sh = F.Solver(g, m, lib=hip)
    sh.init_hotspots(0.4, seed=777, nspots=300, sigma=0.4, Tpeak=0.45)   # built on the device
    r = sh.run_sim_2d(tau0=0.4, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk3", max_steps=1)
    assert r["ok"] and r["nsteps"] == 1 and r["nqueued"] > 0
    so = F.Solver(g, m, lib=oracle)
    so.set_state(sh.get_state())
    so.set_work(sh.get_work())
    dUo, co = so.rhs_2d(r["tau"])
    dUh, ch = sh.rhs_2d(r["tau"])
    keys = ("yT", "phi", "ux", "uy", "n", "ok")
    wo, wh = so.get_work(keys), sh.get_work(keys)
    dT = np.max(np.abs(np.expm1(wh["yT"] - wo["yT"])))
    dphi = np.max(np.abs(wh["phi"] - wo["phi"]))
    du = max(np.max(np.abs(wh["ux"] - wo["ux"])), np.max(np.abs(wh["uy"] - wo["uy"])))
    dn = np.max(np.abs(wh["n"] - wo["n"])) / np.max(np.abs(wo["n"]))
    rd = rel_diff(dUh, dUo)
    print(f"N={N}: queued over the step {r['nqueued']}; dT/T {dT:.1e} du {du:.1e} dn {dn:.1e} "
          f"(dphi {dphi:.1e}); "
          f"dU per field max {np.max(rd):.1e}", flush=True)
    assert np.array_equal(wo["ok"], wh["ok"])
    assert co["nprimfail"] == ch["nprimfail"] and co["nvacuum"] == ch["nvacuum"]
    assert dT <= 1e-12 and du <= 1e-11 and dn <= 1e-11
    assert np.all(np.isfinite(dUh)) and np.max(rd) <= du_bound, rd
    so.close()
    sh.close()


def test_c2_event_three_steps_vs_oracle(oracle, hip):
    g = F.make_grid2d(256, 256, xmax=12.0, ymax=12.0)
    prims = ic.hotspot_event(g, seed=12345, nspots=150, sigma=0.5, Tpeak=0.45)
    m = full_model(m2=False)
    out = []
    for lib in (oracle, hip):
        s = F.Solver(g, m, lib=lib)
        s.init_from_prims(0.6, **prims)
        r = s.run_sim_2d(tau0=0.6, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk2", max_steps=3)
        assert r["ok"]
        out.append((r, s.fields_2d()))
    (ro, fo), (rh, fh) = out
    assert (ro["nsteps"], ro["nhalvings"], ro["nmood"], ro["nprimfail"]) == \
        (rh["nsteps"], rh["nhalvings"], rh["nmood"], rh["nprimfail"])
    for k in ("T", "ux", "uy"):
        d = np.max(np.abs(fh[k] - fo[k])) / np.max(np.abs(fo[k]))
        print(f"{k}: {d:.2e}", flush=True)
        assert d <= 1e-9, (k, d)


def test_512_periodic_charge_and_mirror_symmetry(hip):
    g = F.make_grid2d(512, 512, xmax=15.0, ymax=15.0)
    m = full_model()
    s = F.Solver(g, m, lib=hip)
    s.init_from_prims(0.4, **D.symmetric_event(g, seed=12345, nspots=300, sigma=0.4), bc=F.FIVO_BC_PERIODIC)
    sx, sy = g.interior()
    Q0 = s.get_state()[:, 0].reshape(g.Nxtot, g.Nytot)[sx, sy].sum()
    r = s.run_sim_2d(tau0=0.4, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk3", max_steps=5,
                     bc=F.FIVO_BC_PERIODIC)
    assert r["ok"] and r["nsteps"] == 5
    U = s.get_state()
    Q1 = U[:, 0].reshape(g.Nxtot, g.Nytot)[sx, sy].sum()
    print("charge drift", abs(Q1 - Q0) / Q0, flush=True)
    assert abs(Q1 - Q0) / Q0 < 1e-12
    inside = np.zeros((g.Nxtot, g.Nytot), bool)
    inside[sx, sy] = True
    Ui = U[inside.ravel()]
    Vi = D.transpose_rows(U, g, m.layout)[inside.ravel()]
    asym = np.max(np.abs(Ui - Vi), axis=0) / np.maximum(np.max(np.abs(Ui), axis=0), 1e-300)
    print("x<->y asymmetry", float(np.max(asym)), flush=True)
    assert np.max(asym) < 1e-9


def test_64_event_ensemble_bit_identical_to_one_by_one(hip):
    rng = np.random.Generator(np.random.Philox(2024))
    sizes = rng.choice((128, 192, 256, 320, 384), size=64)
    nsp = rng.integers(50, 301, size=64)
    m = full_model()
    grids = [F.make_grid2d(int(n), int(n), xmax=15.0, ymax=15.0) for n in sizes]
    leg = dict(taufinal=1e9, CFL=0.2, CFLtau=0.05, max_steps=2)
    x = F.Ensemble(grids, m, lib=hip)
    for k, s in enumerate(x.events):
        s.init_hotspots(0.4, seed=2024000 + k, nspots=int(nsp[k]), sigma=0.4, Tpeak=0.45)
    res = x.run_sim_2d(tau0=0.4, **leg)
    Ub = [s.get_state() for s in x.events]
    x.close()
    y = F.Ensemble(grids, m, lib=hip)
    for k, s in enumerate(y.events):
        s.init_hotspots(0.4, seed=2024000 + k, nspots=int(nsp[k]), sigma=0.4, Tpeak=0.45)
        r = s.run_sim_2d(tau0=0.4, integrator="ssprk3", **leg)
        assert r["nsteps"] == res[k]["nsteps"] and r["tau"] == res[k]["tau"]
        assert np.array_equal(s.get_state(), Ub[k]), k
    y.close()


@pytest.mark.parametrize("name,tauf", [("c1", 3.0), ("c2", 8.0)])
def test_full_length_run_vs_oracle(oracle, hip, name, tauf):
    """The benchmark configurations 1 (ideal Gubser 128^2, tau 1 -> 3, SSPRK3) and 2
    (viscous + charm 256^2 hot-spot event, tau 0.6 -> 8, SSPRK2 with MOOD) run to
    their final times on both paths: the same steps and counters, no recovery-flag
    flips, final T and u within 1e-9."""
    import ctypes as C
    import sys
    from conftest import REPO
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import bench
    cfg = bench.CONFIGS[name]
    oracle.dll.orc_set_threads.restype = C.c_int
    oracle.dll.orc_set_threads(16)
    out = []
    for lib in (hip, oracle):
        g, m, prims = bench.make_inputs(cfg)
        s = F.Solver(g, m, lib=lib)
        s.init_from_prims(cfg["tau0"], **prims)
        r = s.run_sim_2d(tau0=cfg["tau0"], taufinal=tauf, CFL=cfg["cfl"], CFLtau=cfg["cfltau"],
                         integrator=cfg["integrator"])
        assert r["ok"]
        out.append((r, s.fields_2d(), s.get_work(("ok",))["ok"]))
        s.close()
    (rh, fh, okh), (ro, fo, oko) = out
    for k in ("nsteps", "nhalvings", "nmood", "nprimfail", "tau"):
        assert rh[k] == ro[k], k
    assert np.array_equal(okh, oko)
    dT = np.max(np.abs(fh["T"] - fo["T"]) / fo["T"])
    du = max(np.max(np.abs(fh[k] - fo[k])) for k in ("ux", "uy"))
    print(f"{name} to tau {tauf}: {rh['nsteps']} steps, max dT/T {dT:.2e}, max du {du:.2e}", flush=True)
    assert dT <= 1e-9 and du <= 1e-9


def test_full_length_c3_vs_oracle_and_conditioning(hip, oracle):
    """Configuration 3 (512^2 full sectors, tau 0.4 -> 10, SSPRK3) to its final time
    on the HIP path and on the oracle from the same input, and on the HIP path again
    with the initial conserved rows perturbed by 1e-13 (relative). HIP and oracle take
    the same 802 steps with the same counters and no recovery-flag flips; their final
    T differs by ~3e-7 (measured 2.9e-7; 5.8e-7 with multiply-add fusion off on the
    device, so the device's contraction is not its source: the two sides' exp and log
    round differently in the last bit). That distance is the reference algorithm's
    conditioning on this lumpy event, not a discrepancy: the 1e-13 perturbation alone
    moves the final T by ~7e-2 with recovery-flag flips, so no implementation that is
    not bit-identical to the reference can agree with it to 1e-9 at this length. The
    test asserts the HIP-vs-oracle distance stays below 1e-5 and below 1e-3 of the
    perturbation's."""
    import ctypes as C
    import sys
    import time
    from conftest import REPO
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import bench
    cfg = bench.CONFIGS["c3"]
    oracle.dll.orc_set_threads.restype = C.c_int
    oracle.dll.orc_set_threads.argtypes = [C.c_int]
    oracle.dll.orc_set_threads(16)
    runs = {}
    for tag, lib, eps in (("hip", hip, 0.0), ("hip_perturbed", hip, 1e-13), ("oracle", oracle, 0.0)):
        g, m, prims = bench.make_inputs(cfg)
        s = F.Solver(g, m, lib=lib)
        s.init_from_prims(cfg["tau0"], **prims)
        if eps:
            U = s.get_state()
            U[:, :4] *= 1 + eps * np.random.default_rng(1).uniform(-1, 1, size=(U.shape[0], 4))
            s.set_state(U)
        t0 = time.perf_counter()
        r = s.run_sim_2d(tau0=cfg["tau0"], taufinal=10.0, CFL=cfg["cfl"], CFLtau=cfg["cfltau"],
                         integrator=cfg["integrator"], dump_dt=2.0,
                         on_dump=lambda tau, _s, tag=tag: print(f"  {tag} tau {tau:.2f}", flush=True))
        assert r["ok"]
        w = s.get_work(("yT", "ok"))
        runs[tag] = (r, np.exp(w["yT"]), w["ok"])
        print(f"{tag}: {r['nsteps']} steps in {time.perf_counter() - t0:.1f} s", flush=True)
        s.close()

    def dist(a, b):
        (_, Ta, oka), (_, Tb, okb) = runs[a], runs[b]
        both = (oka == 1) & (okb == 1)
        return float(np.max(np.abs(Ta - Tb)[both] / Tb[both])), int(np.sum(oka != okb))

    (rh, _, _), (ro, _, _) = runs["hip"], runs["oracle"]
    for k in ("nsteps", "nhalvings", "nmood", "nprimfail", "tau"):
        assert rh[k] == ro[k], k
    d_ho, f_ho = dist("hip", "oracle")
    d_pp, f_pp = dist("hip_perturbed", "hip")
    print(f"c3 to tau 10: HIP vs oracle max dT/T {d_ho:.2e} ({f_ho} flips); HIP with the input perturbed by "
          f"1e-13 vs HIP: {d_pp:.2e} ({f_pp} flips)", flush=True)
    assert f_ho == 0
    assert d_pp > 1e-6
    assert d_ho < 1e-5 and d_ho < 1e-3 * d_pp
