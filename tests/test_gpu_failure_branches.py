"""The forced failure and fallback branches of test_failure_branches.py on the HIP
path, against the oracle on the same inputs: sanitiser and admissibility on 2000
pathological cells, the fresh-seed retry, the staged phi fallback, failed and
halved steps, and SSPRK2's MOOD stage retries on the 512^2 full-sector event.
Each asserts the branch fired (its counter above zero) and that the HIP path took
it exactly as the oracle did."""
import ctypes as C

import numpy as np
import pytest

import fivo2d as F
import failure_cases as fc
from fivo2d import ic
from test_failure_branches import counted_update

pytestmark = pytest.mark.gpu


def _same(a, b, tol):
    """Same NaN/Inf pattern and finite values within tol of the column scale."""
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb)
    assert np.array_equal(a[~fa], b[~fb], equal_nan=True)
    scale = np.maximum(np.max(np.where(fb, np.abs(b), 0.0), axis=0), 1e-300)
    d = np.max(np.where(fb, np.abs(np.where(fa, a, 0) - np.where(fb, b, 0)), 0.0), axis=0) / scale
    assert np.max(d) <= tol, d
    return float(np.max(d))


def test_gpu_pathological_states(hip, oracle):
    out = []
    for lib in (hip, oracle):
        s, tau = fc.event(lib, steps=2)
        out.append(s.get_state())
        s.close()
    Up, _ = fc.pathological_states(out[1], fc.grid(), fc.full_model().layout)
    res = []
    for lib in (hip, oracle):
        s, tau = fc.event(lib, steps=2)
        s.set_state(Up)
        s.sanitize_stage_2d(tau)
        U = s.get_state()
        ok, nbad = s.state_ok_2d(tau)
        res.append((U, ok, nbad))
        s.close()
    (Uh, okh, nh), (Uo, oko, no) = res
    d = _same(Uh, Uo, 1e-12)
    print(f"pathological cells: HIP vs oracle after sanitize {d:.1e}; inadmissible {nh} vs {no}", flush=True)
    assert okh == oko and not okh and nh == no > 0


def test_gpu_fresh_seed_retry(hip, oracle):
    res = []
    for lib in (hip, oracle):
        s, tau = fc.event(lib, steps=2)
        cells = fc.stale_warm_starts(s)
        cnt = counted_update(s, tau)
        res.append((cells, cnt.nretry, s.get_work(("yT", "phi", "ux", "uy", "ok"))))
        s.close()
    (ch, rh, Wh), (co, ro, Wo) = res
    print(f"stale warm starts: {len(co)} cells, retries HIP {rh} oracle {ro}", flush=True)
    assert np.array_equal(ch, co)
    assert rh == ro >= len(co)
    assert np.array_equal(Wh["ok"], Wo["ok"])
    for k in ("yT", "ux", "uy"):
        assert np.max(np.abs(Wh[k][co] - Wo[k][co])) <= 1e-12 * max(1.0, np.max(np.abs(Wo[k][co]))), k


def test_gpu_staged_phi_fallback(hip, oracle):
    res = []
    for lib in (hip, oracle):
        s, tau = fc.event(lib, steps=3)
        cells = fc.charge_row_unsolvable(s, tau)
        cnt = counted_update(s, tau)
        res.append((cells, cnt.nstaged, s.get_work(("phi", "ok", "yT"))))
        s.close()
    (ch, sh, Wh), (co, so, Wo) = res
    print(f"unsolvable charge rows: {len(co)} cells, staged fallbacks HIP {sh} oracle {so}", flush=True)
    assert sh == so > 0
    assert np.array_equal(Wh["ok"], Wo["ok"]) and np.array_equal(Wh["phi"][co], Wo["phi"][co])


@pytest.mark.parametrize("integrator", ["ssprk3", "ssprk2"])
def test_gpu_failed_step(hip, oracle, integrator):
    res = []
    for lib in (hip, oracle):
        s, tau = fc.event(lib)
        fc.superluminal_cell(s)
        U0 = s.get_state()
        r = s.run_sim_2d(tau0=tau, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator=integrator, max_steps=1,
                         max_halve=3)
        assert np.array_equal(s.get_state(), U0, equal_nan=True)
        res.append(r)
        s.close()
    rh, ro = res
    print(f"{integrator}: HIP ok {rh['ok']} halvings {rh['nhalvings']} MOOD {rh['nmood']}; oracle "
          f"{ro['nhalvings']} / {ro['nmood']}", flush=True)
    assert not rh["ok"] and not ro["ok"]
    for k in ("nsteps", "nhalvings", "nmood"):
        assert rh[k] == ro[k], k
    assert rh["nhalvings"] == 4


def test_gpu_halved_steps(hip, oracle):
    res = []
    for lib in (hip, oracle):
        s, tau = fc.event(lib)
        r = s.run_sim_2d(tau0=tau, taufinal=1e9, CFL=0.9, CFLtau=1.0, integrator="ssprk3", max_steps=6)
        res.append((r, s.get_work(("yT", "ok"))))
        s.close()
    (rh, Wh), (ro, Wo) = res
    print(f"CFL 0.9: HIP {rh['nsteps']} steps {rh['nhalvings']} halvings; oracle {ro['nsteps']} / "
          f"{ro['nhalvings']}", flush=True)
    for k in ("ok", "nsteps", "nhalvings"):
        assert rh[k] == ro[k], k
    assert abs(rh["tau"] - ro["tau"]) <= 1e-12 * ro["tau"]      # dt follows the signal speeds' round-off
    assert rh["nhalvings"] > 0
    both = (Wh["ok"] == 1) & (Wo["ok"] == 1)
    assert np.max(np.abs(np.exp(Wh["yT"][both]) - np.exp(Wo["yT"][both])) / np.exp(Wo["yT"][both])) <= 1e-9


def test_gpu_mood_at_512(hip, oracle):
    """SSPRK2 with MOOD on the 512^2 full-sector event at a CFL where stages are
    retried: the same retries, halvings and counters as the oracle, T within 1e-9."""
    g = F.make_grid2d(512, 512, xmax=15.0, ymax=15.0)
    m = fc.full_model()
    prims = ic.hotspot_event(g, seed=12345, nspots=300, sigma=0.4, Tpeak=0.45)
    oracle.dll.orc_set_threads.restype = C.c_int
    oracle.dll.orc_set_threads.argtypes = [C.c_int]
    oracle.dll.orc_set_threads(16)
    res = []
    for lib in (hip, oracle):
        s = F.Solver(g, m, lib=lib)
        s.init_from_prims(0.4, **prims)
        r = s.run_sim_2d(tau0=0.4, taufinal=1e9, CFL=1.0, CFLtau=1.0, integrator="ssprk2", max_steps=4)
        res.append((r, s.get_work(("yT", "ok", "ux", "uy"))))
        s.close()
    (rh, Wh), (ro, Wo) = res
    print(f"512^2 SSPRK2 CFL 1.0: HIP steps {rh['nsteps']} MOOD {rh['nmood']} halvings {rh['nhalvings']}; "
          f"oracle {ro['nsteps']} / {ro['nmood']} / {ro['nhalvings']}", flush=True)
    for k in ("ok", "nsteps", "nmood", "nhalvings", "nprimfail"):
        assert rh[k] == ro[k], k
    assert abs(rh["tau"] - ro["tau"]) <= 1e-12 * ro["tau"]
    assert rh["nmood"] > 0
    assert np.array_equal(Wh["ok"], Wo["ok"])
    both = Wh["ok"] == 1
    assert np.max(np.abs(np.exp(Wh["yT"][both]) - np.exp(Wo["yT"][both])) / np.exp(Wo["yT"][both])) <= 1e-9
