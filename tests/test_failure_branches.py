"""The solver's failure and fallback branches, forced on the oracle (the GPU test
test_gpu_failure_branches.py runs the same inputs on the HIP path against it):

- 2000 pathological cells through sanitize_stage_2d! and state_ok_2d
  (src2d/floors2d.jl:218-245): the sanitiser leaves no NaN or negative value in
  the D and E rows and keeps D >= 0, E >= E_FLOOR; the admissibility predicate
  then fails with inadmissible cells counted (the reference's own note there: the
  predicate must be able to fail on states the sanitiser cannot repair);
- stale warm starts (log T_MIN, the G4 failure mode described at
  src2d/primrec2d.jl:444-463) make the warm solve fail and cons_to_prim_2d!
  retry from the fresh seed: the retry counter fires and the retried cells end
  bit for bit where a cold start ends;
- diffusion currents with nu^tau > J^tau leave no n >= 0 that closes the charge
  row: the staged phi fallback closes the solve (src2d/primrec2d.jl:723-738), phi
  at -PHI_CAP with the hydro rows kept;
- a superluminal cell makes every stage inadmissible: SSPRK3 and SSPRK2 (with its
  MOOD retries) halve the step max_halve times, fail with FIVO_STEP_FAILED and
  restore the entry state (src2d/timestepper2d.jl:74-104, 148-188);
- a CFL far above the stable one halves steps that then succeed.
"""
import ctypes as C

import numpy as np
import pytest

import fivo2d as F
import failure_cases as fc


def counted_update(s, tau):
    cnt = F._abi.Counts()
    s._chk(s.lib.update_primitives(s.ctx, float(tau), 0, C.byref(cnt)), "update_primitives")
    return cnt


def test_pathological_states_sanitize_and_admissibility(oracle):
    s, tau = fc.event(oracle, steps=2)
    Up, kinds = fc.pathological_states(s.get_state(), s.g, s.model.layout)
    assert int(np.sum(kinds >= 0)) == 2000
    s.set_state(Up)
    s.sanitize_stage_2d(tau)
    U = s.get_state()
    ok, nbad = s.state_ok_2d(tau)
    base = U[:, :4]
    assert np.all(np.isfinite(base[:, [0, 3]]))
    assert U[:, 0].min() >= 0.0 and U[:, 3].min() >= F.E_FLOOR
    assert not ok and nbad > 0
    print(f"2000 pathological cells: after sanitize every D, E finite and floored; state_ok false, "
          f"{nbad} inadmissible cells")


def test_stale_warm_start_forces_fresh_seed_retry(oracle):
    s, tau = fc.event(oracle, steps=2)
    cells = fc.stale_warm_starts(s)
    cnt = counted_update(s, tau)
    Wr = s.get_work(("yT", "phi", "ux", "uy", "ok"))
    s2, tau2 = fc.event(oracle, steps=2)
    fc.cold_starts(s2, cells)
    s2.update_primitives_2d(tau2)
    Wc = s2.get_work(("yT", "phi", "ux", "uy", "ok"))
    print(f"stale warm starts on {len(cells)} cells: {cnt.nretry} retries")
    assert cnt.nretry >= len(cells)
    assert np.all(Wr["ok"][cells] == 1)
    for k in Wr:
        assert np.array_equal(Wr[k][cells], Wc[k][cells]), k


def test_unsolvable_charge_row_takes_staged_fallback(oracle):
    s, tau = fc.event(oracle, steps=3)
    cells = fc.charge_row_unsolvable(s, tau)
    cnt = counted_update(s, tau)
    W = s.get_work(("phi", "ok"))
    staged = cells[(W["phi"][cells] == -1e4) & (W["ok"][cells] == 1)]
    print(f"nu^tau = 3 J^tau on {len(cells)} cells: {cnt.nstaged} staged fallbacks, {len(staged)} recovered "
          f"with phi = -PHI_CAP")
    assert len(staged) > 0.5 * len(cells) and cnt.nstaged >= len(staged)


@pytest.mark.parametrize("integrator", ["ssprk3", "ssprk2"])
def test_inadmissible_stage_halves_and_fails(oracle, integrator):
    s, tau = fc.event(oracle)
    fc.superluminal_cell(s)
    U0 = s.get_state()
    r = s.run_sim_2d(tau0=tau, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator=integrator, max_steps=1,
                     max_halve=3)
    print(f"{integrator}: ok {r['ok']} halvings {r['nhalvings']} MOOD retries {r['nmood']}")
    assert not r["ok"] and r["nsteps"] == 0 and r["tau"] == tau
    assert r["nhalvings"] == 4
    assert r["nmood"] == (8 if integrator == "ssprk2" else 0)
    assert np.array_equal(s.get_state(), U0, equal_nan=True)


def test_large_cfl_halves_and_succeeds(oracle):
    s, tau = fc.event(oracle)
    r = s.run_sim_2d(tau0=tau, taufinal=1e9, CFL=0.9, CFLtau=1.0, integrator="ssprk3", max_steps=6)
    print(f"CFL 0.9: {r['nsteps']} steps, {r['nhalvings']} halvings")
    assert r["ok"] and r["nsteps"] == 6 and r["nhalvings"] > 0
