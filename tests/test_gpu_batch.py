"""GPU tests of ensembles (fivo_batch_*): every event of an ensemble must end
where the one-event run_sim_2d! (fivo_run) takes it alone, and where the CPU
oracle takes it (final T and u within 1e-9).

The ensemble runs the same device bodies as the one-event kernels, on the same
per-event arrays, with the same per-event dt / halving decisions made on the
device: step counts, halvings, final times and states must agree bit for bit
(multiply-adds fuse only within one source expression, so a body rounds the
same way in both kernels)."""
import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic
from helpers import rel_diff

pytestmark = pytest.mark.gpu


def full_model():
    """The full-sector model of the 512^2 workload (shear, bulk, diffusion, charm m2)."""
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                            enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                            kappa_coeff=0.1163, pi_clip_factor=1.0, consistent_m2=True)


# uneven grids and hot-spot counts, like the 64-event ensemble in miniature
EVENTS = [((40, 40), 7, 60), ((56, 48), 11, 120), ((32, 64), 13, 40)]


def events(L=12.0):
    out = []
    for (nx, ny), seed, nspots in EVENTS:
        g = F.make_grid2d(nx, ny, xmax=L, ymax=L * ny / nx)
        out.append((g, ic.hotspot_event(g, seed=seed, nspots=nspots, sigma=0.5, Tpeak=0.45)))
    return out


RUN = dict(tau0=0.6, taufinal=0.9, CFL=0.2, CFLtau=0.05)


def run_alone(lib, g, m, prims):
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(RUN["tau0"], **prims)
    r = s.run_sim_2d(integrator="ssprk3", **RUN)
    return r, s.get_state(), s.get_work(("yT", "ux", "uy"))


def run_ensemble(hip, m, evs, tile=None, autotune=False):
    ens = F.Ensemble([g for g, _ in evs], m, lib=hip)
    for s, (g, prims) in zip(ens.events, evs):
        s.init_from_prims(RUN["tau0"], **prims)
    if autotune:
        tile, ms = ens.autotune(RUN["tau0"])
        assert tile in F._abi.BATCH_TILES and all(v > 0 for v in ms.values())
    elif tile is not None:
        ens.set_tile(tile)
    res = ens.run_sim_2d(**RUN)
    out = [(r, s.get_state(), s.get_work(("yT", "ux", "uy"))) for r, s in zip(res, ens.events)]
    ens.close()
    return out


@pytest.mark.parametrize("tile", [(7, 8), (15, 16)])
def test_ensemble_equals_one_event_runs(hip, tile):
    m = full_model()
    evs = events()
    alone = [run_alone(hip, g, m, prims) for g, prims in evs]
    together = run_ensemble(hip, m, evs, tile=tile)
    for k, ((ra, Ua, Wa), (rb, Ub, Wb)) in enumerate(zip(alone, together)):
        print(f"event {k}: alone nsteps={ra['nsteps']} tau={ra['tau']:.6f} halvings={ra['nhalvings']}; "
              f"ensemble nsteps={rb['nsteps']} tau={rb['tau']:.6f}", flush=True)
        assert ra["ok"] and rb["ok"]
        for key in ("nsteps", "nhalvings", "nprimfail", "nvacuum", "nrhs"):
            assert ra[key] == rb[key], (k, key, ra[key], rb[key])
        assert ra["tau"] == rb["tau"]
        r = rel_diff(Ub, Ua)
        nbits = int(np.sum(Ua != Ub))
        print(f"    state: {nbits} of {Ua.size} values differ, max rel diff per field {np.max(r):.2e}", flush=True)
        assert nbits == 0, (k, r)
        for f in ("yT", "ux", "uy"):
            assert np.array_equal(Wa[f], Wb[f]), (k, f)
        assert ra["Q1"] == rb["Q1"] and ra["max_shear_res"] == rb["max_shear_res"]


def test_face_overlap_equals_serial_order(hip):
    """Face tiles that read no queued cell run on a second stream beside the
    cooperative primitive kernel, the rest after it. Kernel timing serialises the
    two streams, so a timed run gives the serial order: the bits must not change."""
    m = full_model()
    for k, (g, prims) in enumerate(events()):
        ra, Ua, Wa = run_alone(hip, g, m, prims)
        s = F.Solver(g, m, lib=hip)
        assert hip.enable_timing(s.ctx, 1) == 0
        s.init_from_prims(RUN["tau0"], **prims)
        rb = s.run_sim_2d(integrator="ssprk3", **RUN)
        print(f"event {k}: {ra['nqueued']} cells queued over {ra['nsteps']} steps", flush=True)
        assert ra["nqueued"] > 0 and ra["nqueued"] == rb["nqueued"] and ra["tau"] == rb["tau"]
        assert np.array_equal(Ua, s.get_state()), k
        Wb = s.get_work(("yT", "ux", "uy"))
        for f in ("yT", "ux", "uy"):
            assert np.array_equal(Wa[f], Wb[f]), (k, f)
        s.close()


def test_ensemble_autotuned_matches_oracle(oracle, hip):
    """Autotuned ensemble vs the CPU oracle per event: final T and u within 1e-9."""
    m = full_model()
    evs = events()
    together = run_ensemble(hip, m, evs, autotune=True)
    g0 = [g for g, _ in evs]
    for k, ((g, prims), (rb, Ub, Wb)) in enumerate(zip(evs, together)):
        ro, Uo, Wo = run_alone(oracle, g, m, prims)
        assert ro["ok"] and rb["ok"]
        assert ro["nsteps"] == rb["nsteps"] and ro["nhalvings"] == rb["nhalvings"], (k, ro, rb)
        sx, sy = g.interior()
        T = lambda W: np.exp(W["yT"].reshape(g.Nxtot, g.Nytot)[sx, sy])
        u = lambda W, f: W[f].reshape(g.Nxtot, g.Nytot)[sx, sy]
        dT = np.max(np.abs(T(Wb) - T(Wo)) / T(Wo))
        du = max(np.max(np.abs(u(Wb, f) - u(Wo, f))) for f in ("ux", "uy"))
        print(f"event {k} ({g.Nx}x{g.Ny}): nsteps {rb['nsteps']} dT/T {dT:.2e} du {du:.2e}", flush=True)
        assert dT <= 1e-9 and du <= 1e-9, (k, dT, du)
        r = rel_diff(Ub, Uo)
        assert np.max(r) <= 1e-7, (k, r)


def test_device_hotspots_match_host_generator(hip):
    """fivo_init_hotspots builds the same event as ic.hotspot_event + set_cell!
    (only the last bits of exp / cbrt / log may differ)."""
    m = full_model()
    g = F.make_grid2d(48, 40, xmax=12.0, ymax=10.0)
    prims = ic.hotspot_event(g, seed=5, nspots=80, sigma=0.5, Tpeak=0.45)
    a = F.Solver(g, m, lib=hip)
    nbad_host = a.init_from_prims(0.6, **prims)
    b = F.Solver(g, m, lib=hip)
    nbad = b.init_hotspots(0.6, seed=5, nspots=80, sigma=0.5, Tpeak=0.45)
    assert nbad == nbad_host          # cells left at zero by set_cell! (T at the floor)
    r = rel_diff(b.get_state(), a.get_state())
    print("device vs host event: max rel diff per field", np.max(r), flush=True)
    assert np.max(r) <= 1e-12, r


def test_ensemble_continues_from_per_event_times(hip):
    """Two ensemble legs (the second from each event's own stop time, histories
    kept) equal the same two legs of each event alone, bit for bit."""
    m = full_model()
    evs = events()
    ens = F.Ensemble([g for g, _ in evs], m, lib=hip)
    for s, (g, prims) in zip(ens.events, evs):
        s.init_from_prims(RUN["tau0"], **prims)
    leg = dict(taufinal=1e9, CFL=RUN["CFL"], CFLtau=RUN["CFLtau"])
    r1 = ens.run_sim_2d(tau0=RUN["tau0"], max_steps=4, **leg)
    taus = [r["tau"] for r in r1]
    assert len(set(taus)) > 1                       # the events stopped at different times
    r2 = ens.run_sim_2d(tau0=taus, max_steps=3, reset_history=False, **leg)
    for k, (g, prims) in enumerate(evs):
        s = F.Solver(g, m, lib=hip)
        s.init_from_prims(RUN["tau0"], **prims)
        a1 = s.run_sim_2d(tau0=RUN["tau0"], max_steps=4, integrator="ssprk3", **leg)
        a2 = s.run_sim_2d(tau0=a1["tau"], max_steps=3, integrator="ssprk3", reset_history=False, **leg)
        assert a1["tau"] == r1[k]["tau"] and a2["tau"] == r2[k]["tau"] and a2["nsteps"] == r2[k]["nsteps"] == 3
        assert np.array_equal(s.get_state(), ens.events[k].get_state()), k
    ens.close()


def test_autotune_leaves_the_solver_state_unchanged(hip):
    """Tuning runs a primitive pass to time the face tiles on real states; the warm
    starts that pass moves are restored, so a run after tuning is the run without."""
    m = full_model()
    evs = events()
    ens = F.Ensemble([g for g, _ in evs], m, lib=hip)
    for s, (g, prims) in zip(ens.events, evs):
        s.init_from_prims(RUN["tau0"], **prims)
    keys = ("x0_yT", "x0_phi", "x0_ux", "x0_uy")
    before = [s.get_work(keys) for s in ens.events]
    U0 = [s.get_state() for s in ens.events]
    ens.autotune(RUN["tau0"])
    for k, s in enumerate(ens.events):
        after = s.get_work(keys)
        for f in keys:
            assert np.array_equal(before[k][f], after[f], equal_nan=True), (k, f)
        assert np.array_equal(U0[k], s.get_state()), k
    ens.close()


def test_ensemble_ssprk2_with_mood_equals_one_event_runs(hip):
    """SSPRK2 ensembles (step_ssprk2_2d!: MOOD stage retries on the expanded bad mask,
    then halving) against each event run alone: the sharp event of the MOOD tests,
    which retries stages, beside two milder events; the same counts (MOOD retries
    included) and the same final state, bit for bit."""
    # the model and the sharp event of the MOOD test (test_oracle_mood.mood_run)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                         enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                         kappa_coeff=0.1163, pi_clip_factor=1.0)
    evs = []
    g = F.make_grid2d(64, 64, xmax=8.0, ymax=8.0)
    evs.append((g, ic.hotspot_event(g, seed=3, nspots=60, sigma=0.25, Tpeak=0.5)))
    evs += events()[:2]
    run = dict(tau0=0.6, taufinal=1.6, CFL=0.6, CFLtau=0.05)
    ens = F.Ensemble([g for g, _ in evs], m, lib=hip)
    for s, (g, prims) in zip(ens.events, evs):
        s.init_from_prims(run["tau0"], **prims)
    together = ens.run_sim_2d(integrator="ssprk2", max_mood=2, **run)
    states = [s.get_state() for s in ens.events]
    ens.close()
    total_mood = 0
    for k, (g, prims) in enumerate(evs):
        s = F.Solver(g, m, lib=hip)
        s.init_from_prims(run["tau0"], **prims)
        ra = s.run_sim_2d(integrator="ssprk2", max_mood=2, **run)
        rb = together[k]
        print(f"event {k}: alone steps {ra['nsteps']} MOOD {ra['nmood']} halvings {ra['nhalvings']}; "
              f"ensemble steps {rb['nsteps']} MOOD {rb['nmood']} halvings {rb['nhalvings']}", flush=True)
        for key in ("ok", "nsteps", "nhalvings", "nmood", "nprimfail", "nvacuum", "nrhs", "tau"):
            assert ra[key] == rb[key], (k, key, ra[key], rb[key])
        assert np.array_equal(s.get_state(), states[k]), k
        total_mood += ra["nmood"]
        s.close()
    assert total_mood >= 1
