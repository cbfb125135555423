"""The benchmark's learned paths on the GPU, against the plain path on the same
inputs and steps:

- the 512^2 full-sector event (the headline workload) through the ensemble path
  with the learned warm start (kb_seed, learned-mode kb_coop) against the plain
  one-event run, over the benchmark's 5 warm-up and 20 timed steps;
- 32 events of 384^2 (the learned-seed configuration) as a learned ensemble
  against the same ensemble without the network (which is bit-identical to the
  events run one by one, test_gpu_batch.py).

Final T within 1e-9 of the plain path on every cell whose recovery converged in
both runs, no recovery-flag flips, learned starts used and the fallback rate
bounded.
"""
import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic
from test_gpu_mlp import full_model, trained_mlp

pytestmark = pytest.mark.gpu


def _T(s):
    w = s.get_work(("yT", "ok"))
    return np.exp(w["yT"]), w["ok"]


def test_batched_learned_512_matches_plain(hip):
    m = full_model()
    net = trained_mlp(hip, m)
    net.train(epochs=6, batch=4096, lr=3e-3)
    g = F.make_grid2d(512, 512, xmax=15.0, ymax=15.0)
    prims = ic.hotspot_event(g, seed=12345, nspots=300, sigma=0.4, Tpeak=0.45)
    leg = dict(taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk3")
    s = F.Solver(g, m, lib=hip)
    s.init_from_prims(0.4, **prims)
    r = s.run_sim_2d(tau0=0.4, max_steps=5, **leg)
    rp = s.run_sim_2d(tau0=r["tau"], max_steps=20, reset_history=False, **leg)
    Tp, okp = _T(s)
    s.close()
    x = F.Ensemble([g], m, lib=hip)
    x.events[0].init_from_prims(0.4, **prims)
    x.set_learned_seed(net)
    x.autotune(0.4)
    r0 = x.run_sim_2d(tau0=0.4, max_steps=5, **{k: v for k, v in leg.items() if k != "integrator"})
    rl = x.run_sim_2d(tau0=[r0[0]["tau"]], max_steps=20, reset_history=False,
                      **{k: v for k, v in leg.items() if k != "integrator"})[0]
    Tl, okl = _T(x.events[0])
    nl = r0[0]["nlearned"] + rl["nlearned"]
    fb = r0[0]["nlearned_fallback"] + rl["nlearned_fallback"]
    x.close()
    net.close()
    both = (okp == 1) & (okl == 1)
    dT = float(np.max(np.where(both, np.abs(Tl - Tp) / Tp, 0.0)))
    print(f"512^2 batched+learned vs plain: dT/T {dT:.2e}, flips {int(np.sum(okp != okl))}, learned starts {nl}, "
          f"fallbacks {fb}, Newton iterations {rl['newton_iters']} vs {rp['newton_iters']}", flush=True)
    assert rl["ok"] and rl["nsteps"] == 20 == rp["nsteps"]
    assert int(np.sum(okp != okl)) == 0
    assert dT <= 1e-9
    assert nl > 0 and fb <= 0.15 * nl
    assert rl["newton_iters"] <= rp["newton_iters"]


def test_learned_ensemble_32x384_matches_plain(hip):
    m = full_model()
    net = trained_mlp(hip, m)
    net.train(epochs=6, batch=4096, lr=3e-3)
    rng = np.random.default_rng(4048)
    spots = rng.integers(50, 301, size=32)
    g = F.make_grid2d(384, 384, xmax=15.0, ymax=15.0)
    leg = dict(taufinal=1e9, CFL=0.2, CFLtau=0.05)
    out = {}
    for learned in (False, True):
        x = F.Ensemble([g] * 32, m, lib=hip)
        for k, s in enumerate(x.events):
            s.init_hotspots(0.4, seed=4048000 + k, nspots=int(spots[k]), sigma=0.4, Tpeak=0.45)
        if learned:
            x.set_learned_seed(net)
        r0 = x.run_sim_2d(tau0=0.4, max_steps=3, **leg)
        r1 = x.run_sim_2d(tau0=[r["tau"] for r in r0], max_steps=5, reset_history=False, **leg)
        assert all(r["ok"] and r["nsteps"] == 5 for r in r1)
        out[learned] = ([_T(s) for s in x.events], r0, r1)
        x.close()
    net.close()
    (P, _, p1), (L, l0, l1) = out[False], out[True]
    worst, flips = 0.0, 0
    for (Tp, okp), (Tl, okl) in zip(P, L):
        both = (okp == 1) & (okl == 1)
        worst = max(worst, float(np.max(np.where(both, np.abs(Tl - Tp) / Tp, 0.0))))
        flips += int(np.sum(okp != okl))
    nl = sum(r["nlearned"] for r in l0 + l1)
    fb = sum(r["nlearned_fallback"] for r in l0 + l1)
    ip = sum(r["newton_iters"] for r in p1)
    il = sum(r["newton_iters"] for r in l1)
    print(f"32 x 384^2 learned vs plain: dT/T {worst:.2e}, flips {flips}, learned starts {nl}, fallbacks {fb}, "
          f"Newton iterations {il} vs {ip}", flush=True)
    assert flips == 0 and worst <= 1e-9
    assert nl > 0 and fb <= 0.15 * nl
    assert il <= ip


def test_learned_seed_reaches_the_cold_seed_root_on_a_sensitive_event(hip):
    """A configuration-4 event (192^2 on [-15, 15] fm, 90 spots, seed 2024057) on
    which a learned run can part from the plain run by O(0.1) within 13 steps.
    Along the plain trajectory, the state recovered from a full cold start with
    the learned seed equals the cold seed's recovery on every cell (same root, to
    the solver's tolerance, no flag flips); and 1e-13 perturbations of the initial
    state alone make the plain run part from itself by more than 1e-9 with flag
    flips over the same steps. So a learned run's departure on such an event is a
    round-off-level difference amplified by the reference algorithm, not a
    different root."""
    m = full_model()
    net = trained_mlp(hip, m)
    net.train(epochs=6, batch=4096, lr=3e-3)
    g = F.make_grid2d(192, 192, xmax=15.0, ymax=15.0)
    leg = dict(taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk3")

    def event():
        s = F.Solver(g, m, lib=hip)
        s.init_hotspots(0.4, seed=2024057, nspots=90, sigma=0.4, Tpeak=0.45)
        return s

    P = event()
    U0 = P.get_state()
    keys = ("x0_yT", "x0_phi", "x0_ux", "x0_uy")
    worst_T, worst_u, flips_root, tau = 0.0, 0.0, 0, 0.4
    twins = [F.Solver(g, m, lib=hip), F.Solver(g, m, lib=hip)]
    twins[1].set_learned_seed(net)
    for st in range(13):
        r = P.run_sim_2d(tau0=tau, max_steps=1, reset_history=(st == 0), **leg)
        tau = r["tau"]
        U, W = P.get_state(), P.get_work()
        rec = []
        for t in twins:
            t.set_state(U)
            t.set_work(dict(W, **{k: np.full_like(W[k], np.nan) for k in keys}))
            t.update_primitives_2d(tau)
            rec.append(t.get_work(("yT", "ok", "ux", "uy")))
        ok2 = (rec[0]["ok"] == 1) & (rec[1]["ok"] == 1)
        Ta = np.exp(rec[0]["yT"])
        worst_T = max(worst_T, float(np.max(np.where(ok2, np.abs(np.exp(rec[1]["yT"]) - Ta) / Ta, 0.0))))
        worst_u = max(worst_u, float(np.max(np.where(ok2, np.abs(rec[1]["ux"] - rec[0]["ux"]) +
                                                     np.abs(rec[1]["uy"] - rec[0]["uy"]), 0.0))))
        flips_root += int(np.sum(rec[0]["ok"] != rec[1]["ok"]))
    Tref, okref = _T(P)
    for t in twins:
        t.close()
    P.close()
    net.close()
    Q = event()
    sign = np.random.default_rng(101).choice([-1.0, 1.0], size=U0.shape)
    Q.set_state(U0 * (1.0 + 1e-13 * sign))
    Q.run_sim_2d(tau0=0.4, max_steps=13, **leg)
    Tq, okq = _T(Q)
    Q.close()
    both = (okq == 1) & (okref == 1)
    d_pert = float(np.max(np.where(both, np.abs(Tq - Tref) / Tref, 0.0)))
    flips_pert = int(np.sum(okq != okref))
    print(f"root check over 13 steps: dT/T {worst_T:.2e}, du {worst_u:.2e}, flips {flips_root}; "
          f"1e-13-perturbed plain run: dT/T {d_pert:.2e}, flips {flips_pert}", flush=True)
    assert worst_T < 1e-11 and worst_u < 1e-11 and flips_root == 0
    assert d_pert > 1e-9 and flips_pert > 0
