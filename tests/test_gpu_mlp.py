"""GPU tests of the learned Newton warm start.

- the BF16-MFMA forward pass and loss gradient against the CPU reference
  (oracle/oracle_mlp.cpp, same BF16 rounding points, double accumulation);
- training on pairs the solver itself produced lowers the loss;
- runs with the learned seed end where the cold-seed runs end (final T and u
  within 1e-9: the seed changes Newton's starting point, not its tolerance),
  start cells from the network, and need fewer Newton iterations.
"""
import ctypes as C

import numpy as np
import pytest

import fivo2d as F
from test_oracle_mlp import oracle_mlp, rand_params, _FP

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_match_oracle(hip):
    lib = oracle_mlp()
    rng = np.random.default_rng(11)
    m = F.Mlp(capacity=0, seed=5, lib=hip)
    p = rand_params(rng, 0.25)
    m.set_params(p)
    n = 200                                   # 12.5 tiles of 16: a ragged last tile
    x = rng.standard_normal((n, 6)).astype(np.float32)
    t = rng.standard_normal((n, 6)).astype(np.float32)
    y = m.forward(x)
    yo = np.zeros((n, 6), np.float32)
    lib.orc_mlp_forward(p.ctypes.data_as(_FP), n, x.ctypes.data_as(_FP), yo.ctypes.data_as(_FP))
    dy = np.max(np.abs(y - yo)) / np.max(np.abs(yo))
    print("forward: max rel diff", dy, flush=True)
    assert dy <= 2e-3
    loss, g = m.grad(x, t)
    go = np.zeros(4998)
    Lo = C.c_double(0)
    lib.orc_mlp_grad(p.ctypes.data_as(_FP), n, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP), C.byref(Lo),
                     go.ctypes.data_as(C.POINTER(C.c_double)))
    assert abs(loss - Lo.value) <= 1e-4 * Lo.value, (loss, Lo.value)
    blocks = {"W1": (0, 384), "b1": (384, 448), "W2": (448, 4544), "b2": (4544, 4608), "W3": (4608, 4992),
              "b3": (4992, 4998)}
    for name, (a, b) in blocks.items():
        r = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        print(f"grad {name}: rel diff {r:.2e}", flush=True)
        assert r <= 1e-2, (name, r)
    m.close()


def full_model():
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                            enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                            kappa_coeff=0.1163, pi_clip_factor=1.0, consistent_m2=True)


def trained_mlp(hip, m, steps=6, N=96):
    """A network trained on pairs from two events of the full-sector model."""
    net = F.Mlp(capacity=2_000_000, seed=3, lib=hip)
    for seed in (21, 22):
        g = F.make_grid2d(N, N, xmax=12.0, ymax=12.0)
        s = F.Solver(g, m, lib=hip)
        s.init_hotspots(0.4, seed=seed, nspots=150, sigma=0.4, Tpeak=0.45)
        tau = 0.4
        s.update_primitives_2d(tau)
        net.collect(s, tau)
        for _ in range(steps):
            r = s.run_sim_2d(tau0=tau, taufinal=1e9, integrator="ssprk3", max_steps=1, reset_history=(tau == 0.4))
            tau = r["tau"]
            net.collect(s, tau)
        s.close()
    net.fit_normalization()
    return net


def test_training_lowers_loss(hip):
    net = trained_mlp(hip, full_model())
    n = net.npairs
    loss = net.train(epochs=6, batch=4096, lr=3e-3)
    print("pairs", n, "loss per epoch", loss, flush=True)
    mse = net.eval()
    print("per-output mse (normalised)", mse, flush=True)
    assert n > 10_000
    assert np.all(np.isfinite(loss)) and loss[-1] < 0.5 * loss[0]
    assert mse[0] < 0.05 and mse[1] < 0.05        # yT and phi, the outputs Newton starts from
    net.close()


def test_learned_seed_run_matches_cold_seed_run(hip):
    m = full_model()
    net = trained_mlp(hip, m)
    net.train(epochs=6, batch=4096, lr=3e-3)
    g = F.make_grid2d(80, 72, xmax=12.0, ymax=10.8)
    out = {}
    for learned in (False, True):
        s = F.Solver(g, m, lib=hip)
        s.init_hotspots(0.4, seed=99, nspots=120, sigma=0.4, Tpeak=0.45)   # an event the network never saw
        if learned:
            s.set_learned_seed(net)
        r = s.run_sim_2d(tau0=0.4, taufinal=0.7, integrator="ssprk3")
        assert r["ok"]
        out[learned] = (r, s.get_work(("yT", "ux", "uy")))
        s.close()
    (r0, W0), (r1, W1) = out[False], out[True]
    print(f"cold seed: nsteps {r0['nsteps']} newton {r0['newton_iters']}; learned: nsteps {r1['nsteps']} "
          f"newton {r1['newton_iters']} learned cells {r1['nlearned']} fallbacks {r1['nlearned_fallback']}", flush=True)
    assert r0["nsteps"] == r1["nsteps"]
    assert r1["nlearned"] > 0 and r0["nlearned"] == 0
    assert r1["nlearned_fallback"] <= 0.15 * r1["nlearned"]
    sx, sy = g.interior()
    T0 = np.exp(W0["yT"].reshape(g.Nxtot, g.Nytot)[sx, sy])
    T1 = np.exp(W1["yT"].reshape(g.Nxtot, g.Nytot)[sx, sy])
    dT = np.max(np.abs(T1 - T0) / T0)
    du = max(np.max(np.abs(W1[f] - W0[f])) for f in ("ux", "uy"))
    print(f"learned vs cold seed: dT/T {dT:.2e} du {du:.2e}", flush=True)
    assert dT <= 1e-9 and du <= 1e-9
    assert r1["newton_iters"] < r0["newton_iters"]
    net.close()
