"""GPU parity of the dissipative sectors: the relaxation kernel and full runs with
shear + bulk + charge diffusion (+ consistent first moment + charm second moment)
against the CPU oracle on identical inputs."""
import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic
from helpers import rel_diff

pytestmark = pytest.mark.gpu


def viscous_model(m2=False, cfm=False):
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                            enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                            kappa_coeff=0.1163, pi_clip_factor=1.0, consistent_m2=m2, consistent_fm=cfm)


def event(N, L=12.0, seed=3, nspots=150, sigma=0.5):
    g = F.make_grid2d(N, N, xmax=L, ymax=L)
    return g, ic.hotspot_event(g, seed=seed, nspots=nspots, sigma=sigma, Tpeak=0.45)


def _pair(oracle, hip, g, m, prims, tau0):
    out = []
    for lib in (oracle, hip):
        s = F.Solver(g, m, lib=lib)
        s.init_from_prims(tau0, **prims)
        out.append(s)
    return out


@pytest.mark.parametrize("m2,cfm", [(False, False), (True, True)])
def test_relax_parity_single_call(oracle, hip, m2, cfm):
    """One relax_dissipative_2d! on an identical evolved state with history."""
    g, prims = event(48)
    m = viscous_model(m2=m2, cfm=cfm)
    so = F.Solver(g, m, lib=oracle)
    so.init_from_prims(0.6, **prims)
    r = so.run_sim_2d(tau0=0.6, taufinal=0.75, CFL=0.2, CFLtau=0.05, integrator="ssprk2")
    assert r["ok"]
    U = so.get_state()
    W = so.get_work()
    sh = F.Solver(g, m, lib=hip)
    sh.set_state(U)
    sh.set_work(W)
    tau, dtau = r["tau"], 0.01
    so.update_primitives_2d(tau)
    sh.update_primitives_2d(tau)
    wo, wh = so.get_work(("yT", "ux", "uy", "ok")), sh.get_work(("yT", "ux", "uy", "ok"))
    assert np.array_equal(wo["ok"], wh["ok"])
    so.relax_dissipative_2d(tau, dtau)
    sh.relax_dissipative_2d(tau, dtau)
    Uo, Uh = so.get_state(), sh.get_state()
    r = rel_diff(Uh, Uo)
    names = m.layout.names
    for a, nm in enumerate(names):
        tol = 1e-9 if nm.startswith("pQ") or nm == "PiQ" else 1e-12
        assert r[a] <= tol, (nm, r[a])
    if m2:
        assert np.max(np.abs(Uo[:, m.layout.idx("PiQ")])) > 0   # the second moment did evolve


@pytest.mark.parametrize("integ,m2", [("ssprk2", False), ("ssprk3", True)])
def test_run_parity_viscous_charm(oracle, hip, integ, m2):
    g, prims = event(64)
    m = viscous_model(m2=m2)
    so, sh = _pair(oracle, hip, g, m, prims, 0.6)
    kw = dict(tau0=0.6, taufinal=1.0, CFL=0.2, CFLtau=0.05, integrator=integ)
    ro, rh = so.run_sim_2d(**kw), sh.run_sim_2d(**kw)
    assert ro["ok"] and rh["ok"]
    assert ro["nsteps"] == rh["nsteps"] and ro["nhalvings"] == rh["nhalvings"] and ro["nmood"] == rh["nmood"]
    assert ro["nprimfail"] == rh["nprimfail"]
    fo, fh = so.fields_2d(), sh.fields_2d()
    for k in ("T", "ux", "uy"):
        d = np.max(np.abs(fh[k] - fo[k])) / np.max(np.abs(fo[k]))
        assert d <= 1e-9, (k, d)
    for k in m.layout.names[4:]:
        d = np.max(np.abs(fh[k] - fo[k])) / max(np.max(np.abs(fo[k])), 1e-300)
        assert d <= 1e-7, (k, d)
