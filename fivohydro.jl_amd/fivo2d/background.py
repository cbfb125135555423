"""export_background2d (tools/export_background2d.jl): a 2+1D run written as the
background a heavy-quark transport code reads.

The tables are T (GeV), the VELOCITY v^i = u^i / u^tau (not the four-velocity the
solver stores; the transport clamps |v| < 1 and boosts with 1/sqrt(1 - v^2)) and
the charm density n, on the interior (x, y) cell centres at nt uniformly spaced
Milne times, as column-major (Nx, Ny, nt) arrays (Julia's T_grid[i, j, k]).

Each slice is written by the library (fivo_background_slice) from the primitives
of a primitive pass at the slice's time: with the HIP library the tables live in
HBM for the whole run and cross to the host once at the end; with a host library
(the CPU oracle) they are host arrays.

The reference saves a JLD2 (HDF5) file. No HDF5 writer exists in this
environment, so `out` is written as a NumPy .npz holding the same keys with the
same shapes and values (strings as 0-d unicode arrays); nothing in it needs
pickle to load.
"""
import math
import time

import numpy as np

from . import _abi
from .model import LatticeHRGEOS, build_model_2d, make_grid2d
from .solver import FivoError, Solver

LABEL = "fivo_2d_background"
VELOCITY_CONVENTION = "v^i = u^i/u^tau (VELOCITY, not four-velocity)"
NOTE = "post-D8 tau_n (bare, g_hq cancels) and post-D9 vacuum-ramp relaxation"


def background_model(DsT=0.1163, eta_over_s=0.10, zeta_over_s=0.10):
    """The exporter's medium (tools/export_background2d.jl:52-57): LatticeHRG EOS,
    shear (C_s 0.2, delta_pipi 4/3), bulk (C_zeta 15), charm diffusion (tau_n
    coefficient 1) and the shear clip at 1."""
    return build_model_2d(eos=LatticeHRGEOS(), enable_shear=True, eta_over_s=eta_over_s, tauShear_coeff=0.2,
                          deltaShear_factor=4 / 3, enable_bulk=True, zeta_over_s=zeta_over_s, tauPi_coeff=15.0,
                          enable_diff=True, kappa_coeff=DsT, tauN_coeff=1.0, pi_clip_factor=1.0)


def _is_device_lib(lib):
    return getattr(lib, "prefix", "fivo_") == "fivo_"


def _init(s, g, ic, tau0):
    """`ic`: a grid CSV path (initialize_from_grid_csv!), a dict of primitive arrays
    over the Ntot cells (set_cell! + finalize_ic!), or a callable g -> such a dict."""
    if isinstance(ic, str):
        s.initialize_from_grid_csv(tau0, ic)
        return ic
    prims = ic(g) if callable(ic) else ic
    s.init_from_prims(tau0, **prims)
    return getattr(ic, "__name__", "primitive arrays")


class _Tables:
    """(Nx, Ny, nt) tables T, vx, vy, n: GPU tensors (HIP library) or Fortran-order
    host arrays (host library)."""

    def __init__(self, g, nt, device):
        self.g, self.nt, self.device = g, nt, device
        if device:
            import torch
            self.t = [torch.empty(nt * g.Ny * g.Nx, dtype=torch.float64, device="cuda") for _ in range(4)]
        else:
            self.t = [np.empty((g.Nx, g.Ny, nt), dtype=np.float64, order="F") for _ in range(4)]

    def slice(self, s, k):
        return s.background_slice(k, self.nt, *self.t)

    def host(self):
        if not self.device:
            return self.t
        g, nt = self.g, self.nt
        return [x.cpu().numpy().reshape(nt, g.Ny, g.Nx).transpose(2, 1, 0) for x in self.t]


def export_background2d(ic, out=None, N=200, tau0=0.4, tauf=8.0, nt=39, xmax=16.0, DsT=0.1163,
                        eta_over_s=0.10, zeta_over_s=0.10, lib=None, CFL=0.2, CFLtau=0.02,
                        integrator="ssprk2", verbose=True):
    """Run the 2-D solver from `ic` and return (and, with `out`, save) the background
    (tools/export_background2d.jl:49-115): keys label, xgrid, ygrid, tgrid, T_grid,
    vx_grid, vy_grid, n_grid and the provenance (ic, N, xmax, tau0, tauf, DsT,
    eta_over_s, zeta_over_s, velocity_convention, note). The slices are on a uniform
    tau grid; each chunk runs run_sim_2d! on the same work arrays (histories kept
    across chunks, as the reference passes its work array), then
    update_primitives_2d! at the slice's time. Raises if a chunk fails or if
    max |v| >= 1 (the check that the four-velocity was converted)."""
    lib = lib if lib is not None else _abi.load_hip()
    g = make_grid2d(N, N, xmax=xmax, ymax=xmax)
    m = background_model(DsT, eta_over_s, zeta_over_s)
    s = Solver(g, m, lib=lib)
    t_start = time.perf_counter()
    try:
        ic_name = _init(s, g, ic, tau0)
        ng = g.nghost
        xs = np.array(g.xC[ng:ng + g.Nx], dtype=np.float64)
        ys = np.array(g.yC[ng:ng + g.Ny], dtype=np.float64)
        ts = np.linspace(tau0, tauf, nt)
        tabs = _Tables(g, nt, _is_device_lib(lib))
        if verbose:
            print(f"FiVo 2-D background: {ic_name}\n  N={N}  box=±{xmax:.1f} fm  tau {tau0:.2f}→{tauf:.2f} "
                  f"in {nt} slices  D_sT={DsT:.4f}", flush=True)
        tau, maxu, maxv, steps = tau0, 0.0, 0.0, 0
        jmax = lambda a, b: math.nan if (math.isnan(a) or math.isnan(b)) else max(a, b)
        for k, tk in enumerate(ts):
            if tk > tau + 1e-12:
                r = s.run_sim_2d(tau0=tau, taufinal=float(tk), CFL=CFL, CFLtau=CFLtau, integrator=integrator,
                                 reset_history=False)
                if not r["ok"]:
                    raise FivoError(f"export_background2d: the solve failed at tau = {r['tau']}")
                tau = r["tau"]
                steps += r["nsteps"]
            s.update_primitives_2d(tau)
            mu, mv = tabs.slice(s, k)
            maxu, maxv = jmax(maxu, mu), jmax(maxv, mv)
        T, vx, vy, n = tabs.host()
    finally:
        s.close()
    if not maxv < 1.0:
        raise FivoError(f"export_background2d: |v| = {maxv} >= 1 — the four-velocity conversion is wrong")
    wall = time.perf_counter() - t_start
    if verbose:
        print(f"  max |u| = {maxu:.3f}  ->  max |v| = {maxv:.4f}   (a table of u would have been clamped to 1)")
        print(f"  T range {T.min():.4f} … {T.max():.4f} GeV")
    res = dict(label=LABEL, xgrid=xs, ygrid=ys, tgrid=ts, T_grid=T, vx_grid=vx, vy_grid=vy, n_grid=n,
               ic=str(ic_name), N=N, xmax=xmax, tau0=tau0, tauf=tauf, DsT=DsT, eta_over_s=eta_over_s,
               zeta_over_s=zeta_over_s, velocity_convention=VELOCITY_CONVENTION, note=NOTE)
    if out is not None:
        np.savez(out, **res)
        if verbose:
            print(f"  wrote {out}  ({len(xs)} × {len(ys)} × {nt})")
    res["_run"] = {"steps": steps, "wall_s": wall, "max_u": maxu, "max_v": maxv}
    return res


def export_backgrounds2d(ics, N=200, tau0=0.4, tauf=8.0, nt=39, xmax=16.0, DsT=0.1163, eta_over_s=0.10,
                         zeta_over_s=0.10, lib=None, CFL=0.2, CFLtau=0.02, integrator="ssprk2", learned=None,
                         autotune=True, verbose=False):
    """export_background2d for many events at once: every event's chunk between two
    slice times is one ensemble run (one launch per phase over all events, the step
    replayed as a hipGraph, each event with its own dt), then each event's slice is
    written into its own tables in HBM. `ics` is a list of what export_background2d
    takes as `ic`; `N` one size or one per event. Each event's tables equal
    export_background2d(...) with the same integrator (SSPRK2 with MOOD retries, the
    reference exporter's, by default) bit for bit. `learned`: an Mlp to seed cold recoveries (results then agree with
    the plain path to the learned seed's tolerance, not bit for bit). Returns one
    result dict per event, with the exporter's keys."""
    from .batch import Ensemble
    lib = lib if lib is not None else _abi.load_hip()
    if not _is_device_lib(lib):
        raise ValueError("export_backgrounds2d needs the HIP library (ensembles run on the device)")
    Ns = [int(N)] * len(ics) if np.ndim(N) == 0 else [int(n) for n in N]
    grids = [make_grid2d(n, n, xmax=xmax, ymax=xmax) for n in Ns]
    m = background_model(DsT, eta_over_s, zeta_over_s)
    ens = Ensemble(grids, m, lib=lib)
    t_start = time.perf_counter()
    try:
        names = [_init(ev, g, ic, tau0) for ev, g, ic in zip(ens.events, grids, ics)]
        if learned is not None:
            ens.set_learned_seed(learned)
        if autotune:
            ens.autotune(tau0)
        ts = np.linspace(tau0, tauf, nt)
        tabs = [_Tables(g, nt, True) for g in grids]
        taus = [tau0] * len(ics)
        steps = [0] * len(ics)
        maxu = [0.0] * len(ics)
        maxv = [0.0] * len(ics)
        jmax = lambda a, b: math.nan if (math.isnan(a) or math.isnan(b)) else max(a, b)
        for k, tk in enumerate(ts):
            if tk > min(taus) + 1e-12:
                rs = ens.run_sim_2d(tau0=taus, taufinal=float(tk), CFL=CFL, CFLtau=CFLtau, reset_history=False,
                                    integrator=integrator)
                for e, r in enumerate(rs):
                    if not r["ok"]:
                        raise FivoError(f"export_backgrounds2d: event {e}'s solve failed at tau = {r['tau']}")
                    taus[e] = r["tau"]
                    steps[e] += r["nsteps"]
            for e, ev in enumerate(ens.events):
                ev.update_primitives_2d(taus[e])
                mu, mv = tabs[e].slice(ev, k)
                maxu[e], maxv[e] = jmax(maxu[e], mu), jmax(maxv[e], mv)
        hosts = [t.host() for t in tabs]
    finally:
        ens.close()
    wall = time.perf_counter() - t_start
    out = []
    for e, (g, (T, vx, vy, n)) in enumerate(zip(grids, hosts)):
        if not maxv[e] < 1.0:
            raise FivoError(f"export_backgrounds2d: event {e}: |v| = {maxv[e]} >= 1")
        ng = g.nghost
        res = dict(label=LABEL, xgrid=np.array(g.xC[ng:ng + g.Nx]), ygrid=np.array(g.yC[ng:ng + g.Ny]),
                   tgrid=ts, T_grid=T, vx_grid=vx, vy_grid=vy, n_grid=n, ic=str(names[e]), N=Ns[e], xmax=xmax,
                   tau0=tau0, tauf=tauf, DsT=DsT, eta_over_s=eta_over_s, zeta_over_s=zeta_over_s,
                   velocity_convention=VELOCITY_CONVENTION, note=NOTE)
        res["_run"] = {"steps": steps[e], "wall_s": wall, "max_u": maxu[e], "max_v": maxv[e]}
        out.append(res)
    if verbose:
        print(f"FiVo 2-D backgrounds: {len(ics)} events, {sum(steps)} event-steps in {wall:.2f} s", flush=True)
    return out


def load_background(path):
    """Read a file written by export_background2d(out=...) (no pickle)."""
    with np.load(path, allow_pickle=False) as z:
        out = {}
        for k in z.files:
            a = z[k]
            out[k] = a.item() if a.ndim == 0 else a
        return out
