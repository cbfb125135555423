"""The 2-D freeze-out surface and its Cooper-Frye pion yield (freezeout2d.jl).

Host mirror of the reference's `freeze_surface` and `pion_yield` over the C ABI
fivo_freezeout_*: the tracker records tau_fo(x, y), the last downward crossing
of T_fo in every interior cell, on the device after each primitive pass, and
Cooper-Frye integrates dN/dy over that surface on the device.

  freeze_surface(tag, ...)   freezeout2d.jl:68-127 (evolution in dtau chunks,
                             record! after each)
  pion_yield(s, ...)         freezeout2d.jl:144-192
  FreezeOut                  the tracker on one Solver

The surface is boost invariant, dSigma_mu = tau_fo (1, -d_x tau_fo, -d_y tau_fo, 0)
dx dy deta; a cell that re-heats above T_fo after freezing loses its tau_fo and
is counted, so the single-valued approximation is measured (freezeout2d.jl:13-27).
"""
import ctypes as C
import weakref

import numpy as np

from . import _abi, ic
from .model import LatticeHRGEOS, build_model_2d, make_grid2d
from .solver import FivoError, Solver

# conventions of the 1-D calibration (freezeout2d.jl:51-59)
T_FO = 0.156
M_PI = 0.138
FYIELD = 2.0 * 2.5
PT_LO, PT_HI = 0.1, 5.0
TAU0 = 0.4


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class FreezeOut:
    """The freeze-out tracker on `solver` (freezeout2d.jl:87-119). Holds the
    solver; close() (or garbage collection) frees the device arrays."""

    def __init__(self, solver, T_fo=T_FO):
        self.solver = solver
        self.lib = solver.lib
        self.T_fo = float(T_fo)
        self.f = C.c_void_p()
        rc = self.lib.freezeout_create(solver.ctx, self.T_fo, C.byref(self.f))
        if rc != 0:
            raise FivoError(f"freezeout_create failed ({rc}): {self.lib.last_error(solver.ctx)}")
        if not hasattr(solver, "_trackers"):
            solver._trackers = weakref.WeakSet()
        solver._trackers.add(self)
        g = solver.g
        self.nx, self.ny = g.Nx, g.Ny
        sx, sy = g.interior()
        self.xs = np.asarray(g.xC)[sx].copy()
        self.ys = np.asarray(g.yC)[sy].copy()
        self.dx, self.dy = g.dx, g.dy

    def close(self):
        if getattr(self, "f", None) is not None and self.f.value:
            self.lib.freezeout_destroy(self.f)
            self.f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise FivoError(f"{what} failed ({rc}): {self.lib.last_error(self.solver.ctx)}")

    def record(self, tau):
        """record! at tau from the solver's current primitives (after rhs_2d,
        update_primitives_2d or run_sim_2d)."""
        self._chk(self.lib.freezeout_record(self.f, float(tau)), "freezeout_record")

    def surface(self):
        """dict: xs, ys, tau_fo (nx, ny; NaN where not frozen), uxf, uyf, dx, dy,
        nhot (cells that were ever hot), nfroze, refroze (re-heating events)."""
        n = self.nx * self.ny
        t, ux, uy = np.empty(n), np.empty(n), np.empty(n)
        cnt = (C.c_int64 * 3)()
        self._chk(self.lib.freezeout_surface(self.f, _dp(t), _dp(ux), _dp(uy), cnt), "freezeout_surface")
        shp = (self.nx, self.ny)
        return dict(xs=self.xs, ys=self.ys, tau_fo=t.reshape(shp), uxf=ux.reshape(shp), uyf=uy.reshape(shp),
                    dx=self.dx, dy=self.dy, nhot=int(cnt[0]), nfroze=int(cnt[1]), refroze=int(cnt[2]))

    def set_surface(self, tau_fo, uxf, uyf):
        """Replace the surface, e.g. to integrate one built elsewhere."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (tau_fo, uxf, uyf)]
        for a in arrs:
            if a.size != self.nx * self.ny:
                raise ValueError(f"set_surface: expected {self.nx}x{self.ny} arrays")
        self._chk(self.lib.freezeout_set_surface(self.f, *[_dp(a) for a in arrs]), "freezeout_set_surface")

    def pion_yield(self, n_eta=40, n_pt=40, n_phi=24, eta_max=6.0, pt_lo=PT_LO, pt_hi=PT_HI, mass=M_PI,
                   yield_factor=FYIELD):
        """pion_yield (freezeout2d.jl:144-192): (N_Boltzmann, N_Bose, negative-dSigma fraction)."""
        o = _abi.CfOpts()
        o.n_eta, o.n_pt, o.n_phi = int(n_eta), int(n_pt), int(n_phi)
        o.eta_max, o.pt_lo, o.pt_hi, o.mass, o.yield_factor = eta_max, pt_lo, pt_hi, mass, yield_factor
        out = (C.c_double * 3)()
        self._chk(self.lib.freezeout_pion_yield(self.f, C.byref(o), out), "freezeout_pion_yield")
        return float(out[0]), float(out[1]), float(out[2])


def freeze_model(eta_s=0.10, zeta_s=0.10):
    """The medium of freeze_surface (freezeout2d.jl:72-75)."""
    return build_model_2d(eos=LatticeHRGEOS(), enable_shear=True, eta_over_s=eta_s, tauShear_coeff=0.2,
                          deltaShear_factor=4 / 3, enable_bulk=True, zeta_over_s=zeta_s, tauPi_coeff=15.0,
                          enable_diff=True, kappa_coeff=0.1163, tauN_coeff=1.0, pi_clip_factor=1.0)


def freeze_surface(tag, N=200, box=18.0, dtau=0.02, taumax=16.0, eta_s=0.10, zeta_s=0.10, profile_csv=None,
                   grid_csv=None, lib=None, T_fo=T_FO, CFL=0.15, CFLtau=0.05, restart_history=False, verbose=True):
    """freeze_surface (freezeout2d.jl:68-127): evolve from tau = 0.4 to taumax in
    chunks of dtau, recording the surface after each. tag "axisym": the radial
    profile of `profile_csv` (load_initial_interpolants, taper 1 fm, linear) at
    rest on every cell; any other tag: initialize_from_grid_csv! of `grid_csv`.
    Returns the FreezeOut tracker (its solver attached as .solver).

    The chunks share one work array, so the relaxation history (the previous
    step's u and alpha behind the d_tau pieces of the Navier-Stokes drives) runs
    on across chunks (main2D.jl:376, reset_history = false with a work array).
    restart_history=True clears it at every chunk instead, as run_sim_2d! did
    before it kept it (TWOD_PROGRAM.md:743, D14): the setting the reference's
    published yields were measured with (TWOD_PROGRAM.md:1847, 2396-2431)."""
    g = make_grid2d(N, N, xmax=box, ymax=box)
    m = freeze_model(eta_s, zeta_s)
    s = Solver(g, m, lib=lib)
    if tag == "axisym":
        if profile_csv is None:
            raise ValueError("freeze_surface('axisym') needs profile_csv (the radial profile r, T0, alpha0)")
        s.init_from_prims(TAU0, **ic.elliptic_profile(g, profile_csv, eps2=0.0, taper_width=1.0))
    else:
        if grid_csv is None:
            raise ValueError(f"freeze_surface({tag!r}) needs grid_csv (x, y, T, alpha)")
        s.initialize_from_grid_csv(TAU0, grid_csv)
    s.reset_history()
    fo = FreezeOut(s, T_fo)
    tau = TAU0
    s.rhs_2d(tau, want_dU=False)
    fo.record(tau)
    while tau < taumax - 1e-9:
        r = s.run_sim_2d(tau0=tau, taufinal=min(tau + dtau, taumax), CFL=CFL, CFLtau=CFLtau,
                         reset_history=bool(restart_history))
        if not r["ok"]:
            raise FivoError(f"freeze_surface: run_sim_2d failed at tau = {r['tau']}")
        tau = r["tau"]
        s.update_primitives_2d(tau)
        fo.record(tau)
    fo.tau = tau
    if verbose:
        sf = fo.surface()
        print(f"  {tag:<12s} cells that were ever hot {sf['nhot']:6d} | froze by tau={taumax:.1f} "
              f"{sf['nfroze']:6d} | still hot {sf['nhot'] - sf['nfroze']:4d} | re-heated after freezing "
              f"{sf['refroze']:5d}", flush=True)
    return fo


def pion_yield(s, n_eta=40, n_pt=40, n_phi=24, eta_max=6.0):
    """pion_yield (freezeout2d.jl:144-192) on the tracker's surface."""
    return s.pion_yield(n_eta=n_eta, n_pt=n_pt, n_phi=n_phi, eta_max=eta_max)


NOALLOC_MAX_CELLS = 1 << 20


def pion_yield_device(tau_fo, uxf, uyf, dx, dy, out, T_fo=T_FO, stream=None, lib=None, n_eta=40, n_pt=40, n_phi=24,
                      eta_max=6.0, pt_lo=PT_LO, pt_hi=PT_HI, mass=M_PI, yield_factor=FYIELD):
    """pion_yield on a surface already on the GPU, through the entry point that
    allocates nothing (fivo_pion_yield_noalloc): tau_fo, uxf, uyf are contiguous
    float64 CUDA tensors of shape (nx, ny), out a float64 CUDA tensor of 3
    elements that receives (N_Boltzmann, N_Bose, negative fraction) on `stream`
    (a torch.cuda.Stream; None: the current stream) without synchronising.
    Calls that overlap on one GPU share its scratch and race."""
    lib = lib if lib is not None else _abi.load_hip()
    for a in (tau_fo, uxf, uyf, out):
        if not (a.is_cuda and a.dtype.is_floating_point and a.element_size() == 8 and a.is_contiguous()):
            raise ValueError("pion_yield_device: arrays must be contiguous float64 CUDA tensors")
    nx, ny = tau_fo.shape
    if tuple(uxf.shape) != (nx, ny) or tuple(uyf.shape) != (nx, ny) or out.numel() < 3:
        raise ValueError("pion_yield_device: shapes do not match")
    o = _abi.CfOpts()
    o.n_eta, o.n_pt, o.n_phi = int(n_eta), int(n_pt), int(n_phi)
    o.eta_max, o.pt_lo, o.pt_hi, o.mass, o.yield_factor = eta_max, pt_lo, pt_hi, mass, yield_factor
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    rc = lib.pion_yield_noalloc(C.c_void_p(tau_fo.data_ptr()), C.c_void_p(uxf.data_ptr()),
                                C.c_void_p(uyf.data_ptr()), int(nx), int(ny), float(dx), float(dy), float(T_fo),
                                C.byref(o), C.c_void_p(out.data_ptr()), C.c_void_p(stream.cuda_stream))
    if rc != 0:
        raise FivoError(f"pion_yield_noalloc failed ({rc}): {lib.last_error(None)}")
