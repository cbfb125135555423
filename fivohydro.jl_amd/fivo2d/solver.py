"""Host-side mirror of the reference's 2+1D driver API over the C ABI.

`Solver` owns one library context (device state + work arrays) and exposes the
reference's sweep and driver functions under their own names:

  rhs_2d               src2d/rhs2d.jl:169-249
  update_primitives_2d src2d/rhs2d.jl:57-65
  apply_bc_2d          src2d/bc2d.jl:14-65
  sanitize_stage_2d    src2d/floors2d.jl:218-227
  state_ok_2d          src2d/timestepper2d.jl:56-66
  compute_dt_2d        src2d/timestepper2d.jl:37-47
  step_ssprk{2,3}_2d   src2d/timestepper2d.jl:74-188
  relax_dissipative_2d src2d/dissipation2d.jl:293-363
  project_shear        main2D.jl:332-358
  run_sim_2d           main2D.jl:371-561
  fields_2d            main2D.jl:588-600
  background_slice     tools/export_background2d.jl:79-94 (one time slice)
  set_cell/finalize_ic main2D.jl:142-171

State arrays cross the boundary in the reference layout (U[i, a] here, i.e.
Julia's U[a, i]); the library converts to its device structure-of-arrays.
"""
import ctypes as C
import math

import numpy as np

from . import _abi
from .model import T_MIN

_DP = C.POINTER(C.c_double)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


class FivoError(RuntimeError):
    pass


class Solver:
    def __init__(self, g, model, lib=None):
        self.g = g
        self.model = model
        self.lib = lib if lib is not None else _abi.load_hip()
        self._gd = g.desc()
        self._md = model.desc()
        ctx = C.c_void_p()
        rc = self.lib.create(C.byref(self._gd), C.byref(self._md), C.byref(ctx))
        if rc != 0 or not ctx.value:
            raise FivoError(f"create failed ({rc}): {self.lib.last_error(None)}")
        self.ctx = ctx
        self.nv = self.lib.nvars(ctx)
        self.Ntot = self.lib.ntot(ctx)
        assert self.nv == model.layout.nv and self.Ntot == g.Ntot

    @classmethod
    def borrowed(cls, g, model, lib, ctx):
        """A Solver over a context owned elsewhere (an event of an ensemble)."""
        s = cls.__new__(cls)
        s.g, s.model, s.lib = g, model, lib
        s._gd, s._md = g.desc(), model.desc()
        s.ctx = C.c_void_p(ctx)
        s._owned = False
        s.nv = lib.nvars(s.ctx)
        s.Ntot = lib.ntot(s.ctx)
        return s

    def close(self):
        # trackers bound to this context (fivo2d.freezeout) go first
        for t in list(getattr(self, "_trackers", ())):
            t.close()
        if getattr(self, "ctx", None) is not None and self.ctx.value and getattr(self, "_owned", True):
            self.lib.destroy(self.ctx)
        self.ctx = None
        self._mlp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            raise FivoError(f"{what} failed ({rc}): {self.lib.last_error(self.ctx)}")
        return rc

    @property
    def E_vac_cut(self):
        return self.lib.e_vac_cut(self.ctx)

    # ---- state ---------------------------------------------------------------
    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.Ntot, self.nv)
        self._chk(self.lib.set_state(self.ctx, _dp(U), self.nv, self.Ntot), "set_state")

    def get_state(self):
        U = np.empty((self.Ntot, self.nv), dtype=np.float64)
        self._chk(self.lib.get_state(self.ctx, _dp(U), self.nv, self.Ntot), "get_state")
        return U

    def get_work(self, names=_abi.WORK_FIELDS):
        out = {}
        v = _abi.WorkView()
        for k in names:
            if k == "ok":
                a = np.zeros(self.Ntot, dtype=np.uint8)
                v.ok = a.ctypes.data_as(C.POINTER(C.c_uint8))
            else:
                a = np.zeros(self.Ntot, dtype=np.float64)
                setattr(v, k, _dp(a))
            out[k] = a
        self._chk(self.lib.get_work(self.ctx, C.byref(v)), "get_work")
        return out

    def set_work(self, arrays):
        v = _abi.WorkView()
        keep = []
        for k, a in arrays.items():
            if k == "ok":
                a = np.ascontiguousarray(a, dtype=np.uint8)
                v.ok = a.ctypes.data_as(C.POINTER(C.c_uint8))
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                setattr(v, k, _dp(a))
            keep.append(a)
        self._chk(self.lib.set_work(self.ctx, C.byref(v)), "set_work")

    def reset_history(self):
        self._chk(self.lib.reset_history(self.ctx), "reset_history")

    # ---- initial conditions ----------------------------------------------------
    def init_from_prims(self, tau0, T, alpha, ux=None, uy=None, Pi=None, pixx=None, pixy=None,
                        piyy=None, nux=None, nuy=None, bc=_abi.FIVO_BC_OUTFLOW):
        """set_cell! on every cell (main2D.jl:142-154) then finalize_ic! (:164-171)."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                for a in (T, alpha, ux, uy, Pi, pixx, pixy, piyy, nux, nuy)]
        for a in arrs:
            assert a is None or a.size == self.Ntot
        nbad = C.c_int64(0)
        self._chk(self.lib.init_from_prims(self.ctx, float(tau0), *[_dp(a) for a in arrs], int(bc),
                                           C.byref(nbad)), "init_from_prims")
        return nbad.value

    def init_hotspots(self, tau0, seed, nspots=150, sigma=0.5, Tpeak=0.45, R=6.6, a=0.55, alpha_floor=-30.0,
                      bc=_abi.FIVO_BC_OUTFLOW):
        """The seeded hot-spot event of ic.hotspot_event, built on the device
        (fivo_init_hotspots), then set_cell! + finalize_ic! (main2D.jl:142-171)."""
        from .ic import hotspot_table
        tab = hotspot_table(seed, nspots, R, a)
        nbad = C.c_int64(0)
        self._chk(self.lib.init_hotspots(self.ctx, float(tau0), int(nspots), _dp(tab), float(sigma), float(Tpeak),
                                         float(alpha_floor), int(bc), C.byref(nbad)), "init_hotspots")
        return nbad.value

    def initialize_from_grid_csv(self, tau0, csv, taper_width=1.0, smooth_fm=0.0, bc=_abi.FIVO_BC_OUTFLOW):
        """initialize_from_grid_csv! (main2D.jl:238-301): the file's T and alpha
        interpolated onto the grid (fivo2d.ic.grid_csv), set_cell! on every cell and
        finalize_ic! on the device. Returns {"nbad", "nvacuum"}: cells whose
        set_cell! failed, split like the reference into vacuum (T <= 10 T_MIN) and
        cells that had matter."""
        from .ic import grid_csv
        prims = grid_csv(self.g, csv, taper_width=taper_width, smooth_fm=smooth_fm)
        vac = prims["T"] <= 10 * T_MIN
        if vac.any():
            # failures among the cells with matter: the same build with the vacuum
            # cells at a state that always succeeds; then the real build
            # (a failed set_cell! leaves its cell as it was, so the state is put back
            # between the two)
            probe = dict(prims, T=np.where(vac, 1.0, prims["T"]), alpha=np.where(vac, 0.0, prims["alpha"]))
            U0 = self.get_state()
            nbad = self.init_from_prims(tau0, bc=bc, **probe)
            self.set_state(U0)
            nfail_all = self.init_from_prims(tau0, bc=bc, **prims)
            return {"nbad": nbad, "nvacuum": nfail_all - nbad}
        return {"nbad": self.init_from_prims(tau0, bc=bc, **prims), "nvacuum": 0}

    def set_learned_seed(self, mlp):
        """Start cells without a warm start from `mlp`'s guess (None: the cold seed).
        The solver keeps a reference to `mlp` while it is attached, so the network's
        device buffers outlive every primitive pass that reads them."""
        if mlp is not None and not (getattr(mlp, "m", None) is not None and mlp.m.value):
            raise FivoError("set_learned_seed: the network is closed")
        self._chk(self.lib.set_learned_seed(self.ctx, None if mlp is None else mlp.m), "set_learned_seed")
        self._mlp = mlp
        if mlp is not None:
            mlp._users.add(self)

    def finalize_ic(self, tau0, bc=_abi.FIVO_BC_OUTFLOW):
        self._chk(self.lib.finalize_ic(self.ctx, float(tau0), int(bc)), "finalize_ic")

    # ---- sweeps --------------------------------------------------------------
    def apply_bc_2d(self, bc=_abi.FIVO_BC_OUTFLOW):
        self._chk(self.lib.apply_bc(self.ctx, int(bc)), "apply_bc")

    def rhs_2d(self, tau, bc=_abi.FIVO_BC_OUTFLOW, force_first_order=None, want_dU=True):
        dU = np.zeros((self.Ntot, self.nv), dtype=np.float64) if want_dU else None
        ff = None
        if force_first_order is not None:
            ff = np.ascontiguousarray(force_first_order, dtype=np.uint8)
            assert ff.size == self.Ntot
        cnt = _abi.Counts()
        self._chk(self.lib.rhs(self.ctx, float(tau), int(bc),
                               None if ff is None else ff.ctypes.data_as(C.POINTER(C.c_uint8)),
                               _dp(dU), C.byref(cnt)), "rhs")
        return dU, {"nprimfail": cnt.nprimfail, "nvacuum": cnt.nvacuum, "amax": cnt.amax}

    def update_primitives_2d(self, tau, bc=_abi.FIVO_BC_OUTFLOW):
        cnt = _abi.Counts()
        self._chk(self.lib.update_primitives(self.ctx, float(tau), int(bc), C.byref(cnt)),
                  "update_primitives")
        return cnt.nprimfail

    def sanitize_stage_2d(self, tau, bc=_abi.FIVO_BC_OUTFLOW):
        self._chk(self.lib.sanitize(self.ctx, float(tau), int(bc)), "sanitize")

    def state_ok_2d(self, tau):
        ok = C.c_int(0)
        nbad = C.c_int64(0)
        self._chk(self.lib.state_ok(self.ctx, float(tau), C.byref(ok), C.byref(nbad)), "state_ok")
        return bool(ok.value), nbad.value

    def compute_dt_2d(self, tau, CFL=0.2, CFLtau=0.05):
        dt = C.c_double(0)
        self._chk(self.lib.compute_dt(self.ctx, float(tau), float(CFL), float(CFLtau), C.byref(dt)),
                  "compute_dt")
        return dt.value

    def step(self, tau, dtau, integrator=_abi.FIVO_SSPRK2, bc=_abi.FIVO_BC_OUTFLOW, max_halve=12,
             max_mood=2):
        ok = C.c_int(0)
        used = C.c_double(0)
        self._chk(self.lib.step(self.ctx, int(integrator), float(tau), float(dtau), int(bc),
                                int(max_halve), int(max_mood), C.byref(ok), C.byref(used)), "step")
        return bool(ok.value), used.value

    def relax_dissipative_2d(self, tau, dtau):
        self._chk(self.lib.relax(self.ctx, float(tau), float(dtau)), "relax")

    def project_shear(self):
        mx = C.c_double(0)
        mean = C.c_double(0)
        self._chk(self.lib.project_shear(self.ctx, C.byref(mx), C.byref(mean)), "project_shear")
        return mx.value, mean.value

    def run_sim_2d(self, tau0=0.4, taufinal=13.0, CFL=0.2, CFLtau=0.05, integrator="ssprk2",
                   bc=_abi.FIVO_BC_OUTFLOW, reset_history=True, on_dump=None, dump_dt=0.5,
                   max_steps=2_000_000, max_halve=12, max_mood=2):
        """run_sim_2d! (main2D.jl:371-561). `on_dump(tau, solver)` like the reference's callback."""
        o = _abi.RunOpts()
        o.tau0, o.taufinal, o.cfl, o.cfltau, o.dump_dt = tau0, taufinal, CFL, CFLtau, dump_dt
        o.integrator = _abi.FIVO_SSPRK3 if integrator in ("ssprk3", 3) else _abi.FIVO_SSPRK2
        o.bc = int(bc)
        o.max_halve, o.max_mood, o.max_steps = max_halve, max_mood, max_steps
        o.reset_history = int(bool(reset_history))
        cb = None
        if on_dump is not None:
            def _cb(user, tau, ctx):
                on_dump(tau, self)
            cb = _abi.DUMP_CB(_cb)
            o.on_dump = cb
        r = _abi.RunResult()
        rc = self.lib.run(self.ctx, C.byref(o), C.byref(r))
        self._chk(rc, "run")
        return r.as_dict()

    # ---- readout -------------------------------------------------------------
    def background_slice(self, k, nt, T, vx, vy, n=None):
        """One time slice of export_background2d (tools/export_background2d.jl:79-94)
        from the primitives of the last primitive pass: T = exp(yT), the velocity
        v^i = u^i / u^tau and n of every interior cell into element (i, j, k) of
        (Nx, Ny, nt) tables, column-major (Julia's T_grid[i, j, k]).

        Host tables: numpy float64 arrays of shape (Nx, Ny, nt) in Fortran order.
        Device tables (the HIP library only): contiguous float64 GPU tensors of
        Nx*Ny*nt values in the same order, e.g. torch tensors of shape (nt, Ny, Nx);
        the slice is written in HBM and nothing crosses to the host.
        Returns (max |u|, max |v|) over the slice (NaN if a cell's is NaN)."""
        g = self.g
        shape = (g.Nx, g.Ny, int(nt))
        tabs = [T, vx, vy, n]
        on_dev = [hasattr(a, "data_ptr") for a in tabs if a is not None]
        if any(on_dev) != all(on_dev):
            raise ValueError("background_slice: the tables must all be on the host or all on the device")
        device = bool(on_dev and on_dev[0])
        ptrs = []
        for a in tabs:
            if a is None:
                ptrs.append(None)
            elif device:
                if not (a.is_cuda and str(a.dtype) == "torch.float64" and a.is_contiguous()
                        and a.numel() == g.Nx * g.Ny * int(nt)):
                    raise ValueError("background_slice: device tables must be contiguous float64 GPU "
                                     f"tensors of {g.Nx}*{g.Ny}*{int(nt)} values")
                ptrs.append(C.cast(C.c_void_p(a.data_ptr()), _DP))
            else:
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == shape
                        and a.flags.f_contiguous):
                    raise ValueError(f"background_slice: host tables must be float64 arrays of shape {shape} "
                                     "in Fortran order")
                ptrs.append(_dp(a))
        mx = (C.c_double * 2)()
        self._chk(self.lib.background_slice(self.ctx, int(k), int(nt), ptrs[0], ptrs[1], ptrs[2], ptrs[3],
                                            int(device), mx), "background_slice")
        return mx[0], mx[1]

    def fields_2d(self):
        """fields_2d (main2D.jl:588-600): interior fields as (Nx, Ny) arrays [ix, iy]."""
        g = self.g
        w = self.get_work(("yT", "mu", "alpha", "n", "e", "P", "ux", "uy"))
        U = self.get_state()
        sx, sy = g.interior()
        grab = lambda a: a.reshape(g.Nxtot, g.Nytot)[sx, sy].copy()
        out = {"x": g.xC[sx].copy(), "y": g.yC[sy].copy(), "T": np.exp(grab(w["yT"])),
               "mu": grab(w["mu"]), "alpha": grab(w["alpha"]), "n": grab(w["n"]), "e": grab(w["e"]),
               "P": grab(w["P"]), "ux": grab(w["ux"]), "uy": grab(w["uy"])}
        for a, name in enumerate(self.model.layout.names[4:], start=4):
            out[name] = grab(U[:, a])
        return out
