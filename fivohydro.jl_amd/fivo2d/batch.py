"""Ensembles of events over the C ABI's fivo_batch_* entry points.

The reference runs fluctuating events one at a time, each through its own
run_sim_2d! (main2D.jl:371-561). `Ensemble` holds many events on their own grids
with one model and runs them together: every phase of the step (SSPRK3, or
SSPRK2 with its MOOD stage retries) is one launch over all events and the step
is a replayed hipGraph, while each event keeps its own dt, halvings, retries and
stopping time. `ensemble.events[k]` is a
`Solver` over event k's context, for initial conditions and readout.
"""
import ctypes as C

import numpy as np

from . import _abi
from .solver import Solver, FivoError


class Ensemble:
    def __init__(self, grids, model, lib=None):
        self.lib = lib if lib is not None else _abi.load_hip()
        self.grids = list(grids)
        self.model = model
        n = len(self.grids)
        gd = (_abi.GridDesc * n)(*[g.desc() for g in self.grids])
        self._md = model.desc()
        b = C.c_void_p()
        rc = self.lib.batch_create(gd, n, C.byref(self._md), C.byref(b))
        if rc != 0 or not b.value:
            raise FivoError(f"batch_create failed ({rc}): {self._err(None)}")
        self.b = b
        self.events = [Solver.borrowed(g, model, self.lib, self.lib.batch_event(b, k))
                       for k, g in enumerate(self.grids)]
        self.tile = None

    def _err(self, b):
        s = self.lib.batch_last_error(b)
        return s.decode() if s else ""

    def _chk(self, rc, what):
        if rc < 0:
            raise FivoError(f"{what} failed ({rc}): {self._err(self.b)}")
        return rc

    def close(self):
        if getattr(self, "b", None) is not None and self.b.value:
            for s in self.events:
                s.ctx = None
                s._mlp = None
            self.lib.batch_destroy(self.b)
            self.b = None
        self._mlp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, tau):
        """Time the fused face kernel for every candidate tile on the current
        states at tau and keep the fastest: returns ((tx, ty), {tile: ms})."""
        tx, ty = C.c_int32(0), C.c_int32(0)
        ms = (C.c_float * _abi.FIVO_BATCH_NTILES)()
        self._chk(self.lib.batch_autotune(self.b, float(tau), C.byref(tx), C.byref(ty), ms), "batch_autotune")
        self.tile = (tx.value, ty.value)
        return self.tile, {_abi.BATCH_TILES[k]: ms[k] for k in range(_abi.FIVO_BATCH_NTILES)}

    def set_learned_seed(self, mlp):
        """Start every event's cells without a warm start from `mlp`'s guess (None: the cold seed).
        The ensemble keeps a reference to `mlp` while it is attached (see Solver.set_learned_seed)."""
        if mlp is not None and not (getattr(mlp, "m", None) is not None and mlp.m.value):
            raise FivoError("set_learned_seed: the network is closed")
        self._chk(self.lib.batch_set_learned_seed(self.b, None if mlp is None else mlp.m), "batch_set_learned_seed")
        self._mlp = mlp
        for s in self.events:
            s._mlp = mlp
        if mlp is not None:
            mlp._users.add(self)

    def set_tile(self, tile):
        k = _abi.BATCH_TILES.index(tuple(tile))
        self._chk(self.lib.batch_set_tile(self.b, k), "batch_set_tile")
        self.tile = tuple(tile)

    def run_sim_2d(self, tau0=0.4, taufinal=13.0, CFL=0.2, CFLtau=0.05, bc=_abi.FIVO_BC_OUTFLOW,
                   reset_history=True, max_steps=2_000_000, max_halve=12, integrator="ssprk3", max_mood=2):
        """run_sim_2d! (main2D.jl:371-561) for every event; one result dict per event.
        integrator "ssprk3" (halving only) or "ssprk2" (MOOD stage retries, up to
        max_mood, then halving), as the one-event run_sim_2d.
        tau0: one start time, or one per event (to continue events that stopped apart)."""
        n = len(self.grids)
        t0s = None
        if np.ndim(tau0) > 0:
            t0s = np.ascontiguousarray(tau0, dtype=np.float64)
            assert t0s.shape == (n,)
            tau0 = float(t0s[0])
        o = _abi.RunOpts()
        o.tau0, o.taufinal, o.cfl, o.cfltau, o.dump_dt = tau0, taufinal, CFL, CFLtau, 0.0
        o.integrator = _abi.FIVO_SSPRK3 if integrator in ("ssprk3", 3) else _abi.FIVO_SSPRK2
        o.bc = int(bc)
        o.max_halve, o.max_mood, o.max_steps = max_halve, max_mood, max_steps
        o.reset_history = int(bool(reset_history))
        res = (_abi.RunResult * n)()
        tp = None if t0s is None else t0s.ctypes.data_as(C.POINTER(C.c_double))
        self._chk(self.lib.batch_run(self.b, C.byref(o), tp, res), "batch_run")
        return [r.as_dict() for r in res]
