"""The learned Newton warm start over the C ABI's fivo_mlp_* entry points.

`Mlp` holds a 6 -> 64 -> 64 -> 6 GELU network (BF16 MFMA kernels, FP32
accumulate) and a device-resident set of (features, targets) pairs collected
from the solver's own converged cells. Train it on the GPU, then hand it to a
Solver or an Ensemble (`set_learned_seed`): cells without a warm start then
start Newton from the network's guess instead of the reference's cold seed
(seed_from_conserved_2d, src2d/primrec2d.jl:323-347), with the cold seed as
the fallback. No counterpart in the reference.
"""
import ctypes as C
import weakref

import numpy as np

from . import _abi

NPARAM = 4998
NNORM = 24
_FP = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(_FP)


class Mlp:
    def __init__(self, capacity=10_000_000, seed=1, lib=None):
        self.lib = lib if lib is not None else _abi.load_hip()
        m = C.c_void_p()
        rc = self.lib.mlp_create(int(capacity), int(seed), C.byref(m))
        if rc != 0 or not m.value:
            raise RuntimeError(f"mlp_create failed ({rc})")
        self.m = m
        # solvers and ensembles whose contexts point at this network's device
        # buffers; the network must not be destroyed while any of them holds it
        self._users = weakref.WeakSet()

    def attached(self):
        """The solvers and ensembles currently seeded by this network."""
        return [u for u in list(self._users) if getattr(u, "_mlp", None) is self]

    def _chk(self, rc, what):
        if rc < 0:
            s = self.lib.mlp_last_error(self.m)
            raise RuntimeError(f"{what} failed ({rc}): {s.decode() if s else ''}")
        return rc

    def close(self):
        """Free the network's device buffers. Raises while a solver or ensemble is
        still seeded by it (their contexts read these buffers in every primitive
        pass): detach them first with set_learned_seed(None), or close them."""
        if getattr(self, "m", None) is not None and self.m.value:
            users = self.attached()
            if users:
                raise RuntimeError(f"Mlp.close: the network still seeds {len(users)} solver(s)/ensemble(s); "
                                   "call set_learned_seed(None) on them (or close them) first")
            self.lib.mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def collect(self, solver, tau):
        """Pairs of the converged interior cells of `solver`'s last primitive pass at tau."""
        n = C.c_int64(0)
        self._chk(self.lib.mlp_collect(self.m, solver.ctx, float(tau), C.byref(n)), "mlp_collect")
        return n.value

    @property
    def npairs(self):
        n = C.c_int64(0)
        self._chk(self.lib.mlp_size(self.m, C.byref(n)), "mlp_size")
        return n.value

    def fit_normalization(self):
        self._chk(self.lib.mlp_fit_normalization(self.m), "mlp_fit_normalization")

    def train(self, epochs=4, batch=8192, lr=2e-3):
        loss = np.zeros(epochs, np.float32)
        self._chk(self.lib.mlp_train(self.m, int(epochs), int(batch), float(lr), _fp(loss)), "mlp_train")
        return loss

    def eval(self):
        """Mean squared error of each normalised output over the held pairs."""
        mse = np.zeros(6, np.float32)
        self._chk(self.lib.mlp_eval(self.m, _fp(mse)), "mlp_eval")
        return mse

    def params(self):
        p = np.zeros(NPARAM, np.float32)
        nrm = np.zeros(NNORM, np.float32)
        self._chk(self.lib.mlp_get_params(self.m, _fp(p), _fp(nrm)), "mlp_get_params")
        return p, nrm

    def set_params(self, p=None, norm=None):
        pp = None if p is None else np.ascontiguousarray(p, np.float32)
        nn = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._chk(self.lib.mlp_set_params(self.m, None if pp is None else _fp(pp), None if nn is None else _fp(nn)),
                  "mlp_set_params")

    def forward(self, x):
        """Normalised features (n, 6) -> normalised outputs (n, 6)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 6), np.float32)
        self._chk(self.lib.mlp_forward(self.m, x.shape[0], _fp(x), _fp(y)), "mlp_forward")
        return y

    def grad(self, x, t):
        """Loss and its gradient w.r.t. the parameters on normalised rows."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_float(0)
        self._chk(self.lib.mlp_grad(self.m, x.shape[0], _fp(x), _fp(t), C.byref(loss), _fp(g)), "mlp_grad")
        return loss.value, g
