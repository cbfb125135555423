"""The learned warm start's lifetime rule on the host side: a solver or ensemble
seeded by a network keeps that network alive, and the network refuses to free
its device buffers while any context still reads them. Runs without a GPU on a
stand-in for the library's C entry points that records what is called."""
import gc

import pytest

import fivo2d as F
from fivo2d import batch as B
from fivo2d import mlp as M
from fivo2d import solver as S


class _Ptr:
    def __init__(self, v):
        self.value = v


class FakeLib:
    """Records create/destroy and seed calls; every call succeeds."""

    def __init__(self):
        self.destroyed = []
        self.seeds = []
        self.next = 100

    def mlp_create(self, cap, seed, ref):
        ref._obj.value = self.next
        self.next += 1
        return 0

    def mlp_destroy(self, m):
        self.destroyed.append(m.value)

    def set_learned_seed(self, ctx, m):
        self.seeds.append(("ctx", ctx.value, None if m is None else m.value))
        return 0

    def batch_set_learned_seed(self, b, m):
        self.seeds.append(("batch", b.value, None if m is None else m.value))
        return 0

    def destroy(self, ctx):
        pass

    def batch_destroy(self, b):
        pass


def _solver(lib, ctx=1):
    s = S.Solver.__new__(S.Solver)
    s.lib, s.ctx, s._owned = lib, _Ptr(ctx), True
    return s


def _ensemble(lib, nev=2):
    e = B.Ensemble.__new__(B.Ensemble)
    e.lib, e.b = lib, _Ptr(7)
    e.events = [_solver(lib, 10 + k) for k in range(nev)]
    for s in e.events:
        s._owned = False
    return e


def test_solver_holds_network_and_close_refuses(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(M.C, "byref", lambda o: type("R", (), {"_obj": o})())
    net = M.Mlp(capacity=4, lib=lib)
    s = _solver(lib)
    s.set_learned_seed(net)
    assert net.attached() == [s]
    with pytest.raises(RuntimeError, match="still seeds"):
        net.close()
    assert lib.destroyed == []
    # the solver's reference keeps the network alive without any other holder
    mval = net.m.value
    del net
    gc.collect()
    assert lib.destroyed == []
    assert s._mlp.m.value == mval
    # detaching releases it; close then frees the buffers
    net = s._mlp
    s.set_learned_seed(None)
    assert net.attached() == []
    net.close()
    assert lib.destroyed == [mval]
    with pytest.raises(F.FivoError, match="closed"):
        s.set_learned_seed(net)


def test_closing_the_solver_detaches(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(M.C, "byref", lambda o: type("R", (), {"_obj": o})())
    net = M.Mlp(capacity=4, lib=lib)
    s = _solver(lib)
    s.set_learned_seed(net)
    s.close()
    assert net.attached() == []
    net.close()
    assert len(lib.destroyed) == 1


def test_ensemble_holds_network(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(M.C, "byref", lambda o: type("R", (), {"_obj": o})())
    net = M.Mlp(capacity=4, lib=lib)
    e = _ensemble(lib)
    e.set_learned_seed(net)
    assert lib.seeds[-1] == ("batch", 7, net.m.value)
    assert all(s._mlp is net for s in e.events)
    with pytest.raises(RuntimeError, match="still seeds"):
        net.close()
    e.close()
    assert net.attached() == []
    net.close()
    assert len(lib.destroyed) == 1
