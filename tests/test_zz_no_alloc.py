"""The entry point that allocates nothing, fivo_pion_yield_noalloc, called through
the C ABI on surfaces in device memory, against fivo_freezeout_pion_yield (the
tracker's entry point, same kernels) on the same surfaces: equal results on two
calls in a row, the second on a smaller surface, on the current stream and on a
stream of its own; oversized surfaces are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from fivo2d import _abi
from fivo2d import freezeout as FO

import freezeout_cases as FC

pytestmark = pytest.mark.gpu


def _surface(n, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(-9.0, 9.0, n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    t = 5.0 + np.cos(0.6 * X) * np.sin(0.4 * Y + seed) + 0.5 * X / 9.0
    keep = (np.hypot(X, Y) < 7.5 + 0.8 * np.sin(3 * np.arctan2(Y, X))) & (rng.random(X.shape) > 0.02)
    t = np.where(keep, t, np.nan)
    return t, 0.5 * np.tanh(X / 4), 0.5 * np.tanh(Y / 4) + 0.1


def _with_tracker(hip, n, t, ux, uy, grid):
    s, fo = FC.tracker(hip, n=n, box=9.0)
    fo.set_surface(t, ux, uy)
    y = fo.pion_yield(**grid)
    dx, dy = fo.dx, fo.dy
    s.close()
    return y, dx, dy


@pytest.mark.parametrize("use_own_stream", [False, True])
def test_noalloc_matches_the_tracker_entry_point(hip, use_own_stream):
    import torch
    dev = torch.device("cuda:0")
    out = torch.full((3,), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream() if use_own_stream else None
    for n, seed, grid in ((200, 1, {}), (64, 2, dict(n_eta=21, n_pt=11, n_phi=16))):
        t, ux, uy = _surface(n, seed)
        ref, dx, dy = _with_tracker(hip, n, t, ux, uy, grid)
        dt = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (t, ux, uy)]
        torch.cuda.synchronize()
        out.fill_(float("nan"))
        FO.pion_yield_device(*dt, dx, dy, out, stream=stream, lib=hip, **grid)
        torch.cuda.synchronize()
        got = tuple(float(v) for v in out.cpu().numpy())
        print(f"  n={n}: no-alloc {got} tracker {ref}", flush=True)
        assert got == ref


def test_noalloc_refuses_oversized_surfaces(hip):
    import torch
    dev = torch.device("cuda:0")
    one = torch.zeros(1, dtype=torch.float64, device=dev)
    out = torch.zeros(3, dtype=torch.float64, device=dev)
    o = _abi.CfOpts()
    o.n_eta, o.n_pt, o.n_phi, o.eta_max, o.pt_lo, o.pt_hi, o.mass, o.yield_factor = 40, 40, 24, 6.0, 0.1, 5.0, 0.138, 5.0
    p = C.c_void_p(one.data_ptr())
    rc = hip.pion_yield_noalloc(p, p, p, 1025, 1024, 0.1, 0.1, 0.156, C.byref(o), C.c_void_p(out.data_ptr()), None)
    assert rc == -1 and "exceeds the scratch limit" in hip.last_error(None)
    rc = hip.pion_yield_noalloc(p, p, p, 2, 50, 0.1, 0.1, 0.156, C.byref(o), C.c_void_p(out.data_ptr()), None)
    assert rc == -1
    o.n_eta = 1
    rc = hip.pion_yield_noalloc(p, p, p, 10, 10, 0.1, 0.1, 0.156, C.byref(o), C.c_void_p(out.data_ptr()), None)
    assert rc == -1
