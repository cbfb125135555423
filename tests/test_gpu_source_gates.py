"""The charm first- and second-moment source gates on the HIP path, through the
library's probe entry points (fivo_probe_fm_source, fivo_probe_m2_source), which
run the device functions the relaxation kernels run: the reference's 1-D limits
(Gc1, Gm1), rotation covariance (Gc2, Gm2), the Bjorken limits (Gc3, Gm4), every
term against brute-force covariant algebra (Gd3 / Gm3), and equality with the
oracle at random flowing states for the default, an all-on and every single-term
switch set."""
import numpy as np
import pytest

import source_gates as G
from fivo2d import probes

pytestmark = pytest.mark.gpu


def test_gpu_fm_gates(hip):
    worst, sy = G.gate_Gc1(hip)
    assert worst < 1e-12 and sy < 1e-14, (worst, sy)
    assert G.gate_Gc2(hip) < 1e-12 and G.gate_Gc2(hip, G.ALL_TERMS) < 1e-12
    d, s4 = G.gate_Gc3_Gc4(hip)
    assert d < 1e-13 and s4[0] == 0.0 and s4[1] == 0.0
    res = G.gate_Gd3_fm(hip)
    assert max(res.values()) < 1e-12, res


def test_gpu_m2_gates(hip):
    worst, wxy = G.gate_Gm1(hip)
    assert worst < 1e-10 and wxy < 1e-12, (worst, wxy)
    for t in (None, G.ALL_TERMS):
        assert G.gate_Gm2(hip, t) < 1e-10
    d, dxy, _ = G.gate_Gm4(hip)
    assert d < 1e-14 and dxy < 1e-14
    res = G.gate_Gd3_m2(hip)
    for name, (wb, wt, we) in res.items():
        assert wb < 1e-11 and wt < 1e-11 and we < 1e-10, (name, wb, wt, we)


def test_gpu_sources_equal_oracle(hip, oracle):
    rows = G.m2_random_rows()
    frows = G.fm_random_rows()
    masks = [None, G.ALL_TERMS] + [1 << G.BIT[n] for n in G.M2_TERMS + list(G.FM_TERMS)]
    for t in masks:
        a, b = probes.m2_source(hip, rows, t), probes.m2_source(oracle, rows, t)
        scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
        assert np.max(np.abs(a - b) / scale) < 1e-12, t
        a, b = probes.fm_source(hip, frows, t), probes.fm_source(oracle, frows, t)
        scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
        assert np.max(np.abs(a - b) / scale) < 1e-12, t
