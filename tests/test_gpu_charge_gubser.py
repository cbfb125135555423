"""Gate G3g on the HIP path at N = 100, 200, 400 (test/test_charge_gubser2d.jl:161-194):
core invariant n/T^3 < 1e-3 and L2(n) < 2e-3 at N = 400, orders 200 -> 400 above 1.5,
x<->y symmetric to 1e-11, charge loss < 3e-4; the published orders (invariant 1.97,
L2(n) 2.05, README2D.md:278) and the measured N = 400 numbers (invariant 5.20e-4,
L2 1.21e-3) to the digits printed; and at N = 100 the HIP run's density equals the
oracle's within 1e-9."""
import math

import numpy as np
import pytest

import test_oracle_charge_gubser as G

pytestmark = pytest.mark.gpu


def test_charge_on_gubser_gate_on_gpu(oracle, hip):
    rs = [G.run_charge(hip, N) for N in (100, 200, 400)]
    for N, r in zip((100, 200, 400), rs):
        print(f"N {N}: core inv {r['worst_c']:.3e} L2c {r['L2c']:.3e} x<->y {r['asym']:.1e} dQ {r['dQ']:.2e}",
              flush=True)
        assert r["nc"] > 100 and r["asym"] < 1e-11 and r["dQ"] < 3e-4
    oi = math.log2(rs[1]["worst_c"] / rs[2]["worst_c"])
    oc = math.log2(rs[1]["L2c"] / rs[2]["L2c"])
    print(f"orders 200 -> 400: invariant {oi:.2f}, L2(n) {oc:.2f}", flush=True)
    assert rs[2]["worst_c"] < 1e-3 and rs[2]["L2c"] < 2e-3 and oi > 1.5 and oc > 1.5
    assert f"{oi:.2f}" == "1.97" and f"{oc:.2f}" == "2.05"
    assert f"{rs[2]['worst_c']:.2e}" == "5.20e-04" and f"{rs[2]['L2c']:.2e}" == "1.21e-03"
    o = G.run_charge(oracle, 100)
    assert np.max(np.abs(rs[0]["n"] - o["n"]) / np.abs(o["n"])) <= 1e-9
