"""Gate G1v on the HIP path at the reference's resolutions: N = 100, 200, 400 for
eta/s = 0.005 and 0.02, tau 1 -> 2 (test/test_gubser_viscous2d.jl:224-241): L2(T) <
5e-3 and L2(pibar) < 8e-2 at N = 400, orders 200 -> 400 above 1.0 (T) and 0.7
(pibar), x<->y symmetric to 1e-11; and at N = 100 the HIP run lands on the
oracle's within 1e-9 (T) and 1e-7 of max |pi^eta_eta|."""
import math

import numpy as np
import pytest

import test_oracle_viscous_gubser as V

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("es", [0.005, 0.02])
def test_viscous_gubser_gate_on_gpu(oracle, hip, es):
    rs = [V.run_visc(hip, N, es) for N in (100, 200, 400)]
    for N, r in zip((100, 200, 400), rs):
        print(f"N {N} eta/s {es}: L2(T) {r['L2T']:.3e} L2(pibar) {r['L2p']:.3e} x<->y {r['asym']:.1e}", flush=True)
        assert r["nc"] > 100 and r["asym"] < 1e-11
    oT = math.log2(rs[1]["L2T"] / rs[2]["L2T"])
    op = math.log2(rs[1]["L2p"] / rs[2]["L2p"])
    print(f"orders 200 -> 400: T {oT:.2f} pibar {op:.2f}", flush=True)
    assert rs[2]["L2T"] < 5e-3 and rs[2]["L2p"] < 8e-2 and oT > 1.0 and op > 0.7
    o = V.run_visc(oracle, 100, es)
    assert np.max(np.abs(rs[0]["T"] - o["T"]) / o["T"]) <= 1e-9
    assert np.max(np.abs(rs[0]["pieta"] - o["pieta"])) <= 1e-7 * np.max(np.abs(o["pieta"]))
