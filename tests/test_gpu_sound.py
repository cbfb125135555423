"""Gate Gs on the HIP path, the reference's whole set (test/test_sound2d.jl:172-217):
ideal c_s (published 0.05 %) and damping floor; Gamma linear in eta/s with the
Navier-Stokes coefficient to 2 % at eta/s = 0.01, 0.02, 0.04, a clean exponential
(half-window split < 1 %); bulk attenuation on the Lorentzian at T = 0.175 and
0.210 within 0.85-1.25 of theory and the same ratio at both temperatures to 8 %;
and the ideal run's amplitude history equal to the oracle's."""
import pytest

import test_oracle_sound as S

pytestmark = pytest.mark.gpu


def test_sound_gate_on_gpu(oracle, hip):
    base = S.sound(hip, 0.35, 0.0, 0.0)
    dcs = abs(base["cs"] - base["cs_exact"]) / base["cs_exact"]
    print(f"ideal: c_s {base['cs']:.5f} vs {base['cs_exact']:.5f} ({100 * dcs:.3f} %), floor {base['G']:.2e}", flush=True)
    assert dcs < 0.02 and abs(base["G"]) < 5e-3 and round(100 * dcs, 2) <= 0.05
    ref = S.sound(oracle, 0.35, 0.0, 0.0)
    assert abs(base["cs"] - ref["cs"]) <= 1e-9 * ref["cs"] and abs(base["G"] - ref["G"]) <= 1e-9
    for es in (0.01, 0.02, 0.04):
        ratio, split = S.shear_ratio(S.sound(hip, 0.35, es, 0.0), base, es)
        print(f"eta/s {es}: Gamma / NS {ratio:.5f}, split {split:.4f}", flush=True)
        assert 0.98 < ratio < 1.02 and split < 0.01
    bl = []
    for T in (0.175, 0.210):
        b = S.sound(hip, T, 0.0, 0.0)
        r = S.sound(hip, T, 0.0, 0.05)
        Gp = 0.5 * r["k"] ** 2 * (0.05 * r["lorbar"] * r["s0"]) / r["eP_nat"]
        bl.append((r["G"] - b["G"]) / Gp)
        print(f"bulk T {T}: <lorentz> {r['lorbar']:.4f} ratio {bl[-1]:.3f}", flush=True)
    assert all(0.85 < x < 1.25 for x in bl) and abs(bl[0] - bl[1]) / bl[0] < 0.08
