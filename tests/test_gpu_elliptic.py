"""Gates G6 and G5 on the HIP path (protocols and measurements in
test_oracle_elliptic.py): the elliptic profile at N = 150 and 300 against the
published anisotropies 0.20786 and 0.20798 (README2D.md:287), the eps2 = 0 control,
equality with the oracle at N = 150, and the round production profile with every
sector to late time at the reference's two resolutions."""
import pytest

from test_oracle_elliptic import check_elliptic, check_production, run_elliptic, run_production

pytestmark = pytest.mark.gpu


def test_g6_anisotropy_converges_to_the_published_values(hip):
    a = run_elliptic(hip, 150)
    b = run_elliptic(hip, 300)
    for o in (a, b):
        check_elliptic(o)
    d1 = abs(b["aniso"] - a["aniso"]) / abs(a["aniso"])
    print(f"  anisotropy {a['aniso']:.5f} -> {b['aniso']:.5f}  (change {100 * d1:.2f}%)", flush=True)
    assert a["aniso"] > 0.15 and d1 < 0.02
    assert f"{a['aniso']:.5f}" == "0.20786" and f"{b['aniso']:.5f}" == "0.20798"
    assert f"{100 * d1:.2f}" == "0.06"


def test_g6_control_is_zero(hip):
    c = run_elliptic(hip, 200, eps2=0.0)
    print(f"  eps2 = 0 control: anisotropy = {c['aniso']:+.3e} (published -2.7e-15)", flush=True)
    assert abs(c["aniso"]) < 1e-10 and c["refx_fo"] < 1e-11 and c["refy_fo"] < 1e-11


def test_g6_hip_equals_oracle(hip, oracle):
    a = run_elliptic(hip, 150)
    b = run_elliptic(oracle, 150)
    print(f"  N=150: HIP {a['aniso']:.15f} oracle {b['aniso']:.15f}", flush=True)
    assert a["nsteps"] == b["nsteps"] and a["nprimfail"] == b["nprimfail"]
    assert abs(a["aniso"] - b["aniso"]) < 1e-9 * abs(b["aniso"])


@pytest.mark.parametrize("N,tauf", [(150, 8.0), (300, 4.0)])
def test_g5_production_all_sectors(hip, N, tauf):
    check_production(run_production(hip, N, tauf))
