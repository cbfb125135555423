"""The reference's fixed-state gates of the charm first and second moments, on the
oracle: they pin the oracle's restatement of consistent_fm_source_2d and
consistent_m2_source_2d (and through them sigma_nu_2d, pi_sigma_contractions_2d,
rank1_traceless_2d and vorticity_coupling_2d) against referees that do not share
its algebra:

- Gc1 / Gm1: the axisymmetric limit against the reference's radial source
  functions (src/hq_consistent_firstmoment.jl:104-124, src/hq_consistent_m2.jl:
  102-252), restated independently in tests/referees.py; the reference's points
  and thresholds (test/test_consistent_fm2d.jl:106-142, test_consistent_m22d.jl:
  127-200);
- Gc2 / Gm2: rotation covariance at the reference's state and four angles, with
  the default and with every term switched on (the default leaves the vorticity
  coupling off);
- Gc3-Gc4 / Gm4: the Bjorken limit and the zero source;
- Gd3 for every term: each term switched on alone against its covariant
  definition built by brute-force 4x4 index algebra in Milne coordinates, over
  300 random flowing states (the reference's Gd3, test/test_dnmr2d.jl:107-139,
  checks pi_sigma this way); and Gm3: the eta-eta component each term's block
  implies by tracelessness equals the brute-force one.
"""
import pytest

import source_gates as G


def test_Gc1_one_dimensional_limit(oracle):
    worst, sy = G.gate_Gc1(oracle)
    assert worst < 1e-12 and sy < 1e-14, (worst, sy)


@pytest.mark.parametrize("terms", [None, G.ALL_TERMS], ids=["default", "all"])
def test_Gc2_rotation(oracle, terms):
    assert G.gate_Gc2(oracle, terms) < 1e-12


def test_Gc3_Gc4_bjorken_and_zero(oracle):
    d, s4 = G.gate_Gc3_Gc4(oracle)
    assert d < 1e-13
    assert s4[0] == 0.0 and s4[1] == 0.0


def test_Gd3_first_moment_terms_brute_force(oracle):
    res = G.gate_Gd3_fm(oracle)
    assert max(res.values()) < 1e-12, res


def test_Gm1_one_dimensional_limit(oracle):
    worst, wxy = G.gate_Gm1(oracle)
    assert worst < 1e-10 and wxy < 1e-12, (worst, wxy)


@pytest.mark.parametrize("terms", [None, G.ALL_TERMS], ids=["default", "all"])
def test_Gm2_rotation_covariance(oracle, terms):
    assert G.gate_Gm2(oracle, terms) < 1e-10


def test_Gm4_bjorken_limit(oracle):
    d, dxy, dPiQ = G.gate_Gm4(oracle)
    assert d < 1e-14 and dxy < 1e-14
    assert dPiQ == dPiQ


def test_Gd3_Gm3_second_moment_terms_brute_force(oracle):
    res = G.gate_Gd3_m2(oracle)
    for name, (wb, wt, we) in res.items():
        assert wb < 1e-11 and wt < 1e-11 and we < 1e-10, (name, wb, wt, we)


def test_Gm2_each_term_rotates(oracle):
    """Rotation covariance of every second-moment term on its own, so a slot error
    in one term cannot hide behind the others' magnitude."""
    for name in G.M2_TERMS:
        assert G.gate_Gm2(oracle, 1 << G.BIT[name]) < 1e-10, name
