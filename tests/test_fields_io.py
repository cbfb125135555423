"""save_fields / load_fields (src/fields_io.jl): the output of a run saved with the
record of what produced it, read back without pickle.

Checked on a short oracle run (host logic; the HIP library writes the same
fields_2d arrays): every field round-trips exactly; the schema keys of
src/fields_io.jl:14-23 are there (fields, solver, terms, model, model_fields, meta,
created, fivo_revision); the terms are the model's resolved switches in register
order (under the "none" preset every term of an enabled sector off); the knobs read back
in Julia's spelling (0.08, true, Inf, 1.0e-5); a file without a model has empty model
records; values that would need pickle are refused.
"""
import datetime

import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic
from fivo2d.fields_io import _jrepr


def _run(oracle, terms="default"):
    g = F.make_grid2d(12, 10, xmax=4.0, ymax=3.5)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                         enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                         kappa_coeff=0.1163, pi_clip_factor=1.0, terms=terms)
    s = F.Solver(g, m, lib=oracle)
    s.init_from_prims(0.6, **ic.hotspot_event(g, seed=5, nspots=20, sigma=0.5, Tpeak=0.4))
    r = s.run_sim_2d(tau0=0.6, taufinal=0.7, CFL=0.2, CFLtau=0.05, integrator="ssprk2")
    assert r["ok"]
    f = s.fields_2d()
    s.close()
    return m, f


def test_round_trip_and_schema(oracle, tmp_path):
    m, f = _run(oracle)
    path = str(tmp_path / "out" / "run.npz")
    assert F.save_fields(path, f, model=m, meta={"event": 3, "note": "vorticity on", "b": 1.5}) == path
    d = F.load_fields(path)
    assert set(d) == {"fields", "solver", "terms", "model", "model_fields", "meta", "created", "fivo_revision"}
    assert set(d["fields"]) == set(f)
    for k in f:
        assert np.array_equal(d["fields"][k], f[k]) and d["fields"][k].dtype == np.asarray(f[k]).dtype, k
    assert d["solver"] == "FiVo 2+1D"
    names = [t[0] for t in F.TERM_REGISTER]
    assert list(d["terms"]) == names
    assert d["terms"] == {n: bool((m.terms >> i) & 1) for i, n in enumerate(names)}
    assert any(d["terms"].values())
    mf = d["model_fields"]
    assert mf["eta_over_s"] == "0.08" and mf["enable_shear"] == "true" and mf["eos"] == "LatticeHRGEOS"
    assert all(isinstance(v, str) for v in mf.values())
    assert "[x]" in d["model"] and "LatticeHRGEOS" in d["model"] and "eta_over_s = 0.08" in d["model"]
    assert d["meta"] == {"event": 3, "note": "vorticity on", "b": 1.5}
    datetime.datetime.fromisoformat(d["created"])
    assert isinstance(d["fivo_revision"], str) and d["fivo_revision"]
    with np.load(path, allow_pickle=False) as z:          # every member loads without pickle
        assert all(z[k].dtype != object for k in z.files)


def test_terms_none_and_no_model(oracle, tmp_path):
    # "none" turns off every term of an enabled sector; the terms of the disabled
    # consistent_fm / consistent_m2 sectors keep their defaults (resolve_terms)
    m, f = _run(oracle, terms="none")
    d = F.load_fields(F.save_fields(str(tmp_path / "a.npz"), f, model=m))
    on = {"enable_shear", "enable_bulk", "enable_diff"}
    assert all(not v for (n, sec, _), v in zip(F.TERM_REGISTER, d["terms"].values()) if sec in on)
    assert d["terms"] == {n: bool((m.terms >> i) & 1) for i, (n, _, _) in enumerate(F.TERM_REGISTER)}
    assert d["meta"] == {}
    d = F.load_fields(F.save_fields(str(tmp_path / "b.npz"), f))
    assert d["solver"] == "FiVo" and d["model"] == "" and d["terms"] is None and d["model_fields"] is None


def test_julia_repr_of_knobs_and_refusals(tmp_path):
    assert [_jrepr(v) for v in (True, False, 0.08, 1.0, 1e-5, float("inf"), float("nan"), 7)] == \
        ["true", "false", "0.08", "1.0", "1.0e-5", "Inf", "NaN", "7"]
    with pytest.raises(TypeError):
        F.save_fields(str(tmp_path / "c.npz"), {"T": np.zeros(3)}, meta={"x": {"nested": 1}})
    with pytest.raises(ValueError):
        F.save_fields(str(tmp_path / "d.npz"), {"a/b": np.zeros(3)})
    with pytest.raises(TypeError):
        F.save_fields(str(tmp_path / "e.npz"), [np.zeros(3)])
