"""save_fields / load_fields (src/fields_io.jl): a run's output saved together with
what produced it, and read back.

The reference writes a JLD2 (HDF5) file. No HDF5 writer exists in this environment,
so the file is a NumPy .npz holding the reference's schema keys, flattened one level
(src/fields_io.jl:14-23):

    fields/<name>        the arrays fields_2d returned
    solver               "FiVo 2+1D" (or the caller's string; "FiVo" without a model)
    terms/<name>         the resolved term switches, one bool per term of the register
    model                a printout of the model: EOS, sectors, every term [x]/[ ], every knob
    model_fields/<name>  every scalar knob of the model, spelled as Julia's repr spells
                         booleans, infinities and exponents; "eos", "eos.<field>"
    meta/<key>           what the caller passed as `meta`
    created              ISO-8601 time stamp
    fivo_revision        git short hash of this package's tree (+ "-dirty"), or "unknown"

Strings are 0-d unicode arrays; nothing in the file needs pickle to load.
"""
import dataclasses
import datetime
import math
import os
import subprocess

import numpy as np

from .model import TERM_REGISTER

SOLVER_2D = "FiVo 2+1D"


def _jrepr(v):
    """A scalar knob spelled as Julia's repr spells it where that differs from
    Python's: true/false, Inf, NaN, a mantissa with a point before the exponent
    (1.0e-5). Other floats keep Python's shortest round-trip digits."""
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        x = float(v)
        if math.isnan(x):
            return "NaN"
        if math.isinf(x):
            return "Inf" if x > 0 else "-Inf"
        s = repr(x)
        if "e" in s:
            mant, ex = s.split("e")
            if "." not in mant:
                mant += ".0"
            s = f"{mant}e{int(ex)}"
        return s
    if isinstance(v, str):
        return repr(v).replace("'", '"')
    return repr(v)


def terms_dict(model):
    """The model's resolved term switches as {name: bool}, in register order
    (terms_namedtuple, src/fields_io.jl:30)."""
    return {name: bool((model.terms >> i) & 1) for i, (name, _, _) in enumerate(TERM_REGISTER)}


def _model_fields(model):
    """Every scalar knob of the model, repr'd, plus the EOS and its fields
    (_model_record, src/fields_io.jl:47-64)."""
    out = {k: _jrepr(v) for k, v in sorted(model.params.items())
           if isinstance(v, (bool, int, float, str, np.bool_, np.integer, np.floating))}
    out["eos"] = type(model.eos).__name__
    if dataclasses.is_dataclass(model.eos):
        for f in dataclasses.fields(model.eos):
            v = getattr(model.eos, f.name)
            if isinstance(v, (bool, int, float, str, np.bool_, np.integer, np.floating)):
                out["eos." + f.name] = _jrepr(v)
    return out


def model_printout(model):
    """A text record of the model: EOS, state layout, sectors, every term [x]/[ ] with
    its sector, every knob."""
    p = model.params
    lines = [f"FiVo 2+1D model, EOS {type(model.eos).__name__}",
             "  state rows: " + ", ".join(model.layout.names)]
    sectors = [k for k in ("enable_shear", "enable_bulk", "enable_diff", "consistent_fm", "consistent_m2") if k in p]
    lines.append("  sectors: " + "  ".join(f"{k} [{'x' if p[k] else ' '}]" for k in sectors))
    lines.append("  terms:")
    on = terms_dict(model)
    for name, sector, ingredients in TERM_REGISTER:
        lines.append(f"    [{'x' if on[name] else ' '}] {name:16s} ({sector}; {', '.join(ingredients)})")
    lines.append("  knobs:")
    for k, v in _model_fields(model).items():
        lines.append(f"    {k} = {v}")
    return "\n".join(lines) + "\n"


def _revision():
    """git short hash of this package's tree, "-dirty" with uncommitted changes
    (_fivo_revision, src/fields_io.jl:34-44)."""
    here = os.path.dirname(os.path.abspath(__file__))
    try:
        rev = subprocess.run(["git", "-C", here, "rev-parse", "--short", "HEAD"], capture_output=True, text=True,
                             timeout=10)
        if rev.returncode != 0:
            return "unknown"
        st = subprocess.run(["git", "-C", here, "status", "--porcelain", "--untracked-files=no"],
                            capture_output=True, text=True, timeout=10)
        r = rev.stdout.strip()
        return r + "-dirty" if st.stdout.strip() else r
    except (OSError, subprocess.SubprocessError):
        return "unknown"


def _check_key(k):
    if not isinstance(k, str) or not k or "/" in k:
        raise ValueError(f"save_fields: key {k!r} must be a non-empty string without '/'")
    return k


def _plain_value(v):
    a = np.asarray(v)
    if a.dtype == object:
        raise TypeError(f"save_fields: {type(v).__name__} values cannot be stored without pickle")
    return a


def save_fields(path, fields, model=None, meta=None, solver=None):
    """Write `fields` (what fields_2d returned: a dict of arrays) to `path` (.npz)
    with the model's term switches, its printout, every scalar knob, the git
    revision and a time stamp (save_fields, src/fields_io.jl:66-93). Returns path."""
    if not isinstance(fields, dict):
        raise TypeError("save_fields: fields must be a dict of arrays (fields_2d's result)")
    solver = solver if solver is not None else (SOLVER_2D if model is not None else "FiVo")
    out = {"solver": np.array(str(solver)), "created": np.array(datetime.datetime.now().isoformat()),
           "fivo_revision": np.array(_revision())}
    for k, v in fields.items():
        out["fields/" + _check_key(k)] = _plain_value(v)
    if model is not None:
        out["model"] = np.array(model_printout(model))
        for k, v in terms_dict(model).items():
            out["terms/" + k] = np.array(v)
        for k, v in _model_fields(model).items():
            out["model_fields/" + k] = np.array(v)
    else:
        out["model"] = np.array("")
    for k, v in (meta or {}).items():
        out["meta/" + _check_key(str(k))] = _plain_value(v)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        np.savez(f, **out)
    return path


def load_fields(path):
    """Read a file written by save_fields (load_fields, src/fields_io.jl:95-109):
    a dict with fields, solver, terms, model, model_fields, meta, created,
    fivo_revision (None where the file has no such key). No pickle."""
    res = {"fields": None, "solver": None, "terms": None, "model": None, "model_fields": None, "meta": None,
           "created": None, "fivo_revision": None}
    with np.load(path, allow_pickle=False) as z:
        for key in z.files:
            a = z[key]
            v = a.item() if a.ndim == 0 else a
            if "/" in key:
                group, name = key.split("/", 1)
                if res.get(group) is None:
                    res[group] = {}
                res[group][name] = v
            else:
                res[key] = v
    if res["meta"] is None:
        res["meta"] = {}
    return res
