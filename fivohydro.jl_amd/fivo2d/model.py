"""Host-side mirror of FiVoHydro.jl's 2+1D construction API.

Names, keyword arguments and defaults follow the reference so a caller of
`hydro2d` finds the same vocabulary:

  make_grid2d        src2d/grid2d.jl:52-74
  ConformalHQEOS     src/eos.jl:39-46
  LatticeHRGEOS      src/eos.jl:318-347
  build_model_2d     main2D.jl:85-123 (+ IdealDiffVisc2DModel defaults,
                     src2d/primitives2d.jl:41-210, and reject_unwired_knobs_2d,
                     src2d/primitives2d.jl:246-272)
  make_layout2d      src2d/state_layout2d.jl:61-84
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi

HBARC = 0.1973269804
T_MIN = 1e-20
E_FLOOR = 1e-20
# I1(N/2)/I0(N/2) at N = 21.55 (src/eos.jl:346-348), evaluated to 25 digits.
LHRG_CANON_FACTOR = 0.9524014587885936


@dataclass
class Grid2D:
    """Cell-centred box [-xmax,xmax] x [-ymax,ymax] with nghost ghost layers
    (src2d/grid2d.jl:19-61). Flat index i = ix*Nytot + iy (0-based, y fastest)."""
    Nx: int
    Ny: int
    xmax: float
    ymax: float
    nghost: int = 3

    @property
    def xmin(self):
        return -self.xmax

    @property
    def ymin(self):
        return -self.ymax

    @property
    def dx(self):
        return (self.xmax - self.xmin) / self.Nx

    @property
    def dy(self):
        return (self.ymax - self.ymin) / self.Ny

    @property
    def Nxtot(self):
        return self.Nx + 2 * self.nghost

    @property
    def Nytot(self):
        return self.Ny + 2 * self.nghost

    @property
    def Ntot(self):
        return self.Nxtot * self.Nytot

    @property
    def xC(self):
        i = np.arange(1, self.Nxtot + 1)
        return self.xmin + (i - self.nghost - 0.5) * self.dx

    @property
    def yC(self):
        j = np.arange(1, self.Nytot + 1)
        return self.ymin + (j - self.nghost - 0.5) * self.dy

    def lin(self, ix, iy):
        return ix * self.Nytot + iy

    def interior(self):
        """Slices (ix, iy) of the interior on an (Nxtot, Nytot) array."""
        ng = self.nghost
        return slice(ng, ng + self.Nx), slice(ng, ng + self.Ny)

    def desc(self):
        return _abi.GridDesc(self.Nx, self.Ny, self.nghost, 0, self.xmax, self.ymax)


def make_grid2d(Nx, Ny, xmax=20.0, ymax=20.0, nghost=3):
    """make_grid2d (src2d/grid2d.jl:52-74)."""
    assert Nx > 8 and Ny > 8 and xmax > 0 and ymax > 0
    if nghost != 3:
        raise ValueError("the 2+1D solver uses nghost = 3")
    return Grid2D(Nx, Ny, float(xmax), float(ymax), nghost)


make_grid_2d = make_grid2d


@dataclass
class ConformalHQEOS:
    """src/eos.jl:39-46."""
    g_eff: float = 40.0
    m_hq: float = 1.5
    g_hq: float = 6.0
    hq_times_fmGeV3: bool = False

    def desc(self):
        d = _abi.EosDesc()
        d.kind = _abi.FIVO_EOS_CONFORMAL_HQ
        d.hq_times_fmGeV3 = int(self.hq_times_fmGeV3)
        d.g_eff, d.m_hq, d.g_hq = self.g_eff, self.m_hq, self.g_hq
        return d


@dataclass
class LatticeHRGEOS:
    """src/eos.jl:318-347 (light-sector fit + charm Boltzmann density)."""
    a1: float = -15.526548963383643
    a2: float = 18.6159584620131
    a3: float = -10.731808109698516
    a4: float = 2.7413302179949346
    b1: float = -3.3147904483107595
    b2: float = 5.310983721567554
    b3: float = -4.653922019495976
    b4: float = 1.8600649533271152
    c: float = -1.0465330501411811
    d: float = 0.09551531822245873
    m_hq: float = 1.5
    g_hq: float = 6.0
    hq_times_fmGeV3: bool = True
    canon_factor: float = LHRG_CANON_FACTOR

    def desc(self):
        d = _abi.EosDesc()
        d.kind = _abi.FIVO_EOS_LATTICE_HRG
        d.hq_times_fmGeV3 = int(self.hq_times_fmGeV3)
        d.m_hq, d.g_hq, d.canon_factor = self.m_hq, self.g_hq, self.canon_factor
        for k in ("a1", "a2", "a3", "a4", "b1", "b2", "b3", "b4", "c", "d"):
            setattr(d, k, getattr(self, k))
        return d


@dataclass
class StateLayout2D:
    """make_layout2d (src2d/state_layout2d.jl:61-84)."""
    names: list

    @property
    def nv(self):
        return len(self.names)

    def idx(self, name):
        return self.names.index(name)

    def has(self, name):
        return name in self.names


def make_layout2d(with_charge=True, with_bulk=True, with_shear=True, with_m2=False):
    names = ["Dtau", "Sx", "Sy", "E"]
    if with_charge:
        names += ["nux", "nuy"]
    if with_bulk:
        names += ["Pi"]
    if with_shear:
        names += ["pixx", "pixy", "piyy", "pieta"]
    if with_m2:
        names += ["pQxx", "pQxy", "pQyy", "pQeta", "PiQ"]
    return StateLayout2D(names)


# Knobs the 2-D reference declares but never reads (src2d/primitives2d.jl:237-242)
_UNWIRED = (("diffusion_drive", "alpha"), ("diff_dt_coeff", 0.0),
            ("shear_dt_coeff", 0.0), ("bulk_dt_coeff", 0.0))


@dataclass
class IdealDiffVisc2DModel:
    eos: object
    layout: StateLayout2D
    params: dict = field(default_factory=dict)
    terms: int = _abi.TERMS_DEFAULT

    def __getattr__(self, k):
        p = self.__dict__.get("params", {})
        if k in p:
            return p[k]
        raise AttributeError(k)

    def desc(self):
        m = _abi.ModelDesc()
        m.eos = self.eos.desc()
        for f in _abi._MODEL_INT_FIELDS:
            setattr(m, f, int(bool(self.params[f])))
        for f in _abi._MODEL_DBL_FIELDS:
            setattr(m, f, float(self.params[f]))
        m.terms = self.terms
        return m


# ---- per-term switches (src/terms.jl) ------------------------------------------------
# (name, sector, ingredients) of every term, in the order of the Terms fields
# (src/terms.jl:114-133); a term's bit in the mask is its index here.
TERM_REGISTER = (
    ("nu_gradalpha", "enable_diff", ("fugacity_gradient",)),
    ("fm_gradT", "consistent_fm", ("temperature_gradient",)),
    ("fm_inertial", "consistent_fm", ("acceleration",)),
    ("fm_nu_gradu", "consistent_fm", ("velocity_gradient",)),
    ("fm_expansion", "consistent_fm", ("expansion",)),
    ("fm_dlnh", "consistent_fm", ("cooling",)),
    ("m2_nu_gradient", "consistent_m2", ("current_gradient",)),
    ("m2_bg_gradu", "consistent_m2", ("shear", "expansion")),
    ("m2_bg_DlnT", "consistent_m2", ("cooling",)),
    ("m2_bg_Dalpha", "consistent_m2", ("fugacity_rate",)),
    ("m2_expansion", "consistent_m2", ("expansion", "cooling")),
    ("m2_pi_sigma", "consistent_m2", ("shear",)),
    ("m2_PiQ_sigma", "consistent_m2", ("shear",)),
    ("m2_vorticity", "consistent_m2", ("vorticity",)),
    ("m2_projector", "consistent_m2", ("acceleration",)),
    ("m2_accel_nu", "consistent_m2", ("acceleration",)),
    ("m2_nu_gradTh", "consistent_m2", ("temperature_gradient",)),
    ("shear_vorticity", "enable_shear", ("vorticity",)),
)
INGREDIENTS = ("fugacity_gradient", "temperature_gradient", "acceleration", "shear", "expansion", "vorticity",
               "velocity_gradient", "cooling", "fugacity_rate", "current_gradient")
INGREDIENT_GROUPS = {
    "inertia": ("acceleration",),
    "velocity_gradient": ("velocity_gradient", "shear", "expansion", "vorticity"),
    "medium_gradients": ("temperature_gradient", "acceleration", "velocity_gradient", "shear", "expansion",
                         "vorticity", "cooling"),
}
TERM_PRESETS = ("default", "full", "homogeneous", "none")


def _term_default(k):
    return bool((_abi.TERMS_DEFAULT >> k) & 1)


def _expand_ingredients(xs):
    out = []
    for x in ((xs,) if isinstance(xs, str) else xs):
        x = x.lstrip(":")
        if x in INGREDIENT_GROUPS:
            out += [y for y in INGREDIENT_GROUPS[x] if y not in out]
        elif x in INGREDIENTS:
            if x not in out:
                out.append(x)
        else:
            raise ValueError(f"Terms: unknown ingredient :{x}. Valid: " + ", ".join(":" + i for i in INGREDIENTS)
                             + "; groups: " + ", ".join(":" + g for g in INGREDIENT_GROUPS))
    return out


def _preset_values(p):
    """terms_preset (src/terms.jl:184-197)."""
    p = p.lstrip(":")
    n = len(TERM_REGISTER)
    if p == "default":
        return [_term_default(k) for k in range(n)]
    if p == "full":
        return [True] * n
    if p == "none":
        return [False] * n
    if p == "homogeneous":
        ingr = _expand_ingredients(("medium_gradients",))
        return [_term_default(k) and not any(i in ingr for i in TERM_REGISTER[k][2]) for k in range(n)]
    raise ValueError(f"Terms: unknown preset :{p}. Valid: " + ", ".join(":" + q for q in TERM_PRESETS))


def without(*ingredients, from_="default"):
    """without(ingredients...; from = :default) (src/terms.jl:199-209): a terms spec
    that drops every term built from any of `ingredients`."""
    return {"preset": from_, "without": tuple(ingredients)}


def resolve_terms(spec="default", sector_on=None, **overrides):
    """resolve_terms (src/terms.jl:211-262): the bit mask of the `Terms` switches
    for a spec: an int mask (taken as is), a preset name ('default', 'full',
    'homogeneous', 'none'), or a dict with optional 'preset', 'without', 'with'
    and term names (applied in that order, explicit names last). A change that
    came from a preset or an ingredient in a disabled sector is reset to the
    default; an explicit name is kept, so that check_terms refuses it."""
    if isinstance(spec, int):
        return spec
    if spec is None:
        spec = "default"
    nt = dict(spec) if isinstance(spec, dict) else {"preset": spec}
    nt.update(overrides)
    names = [t[0] for t in TERM_REGISTER]
    ctrl = ("preset", "without", "with")
    bad = [k for k in nt if k not in names and k not in ctrl]
    if bad:
        raise ValueError(f"Terms: unknown term(s) {bad}. Valid names: " + ", ".join(names)
                         + " — plus `preset`, `without`, `with`.")
    vals = _preset_values(nt.get("preset", "default"))
    for key, on in (("without", False), ("with", True)):
        if key in nt:
            ingr = _expand_ingredients(nt[key])
            for k, (_, _, ing) in enumerate(TERM_REGISTER):
                if any(i in ingr for i in ing):
                    vals[k] = on
    if sector_on is not None:
        for k, (_, sector, _) in enumerate(TERM_REGISTER):
            if not sector_on(sector):
                vals[k] = _term_default(k)
    for key, v in nt.items():
        if key in ctrl:
            continue
        if not isinstance(v, bool):
            raise ValueError(f"Terms: `{key}` must be a Bool, got {v!r}")
        vals[names.index(key)] = v
    return sum(1 << k for k, v in enumerate(vals) if v)


def check_terms(mask, sector_on, solver="build_model_2d"):
    """check_terms (src/terms.jl:272-283): refuse a switch whose sector is disabled."""
    for k, (name, sector, _) in enumerate(TERM_REGISTER):
        v = bool((mask >> k) & 1)
        if v == _term_default(k) or sector_on(sector):
            continue
        raise ValueError(f"{solver}: `{name} = {str(v).lower()}` changes a term of the `{sector}` sector, but "
                         f"`{sector} = false`, so the switch would act on nothing. Enable `{sector}`, or leave "
                         f"`{name}` at its default ({str(_term_default(k)).lower()}).")


def build_model_2d(eos=None, r_domain=math.inf, enable_shear=False, enable_bulk=False,
                   enable_diff=False, with_charge=True, eta_over_s=0.1, tauShear_coeff=0.2,
                   zeta_over_s=0.1, tauPi_coeff=15.0, terms="default", **kw):
    """build_model_2d (main2D.jl:85-123) with the IdealDiffVisc2DModel defaults
    (src2d/primitives2d.jl:41-210). Unwired knobs are refused like
    reject_unwired_knobs_2d (src2d/primitives2d.jl:246-272)."""
    if eos is None:
        eos = LatticeHRGEOS()
    for knob, default in _UNWIRED:
        if knob in kw and kw[knob] != default:
            raise ValueError(f"IdealDiffVisc2DModel: `{knob} = {kw[knob]}` is NOT implemented in "
                             f"the 2-D solver. Use the default ({default}).")
        kw.pop(knob, None)
    p = dict(
        with_charge=with_charge, enable_shear=enable_shear, enable_bulk=enable_bulk,
        enable_diff=enable_diff, consistent_fm=False, consistent_m2=False,
        advect_nu=False, relax_advect_nu=True, advect_Pi=False, relax_advect_Pi=True,
        advect_pi=False, relax_advect_pi=True, relax_advect_m2=True, vacuum_ramp_relax=True,
        shear_projected_deriv=True, diff_projected_deriv=True, shear_constraint_project=True,
        eta_over_s=eta_over_s if enable_shear else 0.0,
        tauShear_coeff=tauShear_coeff if enable_shear else 0.0,
        deltaShear_factor=4.0 / 3.0, taupi_pi_factor=0.0, pi_clip_factor=-1.0,
        zeta_over_s=zeta_over_s if enable_bulk else 0.0,
        tauPi_coeff=tauPi_coeff if enable_bulk else 0.0,
        deltaPi_factor=0.0, lambda_Pi_pi_factor=0.0, lambda_pi_Pi_factor=0.0, Pi_clip_factor=-1.0,
        kappa_coeff=0.0, transport_mass=0.0, tauN_coeff=1.0, deltaN_factor=0.0,
        nu_clip_factor=-1.0, vacuum_n_lo=1e-6, vacuum_n_hi=2e-3,
        r_domain=r_domain, T_vac_cut=0.05, E_vac_cut=-1.0,
    )
    if "shear_constraint" in kw:
        sc = kw.pop("shear_constraint")
        p["shear_constraint_project"] = (sc in ("project", ":project"))
    for k, v in kw.items():
        if k not in p:
            raise ValueError(f"build_model_2d: unknown keyword {k!r}")
        p[k] = v
    for a, b in (("advect_nu", "relax_advect_nu"), ("advect_Pi", "relax_advect_Pi"),
                 ("advect_pi", "relax_advect_pi")):
        if p[a] and p[b]:
            raise ValueError(f"Pick ONE for transport: {a} OR {b} (not both).")
    if p["consistent_fm"] and p["tauN_coeff"] != 1.0:
        raise ValueError("`tauN_coeff` cannot be combined with `consistent_fm = true`.")
    if p["consistent_m2"] and not enable_diff:
        raise ValueError("consistent_m2 requires enable_diff (src2d/dissipation2d.jl:339-342).")
    layout = make_layout2d(with_charge=with_charge, with_bulk=enable_bulk,
                           with_shear=enable_shear, with_m2=p["consistent_m2"])
    # terms2d / sector_on_2d (main2D.jl:117-121): the sectors this model carries
    sectors = {"enable_diff": bool(enable_diff), "enable_shear": bool(enable_shear),
               "consistent_fm": bool(p["consistent_fm"]), "consistent_m2": bool(p["consistent_m2"])}
    tmask = resolve_terms(terms, sector_on=sectors.get)
    check_terms(tmask, sectors.get)
    return IdealDiffVisc2DModel(eos=eos, layout=layout, params=p, terms=tmask)


def allocate_state(g, model):
    """allocate_state (main2D.jl:125-126): zeros, reference layout U[a + nv*i],
    returned as an (Ntot, nv) C-ordered array so U[i, a] is element a of cell i."""
    return np.zeros((g.Ntot, model.layout.nv), dtype=np.float64)
