"""ctypes mirror of include/fivo2d.h and the loader of the HIP library.

The structures here are byte-for-byte the C structs of the ABI. `load_hip()`
opens the in-tree libfivo2d.so built for gfx950 and fails loudly when it is
missing: there is no CPU fallback behind this package.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
HIP_LIB_PATH = os.path.join(PKG_ROOT, "csrc", "libfivo2d.so")

FIVO_OK = 0
FIVO_STEP_FAILED = 1
FIVO_EOS_CONFORMAL_HQ = 0
FIVO_EOS_LATTICE_HRG = 1
FIVO_BC_OUTFLOW = 0
FIVO_BC_PERIODIC = 1
FIVO_SSPRK2 = 2
FIVO_SSPRK3 = 3
FIVO_BATCH_NTILES = 4
BATCH_TILES = ((7, 8), (8, 28), (12, 20), (15, 16))   # candidate face tiles, FIVO_BATCH_NTILES of them

# Terms field order of src/terms.jl:82-106 -> bit index
TERM_NAMES = (
    "nu_gradalpha", "fm_gradT", "fm_inertial", "fm_nu_gradu", "fm_expansion", "fm_dlnh",
    "m2_nu_gradient", "m2_bg_gradu", "m2_bg_DlnT", "m2_bg_Dalpha", "m2_expansion",
    "m2_pi_sigma", "m2_PiQ_sigma", "m2_vorticity", "m2_projector", "m2_accel_nu",
    "m2_nu_gradTh", "shear_vorticity",
)
TERMS_DEFAULT = (0x3FFFF & ~((1 << 13) | (1 << 17)))


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nghost", C.c_int32), ("_pad", C.c_int32),
                ("xmax", C.c_double), ("ymax", C.c_double)]


class EosDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("hq_times_fmGeV3", C.c_int32),
                ("g_eff", C.c_double), ("m_hq", C.c_double), ("g_hq", C.c_double),
                ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double), ("a4", C.c_double),
                ("b1", C.c_double), ("b2", C.c_double), ("b3", C.c_double), ("b4", C.c_double),
                ("c", C.c_double), ("d", C.c_double), ("canon_factor", C.c_double)]


_MODEL_INT_FIELDS = ("with_charge", "enable_shear", "enable_bulk", "enable_diff",
                     "consistent_fm", "consistent_m2", "advect_nu", "relax_advect_nu",
                     "advect_Pi", "relax_advect_Pi", "advect_pi", "relax_advect_pi",
                     "relax_advect_m2", "vacuum_ramp_relax", "shear_projected_deriv",
                     "diff_projected_deriv", "shear_constraint_project")
_MODEL_DBL_FIELDS = ("eta_over_s", "tauShear_coeff", "deltaShear_factor", "taupi_pi_factor",
                     "pi_clip_factor", "zeta_over_s", "tauPi_coeff", "deltaPi_factor",
                     "lambda_Pi_pi_factor", "lambda_pi_Pi_factor", "Pi_clip_factor",
                     "kappa_coeff", "transport_mass", "tauN_coeff", "deltaN_factor",
                     "nu_clip_factor", "vacuum_n_lo", "vacuum_n_hi", "r_domain",
                     "T_vac_cut", "E_vac_cut")


class ModelDesc(C.Structure):
    _fields_ = ([("eos", EosDesc)] + [(f, C.c_int32) for f in _MODEL_INT_FIELDS] +
                [("terms", C.c_uint32)] + [(f, C.c_double) for f in _MODEL_DBL_FIELDS])


_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)


class WorkView(C.Structure):
    _fields_ = ([(f, _DP) for f in ("yT", "phi", "mu", "alpha", "ux", "uy", "P", "n", "e")] +
                [("ok", C.POINTER(C.c_uint8))] +
                [(f, _DP) for f in ("x0_yT", "x0_phi", "x0_ux", "x0_uy", "ux_prev", "uy_prev",
                                    "alpha_prev", "T_prev", "nux_prev", "nuy_prev", "shear_res")])


WORK_FIELDS = ("yT", "phi", "mu", "alpha", "ux", "uy", "P", "n", "e", "ok", "x0_yT", "x0_phi",
               "x0_ux", "x0_uy", "ux_prev", "uy_prev", "alpha_prev", "T_prev", "nux_prev",
               "nuy_prev", "shear_res")


class Counts(C.Structure):
    _fields_ = [("nprimfail", C.c_int64), ("nvacuum", C.c_int64), ("amax", C.c_double),
                ("nretry", C.c_int64), ("nstaged", C.c_int64)]


DUMP_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_void_p)


class RunOpts(C.Structure):
    _fields_ = [("tau0", C.c_double), ("taufinal", C.c_double), ("cfl", C.c_double),
                ("cfltau", C.c_double), ("dump_dt", C.c_double),
                ("integrator", C.c_int32), ("bc", C.c_int32),
                ("max_halve", C.c_int32), ("max_mood", C.c_int32),
                ("max_steps", C.c_int64), ("reset_history", C.c_int32), ("_pad", C.c_int32),
                ("on_dump", DUMP_CB), ("user", C.c_void_p)]


class RunResult(C.Structure):
    _fields_ = [("ok", C.c_int32), ("_pad", C.c_int32), ("tau", C.c_double),
                ("nsteps", C.c_int64), ("nprimfail", C.c_int64), ("nvacuum", C.c_int64),
                ("max_shear_res", C.c_double), ("maxu", C.c_double), ("dQ", C.c_double),
                ("Q0", C.c_double), ("Q1", C.c_double), ("minPtot", C.c_double),
                ("nrhs", C.c_int64), ("nhalvings", C.c_int64), ("nmood", C.c_int64),
                ("newton_iters", C.c_int64), ("nlearned", C.c_int64), ("nlearned_fallback", C.c_int64),
                ("nqueued", C.c_int64), ("nretry", C.c_int64), ("nstaged", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("_")}


class CfOpts(C.Structure):
    _fields_ = [("n_eta", C.c_int32), ("n_pt", C.c_int32), ("n_phi", C.c_int32), ("_pad", C.c_int32),
                ("eta_max", C.c_double), ("pt_lo", C.c_double), ("pt_hi", C.c_double),
                ("mass", C.c_double), ("yield_factor", C.c_double)]


# Entry points shared by the HIP library (prefix fivo_) and any library with the
# same ABI. (name, restype, argtypes)
_VP = C.c_void_p
ABI_FUNCS = (
    ("create", C.c_int, [C.POINTER(GridDesc), C.POINTER(ModelDesc), C.POINTER(C.c_void_p)]),
    ("destroy", None, [_VP]),
    ("nvars", C.c_int, [_VP]),
    ("ntot", C.c_int64, [_VP]),
    ("e_vac_cut", C.c_double, [_VP]),
    ("set_state", C.c_int, [_VP, _DP, C.c_int, C.c_int64]),
    ("get_state", C.c_int, [_VP, _DP, C.c_int, C.c_int64]),
    ("set_work", C.c_int, [_VP, C.POINTER(WorkView)]),
    ("get_work", C.c_int, [_VP, C.POINTER(WorkView)]),
    ("reset_history", C.c_int, [_VP]),
    ("init_from_prims", C.c_int, [_VP, C.c_double] + [_DP] * 10 + [C.c_int, C.POINTER(C.c_int64)]),
    ("finalize_ic", C.c_int, [_VP, C.c_double, C.c_int]),
    ("apply_bc", C.c_int, [_VP, C.c_int]),
    ("rhs", C.c_int, [_VP, C.c_double, C.c_int, C.POINTER(C.c_uint8), _DP, C.POINTER(Counts)]),
    ("update_primitives", C.c_int, [_VP, C.c_double, C.c_int, C.POINTER(Counts)]),
    ("background_slice", C.c_int, [_VP, C.c_int32, C.c_int32, _DP, _DP, _DP, _DP, C.c_int32, _DP]),
    ("sanitize", C.c_int, [_VP, C.c_double, C.c_int]),
    ("state_ok", C.c_int, [_VP, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    ("compute_dt", C.c_int, [_VP, C.c_double, C.c_double, C.c_double, _DP]),
    ("step", C.c_int, [_VP, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                       C.POINTER(C.c_int), _DP]),
    ("relax", C.c_int, [_VP, C.c_double, C.c_double]),
    ("project_shear", C.c_int, [_VP, _DP, _DP]),
    ("run", C.c_int, [_VP, C.POINTER(RunOpts), C.POINTER(RunResult)]),
    ("probe_fm_source", C.c_int, [C.c_uint32, C.c_int64, _DP, _DP]),
    ("probe_m2_source", C.c_int, [C.c_uint32, C.c_int64, _DP, _DP]),
    ("freezeout_create", C.c_int, [_VP, C.c_double, C.POINTER(C.c_void_p)]),
    ("freezeout_destroy", None, [_VP]),
    ("freezeout_record", C.c_int, [_VP, C.c_double]),
    ("freezeout_surface", C.c_int, [_VP, _DP, _DP, _DP, C.POINTER(C.c_int64)]),
    ("freezeout_set_surface", C.c_int, [_VP, _DP, _DP, _DP]),
    ("freezeout_pion_yield", C.c_int, [_VP, C.POINTER(CfOpts), _DP]),
)


class Lib:
    """A loaded library exposing the fivo2d ABI under `prefix` (e.g. 'fivo_')."""

    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise RuntimeError(f"fivo2d: library {path} is not built (run __graft_entry__.build())")
        self.path = path
        self.prefix = prefix
        self.dll = C.CDLL(path)
        for name, res, args in ABI_FUNCS:
            f = getattr(self.dll, prefix + name)
            f.restype = res
            f.argtypes = args
            setattr(self, name, f)
        le = getattr(self.dll, prefix + "last_error", None)
        if le is not None:
            le.restype = C.c_char_p
            le.argtypes = [_VP]
        self._last_error = le

    def last_error(self, ctx):
        if self._last_error is None:
            return ""
        s = self._last_error(ctx)
        return s.decode() if s else ""


_HIP = None


def load_hip():
    """Load the gfx950 HIP library. Raises if it is not built."""
    global _HIP
    if _HIP is None:
        lib = Lib(HIP_LIB_PATH, "fivo_")
        for name, res, args in (("kernel_time", C.c_int64, [C.c_void_p, C.c_int, _DP]),
                                ("enable_timing", C.c_int, [C.c_void_p, C.c_int]),
                                ("get_histograms", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
                                ("batch_create", C.c_int, [C.POINTER(GridDesc), C.c_int32, C.POINTER(ModelDesc),
                                                           C.POINTER(C.c_void_p)]),
                                ("batch_destroy", None, [C.c_void_p]),
                                ("batch_last_error", C.c_char_p, [C.c_void_p]),
                                ("batch_event", C.c_void_p, [C.c_void_p, C.c_int32]),
                                ("batch_autotune", C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int32),
                                                             C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
                                ("batch_set_tile", C.c_int, [C.c_void_p, C.c_int32]),
                                ("batch_run", C.c_int, [C.c_void_p, C.POINTER(RunOpts), _DP, C.POINTER(RunResult)]),
                                ("init_hotspots", C.c_int, [C.c_void_p, C.c_double, C.c_int32, _DP, C.c_double,
                                                            C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int64)]),
                                ("pion_yield_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                                                 C.c_int32, C.c_double, C.c_double, C.c_double,
                                                                 C.POINTER(CfOpts), C.c_void_p, C.c_void_p]),
                                ("mlp_create", C.c_int, [C.c_int64, C.c_uint64, C.POINTER(C.c_void_p)]),
                                ("mlp_destroy", None, [C.c_void_p]),
                                ("mlp_last_error", C.c_char_p, [C.c_void_p]),
                                ("mlp_collect", C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_int64)]),
                                ("mlp_size", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
                                ("mlp_fit_normalization", C.c_int, [C.c_void_p]),
                                ("mlp_train", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, _FP]),
                                ("mlp_get_params", C.c_int, [C.c_void_p, _FP, _FP]),
                                ("mlp_set_params", C.c_int, [C.c_void_p, _FP, _FP]),
                                ("mlp_forward", C.c_int, [C.c_void_p, C.c_int64, _FP, _FP]),
                                ("mlp_eval", C.c_int, [C.c_void_p, _FP]),
                                ("mlp_grad", C.c_int, [C.c_void_p, C.c_int64, _FP, _FP, _FP, _FP]),
                                ("set_learned_seed", C.c_int, [C.c_void_p, C.c_void_p]),
                                ("batch_set_learned_seed", C.c_int, [C.c_void_p, C.c_void_p])):
            f = getattr(lib.dll, "fivo_" + name)
            f.restype = res
            f.argtypes = args
            setattr(lib, name, f)
        _HIP = lib
    return _HIP
