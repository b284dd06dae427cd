"""ctypes binding of libslx.so (the C-ABI declared in include/slx.h)."""
from __future__ import annotations

import ctypes as C
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libslx.so")
_lib = None

c_int, c_double, c_void_p = C.c_int, C.c_double, C.c_void_p
P_int = C.POINTER(C.c_int)
P_double = C.POINTER(C.c_double)


class SlxSolverParams(C.Structure):
    _fields_ = [("lmax", c_int), ("kmax", c_int), ("nmod2", c_int),
                ("tpw", c_int), ("checkmarine", c_int), ("calcRG", c_int), ("extrap", c_int),
                ("epsilon1", c_double),
                ("radius", c_double), ("mass", c_double), ("rhoi", c_double), ("rhow", c_double),
                ("gacc", c_double), ("omega", c_double), ("moiA", c_double), ("moiC", c_double), ("kf", c_double),
                ("kTE2", c_double), ("hTE2", c_double), ("kE2", c_double),
                ("nmod", P_int), ("s", P_double), ("a_sl", P_double), ("a_rr", P_double),
                ("a_il", P_double), ("a_m", P_double), ("a_tsl", P_double), ("a_trr", P_double),
                ("ttl", P_double), ("ttlekhl", P_double), ("ttlhe", P_double),
                ("ramp", c_int), ("a_hl", P_double), ("a_thl", P_double), ("ttlle", P_double), ("lTE2", c_double),
                ("anderson", c_int), ("anderson_depth", c_int), ("anderson_freeze", c_int),
                ("max_mask_refresh", c_int), ("anderson_keep", c_int),
                ("precision", c_int), ("eps_switch", c_double), ("max_inner", c_int)]


class SlxStatus(C.Structure):
    _fields_ = [("numiter", c_int), ("done", c_int), ("failed", c_int), ("nacc", c_int),
                ("reguess", c_int), ("ice_analyses", c_int), ("step", c_int), ("finished", c_int),
                ("xi", c_double), ("conserv", c_double),
                ("il", c_double * 9), ("mm", c_double * 3), ("lam", c_double * 18),
                ("ocean_area", c_double), ("oceanBeta_area", c_double), ("gmslc_ocn", c_double),
                ("gmslc_ocnBeta", c_double), ("ice_volume", c_double),
                ("ocean_area0", c_double), ("oceanBeta_area0", c_double), ("ice_volume0", c_double),
                ("aa_mixed", c_int), ("aa_plain", c_int), ("aa_fallback", c_int), ("aa_singular", c_int),
                ("aa_mask_refresh", c_int), ("aa_frozen_iters", c_int), ("sweeps32", c_int), ("pad_", c_int)]


class SlxFields(C.Structure):
    _fields_ = [(n, c_void_p) for n in ("icestar", "tinit", "cstar0", "cxy0", "beta0", "dS_converged", "p_load",
                                        "p_ice", "p_ocean", "prev_ice", "load_inc", "modal_R")]


class SlxOutputs(C.Structure):
    _fields_ = [(n, c_void_p) for n in ("tgrid", "ocean", "beta", "R", "G")]


class SlxReadouts(C.Structure):
    _fields_ = [(n, c_void_p) for n in ("dSLdt", "dRdt", "dSLdt_lm", "dRdt_lm", "rot_rates", "horiz_lm")]


# name -> (restype, argtypes)
SIGNATURES = {
    "slx_last_error": (C.c_char_p, []),
    "slx_abi_version": (c_int, []),
    "slx_sht_create": (c_int, [c_int, c_int, c_int, c_int, c_int, C.POINTER(c_void_p)]),
    "slx_sht_destroy": (c_int, [c_void_p]),
    "slx_sht_info": (c_int, [c_void_p, P_int, P_double, P_double]),
    "slx_sht_analysis": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "slx_sht_analysis_contraction": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "slx_sht_analysis_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "slx_sht_synthesis": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "slx_sht_tri_to_dense": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "slx_sht_dense_to_tri": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "slx_modal_advance_read": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "slx_sht_analysis_adjoint": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "slx_sht_synthesis_adjoint": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "slx_modal_reverse": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "slx_solver_create": (c_int, [c_void_p, C.POINTER(SlxSolverParams), C.POINTER(SlxOutputs), C.POINTER(c_void_p)]),
    "slx_solver_destroy": (c_int, [c_void_p]),
    "slx_solver_init": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "slx_solver_set_dt": (c_int, [c_void_p, c_double]),
    "slx_step_begin": (c_int, [c_void_p, c_void_p, c_double, c_double, c_double, c_void_p]),
    "slx_step_begin_warm": (c_int, [c_void_p, c_void_p, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p]),
    "slx_step_iterate": (c_int, [c_void_p, c_int, c_void_p]),
    "slx_step_finish": (c_int, [c_void_p, c_void_p]),
    "slx_step_finish_readouts": (c_int, [c_void_p, C.POINTER(SlxReadouts), c_void_p]),
    "slx_solver_status": (c_int, [c_void_p, C.POINTER(SlxStatus), c_void_p]),
    "slx_run_steps": (c_int, [c_void_p, c_void_p, C.c_long, c_int, P_double, P_double, P_int,
                              C.POINTER(SlxStatus), c_void_p]),
    "slx_solver_fields": (c_int, [c_void_p, C.POINTER(SlxFields)]),
    "slx_solver_profile": (c_int, [c_void_p, c_int]),
    "slx_solver_stage_times": (c_int, [c_void_p, C.POINTER(c_double)]),
}


PREC_FP64, PREC_FP32 = 0, 1
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_FP32, "fp32": PREC_FP32}


class SlxError(RuntimeError):
    pass


def lib_path() -> str:
    return _PATH


def load():
    """Load libslx.so (raises if it has not been built: there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_PATH):
        raise ImportError(f"{_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().slx_last_error().decode(errors="replace")
        raise SlxError(f"{what or 'libslx'} failed with code {rc}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
