"""ctypes binding of ``libmantlepar_amd.so`` (the C ABI in include/mantlepar_amd.h).

The library is loaded from this package directory only; a missing build is an
error, never a silent fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmantlepar_amd.so")

MP_OK = 0
STATUS_NAMES = {1: "MP_ERR_ARG", 2: "MP_ERR_FP", 3: "MP_ERR_STATE", 4: "MP_ERR_PHASE",
                5: "MP_ERR_DEVICE", 6: "MP_ERR_MEMORY", 7: "MP_ERR_BACKFLOW",
                8: "MP_ERR_COMPOSITION", 9: "MP_ERR_PENDING", 10: "MP_ERR_BALANCE"}
MP_ERR_PENDING = 9
BATCH_EAGER, BATCH_GRAPH = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
BC_ZERO_FLUX, BC_OUTFLOW, BC_INFLOW_OUTFLOW = 0, 1, 2
SCHEME_FE, SCHEME_SSPRK2, SCHEME_SSPRK3 = 1, 2, 3
LAUNCH_PER_STAGE, LAUNCH_PER_STEP, LAUNCH_STALL_CHECK, LAUNCH_AUTO = 0, 1, 2, 3


class MantleError(RuntimeError):
    """A nonzero mp_status, with the library's message."""

    def __init__(self, code, message):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("q", C.c_int32)]


class WenoOptions(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("s", C.c_int32), ("large_weight", C.c_double),
                ("small_weight", C.c_double * 4), ("use_constant", C.c_int32),
                ("constant_linear_weight", C.c_double), ("lf_mode", C.c_int32),
                ("global_speed", C.c_double)]


class PhaseParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("Tm0", "Te0", "nu", "L", "cp", "Xe", "rho", "rhor", "mus", "mul", "k0", "g",
                 "alpha0", "pressure")] + [("hypereutectic", C.c_int32)]


class StepReport(C.Structure):
    _fields_ = [("dt", C.c_double), ("max_rate", C.c_double), ("courant", C.c_double),
                ("total_H", C.c_double), ("total_C", C.c_double), ("boundary_H", C.c_double),
                ("boundary_C", C.c_double), ("error_bits", C.c_uint32),
                ("steps_done", C.c_int32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [("mesh", MeshDesc), ("H", C.c_void_p), ("C", C.c_void_p), ("vnx", C.c_void_p),
                ("vny", C.c_void_p), ("bc", C.c_void_p * 4), ("cfl", C.c_double),
                ("nsteps", C.c_int32), ("H_out", C.c_void_p), ("C_out", C.c_void_p),
                ("T", C.c_void_p), ("phi", C.c_void_p), ("Cl", C.c_void_p), ("Cs", C.c_void_p),
                ("region", C.c_void_p), ("lithostatic", C.c_int32), ("surface_y", C.c_double),
                ("limiter", C.c_int32), ("cell_points_per_axis", C.c_int32),
                ("diffusion_samples", C.c_int32), ("extent_fraction", C.c_double),
                ("thermal_diffusivity", C.c_double), ("thermal_types", C.c_void_p),
                ("thermal_values", C.c_void_p)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("error_bits", C.c_uint32),
                ("steps_done", C.c_int32), ("dt", C.c_double), ("total_H", C.c_double),
                ("total_C", C.c_double), ("boundary_H", C.c_double), ("boundary_C", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int64), ("resident", C.c_int64), ("cell_steps", C.c_double),
                ("elapsed_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("hbm_in_use", C.c_uint64),
                ("hbm_peak", C.c_uint64), ("hbm_budget", C.c_uint64), ("epochs", C.c_int64),
                ("max_resident", C.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class HandoffMeta(C.Structure):
    _fields_ = [("family", C.c_char * 64), ("nx", C.c_int32), ("ny", C.c_int32),
                ("domain", C.c_double * 4), ("perturbation", C.c_double), ("seed", C.c_uint64),
                ("enthalpy_Jkg", C.c_double), ("length_m", C.c_double),
                ("time_scale_s", C.c_double), ("time", C.c_double), ("accepted_steps", C.c_int64)]


class EdgeSample(C.Structure):
    _fields_ = [("edge_id", C.c_int64), ("q", C.c_int32), ("vx", C.c_double), ("vy", C.c_double)]


class CoupledParams(C.Structure):
    _fields_ = [("theta", C.c_double), ("cooling", C.c_double),
                ("thermal_diffusivity", C.c_double), ("cell_points_per_axis", C.c_int32),
                ("diffusion_samples", C.c_int32), ("extent_fraction", C.c_double),
                ("lf_mode", C.c_int32), ("global_speed", C.c_double)]


class CoupledRates(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("boundary_H", "boundary_C", "source_H",
                                          "maximum_rate", "divergence")]


class CoupledTime(C.Structure):
    _fields_ = [("cfl", C.c_double), ("initial_step", C.c_double), ("maximum_step", C.c_double),
                ("minimum_step", C.c_double), ("end_time", C.c_double), ("fixed", C.c_int32)]


class CoupledStep(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "courant", "balance_H", "balance_C", "boundary_H",
                                          "boundary_C", "source_H", "maximum_rate")] + \
               [("retries", C.c_int32), ("pad", C.c_int32)]


_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

# (name, restype, argtypes) for every entry point of include/mantlepar_amd.h
SIGNATURES = [
    ("mp_default_weno_options", None, [C.POINTER(WenoOptions)]),
    ("mp_default_phase_params", None, [C.POINTER(PhaseParams)]),
    ("mp_last_error", C.c_char_p, []),
    ("mp_version", C.c_char_p, []),
    ("mp_report_error", C.c_int32, [C.c_int32, C.c_char_p]),
    ("mp_transport_create", C.c_int32, [C.POINTER(MeshDesc), C.POINTER(WenoOptions),
                                        C.POINTER(PhaseParams), C.c_int32, _vp, C.POINTER(_vp)]),
    ("mp_transport_destroy", C.c_int32, [_vp]),
    ("mp_transport_upload_state", C.c_int32, [_vp, _dp, _dp]),
    ("mp_transport_download_state", C.c_int32, [_vp, _dp, _dp]),
    ("mp_transport_set_face_velocity", C.c_int32, [_vp, _dp, _dp]),
    ("mp_transport_set_boundary", C.c_int32, [_vp, _u8p, _u8p, _u8p, _u8p]),
    ("mp_transport_set_inflow", C.c_int32, [_vp, _dp, _dp]),
    ("mp_transport_stable_dt", C.c_int32, [_vp, C.c_double, _dp]),
    ("mp_transport_step", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_double,
                                      C.POINTER(StepReport)]),
    ("mp_transport_step_noalloc", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_double,
                                              C.POINTER(StepReport)]),
    ("mp_transport_phase", C.c_int32, [_vp, _dp, _dp, _dp, _dp, _u8p]),
    ("mp_transport_rhs", C.c_int32, [_vp, _dp, _dp, _dp]),
    ("mp_transport_reconstruct", C.c_int32, [_vp, C.c_int32, _dp, _dp, _dp]),
    ("mp_transport_evaluate", C.c_int32, [_vp, C.c_int64, C.c_void_p, _dp, _dp, _dp, _dp]),
    ("mp_transport_device_state", C.c_int32, [_vp, C.POINTER(_dp), C.POINTER(_dp)]),
    ("mp_transport_synchronize", C.c_int32, [_vp]),
    ("mp_transport_profile", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_double, _dp, _dp, _dp]),
    ("mp_transport_set_tile", C.c_int32, [_vp, C.c_int32]),
    ("mp_transport_set_balance_tolerance", C.c_int32, [_vp, C.c_double]),
    ("mp_transport_set_launch", C.c_int32, [_vp, C.c_int32]),
    ("mp_transport_layout", C.c_int32, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("mp_transport_set_limiter", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_double]),
    ("mp_transport_limits", C.c_int32, [_vp, _dp, _dp]),
    ("mp_transport_set_diffusion", C.c_int32, [_vp, C.c_double, C.c_int32, C.c_double]),
    ("mp_default_coupled_params", None, [C.POINTER(CoupledParams)]),
    ("mp_coupled_create", C.c_int32, [C.POINTER(MeshDesc), C.POINTER(PhaseParams),
                                      C.POINTER(CoupledParams), _vp, C.POINTER(_vp)]),
    ("mp_coupled_destroy", C.c_int32, [_vp]),
    ("mp_coupled_set_flow", C.c_int32, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("mp_coupled_set_boundary", C.c_int32, [_vp, C.c_int32, _u8p, _u8p, _u8p, _u8p]),
    ("mp_coupled_set_inflow", C.c_int32, [_vp, _dp, _dp]),
    ("mp_coupled_set_thermal_boundary", C.c_int32, [_vp, _u8p, _dp]),
    ("mp_coupled_upload_state", C.c_int32, [_vp, _dp, _dp]),
    ("mp_coupled_download_state", C.c_int32, [_vp, _dp, _dp]),
    ("mp_coupled_rhs", C.c_int32, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(CoupledRates)]),
    ("mp_coupled_run", C.c_int32, [_vp, C.POINTER(CoupledTime), _dp, C.c_int32,
                                   C.POINTER(CoupledStep), C.POINTER(C.c_int32)]),
    ("mp_coupled_run_batch", C.c_int32, [C.POINTER(_vp), C.c_int32, C.POINTER(CoupledTime), _dp, C.c_int32,
                                         C.POINTER(CoupledStep), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]),
    ("mp_transport_set_thermal_boundary", C.c_int32, [_vp, _u8p, _dp]),
    ("mp_transport_set_lithostatic", C.c_int32, [_vp, C.c_int32, C.c_double]),
    ("mp_coupled_set_lithostatic", C.c_int32, [_vp, C.c_int32, C.c_double]),
    ("mp_batch_set_lithostatic", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_double]),
    ("mp_batch_set_limiter", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double]),
    ("mp_batch_set_diffusion", C.c_int32, [_vp, C.c_int32, C.c_double, C.c_int32, C.c_double]),
    ("mp_batch_set_thermal_boundary", C.c_int32, [_vp, C.c_int32, _vp, _vp]),
    ("mp_batch_create", C.c_int32, [C.c_int32, C.POINTER(WenoOptions), C.POINTER(PhaseParams),
                                    C.c_int32, C.c_uint64, _vp, C.POINTER(_vp)]),
    ("mp_batch_destroy", C.c_int32, [_vp]),
    ("mp_batch_add_mesh", C.c_int32, [_vp, C.POINTER(MeshDesc), C.POINTER(C.c_int32)]),
    ("mp_batch_remove_mesh", C.c_int32, [_vp, C.c_int32]),
    ("mp_batch_upload_state", C.c_int32, [_vp, C.c_int32, _dp, _dp]),
    ("mp_batch_download_state", C.c_int32, [_vp, C.c_int32, _dp, _dp]),
    ("mp_batch_set_face_velocity", C.c_int32, [_vp, C.c_int32, _dp, _dp]),
    ("mp_batch_set_boundary", C.c_int32, [_vp, C.c_int32, _u8p, _u8p, _u8p, _u8p]),
    ("mp_batch_set_inflow", C.c_int32, [_vp, C.c_int32, _dp, _dp]),
    ("mp_batch_stable_dt", C.c_int32, [_vp, C.c_int32, C.c_double, _dp]),
    ("mp_batch_set_dt", C.c_int32, [_vp, C.c_int32, C.c_double]),
    ("mp_batch_step", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(StepReport)]),
    ("mp_batch_phase", C.c_int32, [_vp, C.c_int32, _dp, _dp, _dp, _dp, _u8p]),
    ("mp_batch_set_schedule", C.c_int32, [_vp, C.c_int32, C.c_int32]),
    ("mp_batch_get_schedule", C.c_int32, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("mp_batch_set_launch", C.c_int32, [_vp, C.c_int32]),
    ("mp_batch_set_balance_tolerance", C.c_int32, [_vp, C.c_double]),
    ("mp_batch_get_launch", C.c_int32, [_vp, C.POINTER(C.c_int32)]),
    ("mp_batch_autotune", C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_char_p,
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32), _dp]),
    ("mp_batch_profile", C.c_int32, [_vp, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    ("mp_batch_info", C.c_int32, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
    ("mp_server_create", C.c_int32, [C.c_int32, C.POINTER(WenoOptions), C.POINTER(PhaseParams),
                                     C.c_int32, C.c_int32, C.c_uint64, C.c_int32,
                                     C.POINTER(_vp)]),
    ("mp_server_submit", C.c_int32, [_vp, C.POINTER(Request), C.POINTER(C.c_int64)]),
    ("mp_server_wait", C.c_int32, [_vp, C.c_int64, C.c_double, C.POINTER(RequestResult)]),
    ("mp_server_drain", C.c_int32, [_vp]),
    ("mp_server_stats_get", C.c_int32, [_vp, C.POINTER(ServerStats)]),
    ("mp_server_set_tile", C.c_int32, [_vp, C.c_int32]),
    ("mp_server_destroy", C.c_int32, [_vp]),
    ("mp_server_now", C.c_double, [_vp]),
    ("mp_server_progress", C.c_int32, [_vp, _dp, _dp, C.c_int32, C.POINTER(C.c_int32)]),
    ("mp_handoff_write", C.c_int32, [C.c_char_p, C.POINTER(HandoffMeta), _dp, _dp]),
    ("mp_handoff_read", C.c_int32, [C.c_char_p, C.POINTER(HandoffMeta), C.c_double, _dp, _dp,
                                    C.POINTER(HandoffMeta)]),
    ("mp_face_velocity_from_samples", C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                                                  C.c_void_p, _dp, _dp]),
    ("mp_phase_evaluate", C.c_int32, [C.POINTER(PhaseParams), C.c_int64, _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp, _u8p, _dp, _dp, _dp, _dp]),
    ("mp_phase_boundaries", C.c_int32, [C.POINTER(PhaseParams), C.c_double, C.c_double, _dp,
                                        _dp, _dp, _dp]),
    ("mp_phase_derive", C.c_int32, [C.POINTER(PhaseParams), _dp]),
    ("mp_phase_static_pressure", C.c_int32, [C.POINTER(PhaseParams), C.c_double, C.c_double,
                                             C.c_int32, _dp]),
    ("mp_phase_pressure_corrected_temperature", C.c_int32, [C.POINTER(PhaseParams), C.c_double,
                                                            C.c_double, _dp]),
]

_lib = None


def lib():
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m mantlepar_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != MP_OK:
        raise MantleError(code, lib().mp_last_error().decode())


def dptr(a):
    """double* of a C-contiguous float64 array (or NULL for None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def u8ptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)
