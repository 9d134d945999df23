"""ctypes binding of libmcs.so (declarations in include/mcs.h).

The shared library is built in-tree by ``python -m mcstreamer_amd.build``.  There
is no fallback: if the library is missing or fails to load, importing any
compute entry point raises, so a GPU run can never silently take a CPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcs.so")

MCS_OK = 0
MCS_E_ARG = 1
MCS_E_HIP = 2
MCS_E_CAPACITY = 3
MCS_E_FIELD = 4
MCS_E_TABLE_RANGE = 5
MCS_E_NONCONVERGED = 6
MCS_E_NAN = 7

PROC_NULL, PROC_ELASTIC, PROC_EXCITATION, PROC_IONIZATION, PROC_ATTACHMENT, PROC_PHOTOEMISSION = range(6)
MAX_PROC = 32
EGRID_UNIFORM, EGRID_LOG = 0, 1
FIELD_UNIFORM, FIELD_MESH = 0, 1
ARR_QFIXED, ARR_QPART, ARR_Q, ARR_U, ARR_EX, ARR_EY, ARR_EZ = range(7)
DEPOSIT_NGP, DEPOSIT_CIC = 0, 1
BC_DIRICHLET, BC_NEUMANN = 0, 1
PREC_F64, PREC_F32 = 0, 1
MAJORANT_GLOBAL, MAJORANT_BANDED = 0, 1

_ERRNAMES = {
    MCS_E_ARG: "invalid argument",
    MCS_E_HIP: "HIP runtime error",
    MCS_E_CAPACITY: "capacity exceeded",
    MCS_E_FIELD: "field limit exceeded",
    MCS_E_TABLE_RANGE: "energy beyond the collision table",
    MCS_E_NONCONVERGED: "multigrid did not converge",
    MCS_E_NAN: "NaN encountered",
}


class MCSError(RuntimeError):
    """A libmcs entry point returned a non-zero status."""

    def __init__(self, code, message):
        super().__init__(f"[{_ERRNAMES.get(code, code)}] {message}")
        self.code = code


class Field(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("e", ctypes.c_double * 3), ("mesh", ctypes.c_void_p)]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int64),
        ("n_out", ctypes.c_int64),
        ("births", ctypes.c_int64),
        ("deaths", ctypes.c_int64),
        ("flights", ctypes.c_int64),
        ("events", ctypes.c_int64 * MAX_PROC),
        ("flags", ctypes.c_int32),
    ]


class Moments(ctypes.Structure):
    _fields_ = [
        ("nparticles", ctypes.c_int64),
        ("nactive", ctypes.c_int64),
        ("weight", ctypes.c_double),
        ("wenergy", ctypes.c_double),
        ("maxenergy", ctypes.c_double),
        ("sum_x", ctypes.c_double * 3),
        ("sum_x2", ctypes.c_double * 3),
        ("sum_v", ctypes.c_double * 3),
        ("sum_energy", ctypes.c_double),
        ("collrate", ctypes.c_double * MAX_PROC),
    ]


class ResampleStats(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int64),
        ("n_out", ctypes.c_int64),
        ("cells_merged", ctypes.c_int64),
        ("clones", ctypes.c_int64),
        ("splits", ctypes.c_int64),
        ("removed", ctypes.c_int64),
    ]


class MGConfigC(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32 * 3),
        ("bc", ctypes.c_int32 * 6),
        ("s", ctypes.c_double),
        ("levels", ctypes.c_int32),
        ("smooth1", ctypes.c_int32),
        ("smooth2", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("maxiter", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("cyl", ctypes.c_int32),
    ]


TALLY_NQ = 4   # v_z, energy, ionization rate, attachment rate


class SwarmParams(ctypes.Structure):
    _fields_ = [
        ("e", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("n0", ctypes.c_int64),
        ("tol", ctypes.c_double),
        ("capacity", ctypes.c_int64),
        ("instance", ctypes.c_uint32),
        ("relax_steps", ctypes.c_int32),
        ("block_steps", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("init_vth", ctypes.c_double),
    ]


class SwarmTally(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("nactive", ctypes.c_int64),
        ("nrecorded", ctypes.c_int64),
        ("nblocks", ctypes.c_int64),
        ("flags", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("sum", ctypes.c_double * TALLY_NQ),
        ("bm_s1", ctypes.c_double * TALLY_NQ),
        ("bm_s2", ctypes.c_double * TALLY_NQ),
        ("last", ctypes.c_double * TALLY_NQ),
        ("diffusion", ctypes.c_double * 3),
        ("window", ctypes.c_int64),
    ]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PU8 = ctypes.POINTER(ctypes.c_uint8)
_PU64 = ctypes.POINTER(ctypes.c_uint64)
_PP = ctypes.POINTER(ctypes.c_void_p)

# name -> argument types (all return int32 status except mcs_last_error)
SIGNATURES = {
    "mcs_ctx_create": [_I32, _PP],
    "mcs_ctx_destroy": [_P],
    "mcs_ctx_synchronize": [_P],
    "mcs_ctx_stream": [_P, _PP],
    "mcs_ctx_profile": [_P, _I32],
    "mcs_ctx_profile_read": [_P, ctypes.c_char_p, _PD, _PI64],
    "mcs_ctx_profile_reset": [_P],
    "mcs_table_create": [_P, _I32, _I32, _I32, _PD, _PD, _PI32, _PD, _D, _PP],
    "mcs_table_destroy": [_P],
    "mcs_pop_create": [_P, _I64, _PP],
    "mcs_pop_create_split": [_P, _I64, _P, _PP],
    "mcs_pop_destroy": [_P],
    "mcs_pop_upload": [_P, _I64, _PD, _PD, _PD, _PD, _PU8, _PU64, _U64],
    "mcs_pop_download": [_P, _I64, _PI64, _PD, _PD, _PD, _PD, _PU8, _PU64, _PU64],
    "mcs_pop_size": [_P, _PI64],
    "mcs_pop_moments": [_P, _P, ctypes.POINTER(Moments)],
    "mcs_pop_moments_noalloc": [_P, _P, ctypes.POINTER(Moments)],
    "mcs_pop_energies": [_P, _I64, _PI64, _PD],
    "mcs_advance": [_P, _P, ctypes.POINTER(Field), _D, _P, _U64, _U32, _U32, ctypes.POINTER(StepStats)],
    "mcs_repack": [_P],
    "mcs_swarm_resample": [_P, _I64, _D, _U64, _U32, _U32, _PI32],
    "mcs_mesh_create": [_P, _PI32, _D, _PP],
    "mcs_mesh_destroy": [_P],
    "mcs_mesh_set": [_P, _I32, _PD],
    "mcs_mesh_get": [_P, _I32, _PD],
    "mcs_mesh_zero": [_P, _I32],
    "mcs_sort_by_cell": [_P, _P],
    "mcs_deposit": [_P, _P, _I32, _I32, _D],
    "mcs_assemble_charge": [_P],
    "mcs_efield_faces": [_P, _D],
    "mcs_resample_cells": [_P, _P, _I32, _I32, _U64, _U32, _U32, ctypes.POINTER(ResampleStats)],
    "mcs_mg_create": [_P, ctypes.POINTER(MGConfigC), _PP],
    "mcs_mg_destroy": [_P],
    "mcs_mg_solve_host": [_P, _PD, _PD, _PI32, _PD],
    "mcs_mg_solve_mesh": [_P, _P, _PI32, _PD],
    "mcs_mg_solve_device": [_P, _P, _P, _PI32, _PD],
    "mcs_mg_setinhom_mesh": [_P, _P, _I32, _PD],
    "mcs_swarm_batch_create": [_P, _I64, _I32, _U64, _I32, _PP],
    "mcs_swarm_batch_destroy": [_P],
    "mcs_swarm_batch_add": [_P, _P, ctypes.POINTER(SwarmParams), _I64, _PD, _PD, _PD, _PD, _PU8, _PI32],
    "mcs_swarm_batch_remove": [_P, _I32],
    "mcs_swarm_batch_step": [_P, _I32],
    "mcs_swarm_batch_tally": [_P, _I32, ctypes.POINTER(SwarmTally)],
    "mcs_swarm_batch_tally_all": [_P, _I32, ctypes.POINTER(SwarmTally)],
    "mcs_swarm_batch_download": [_P, _I32, _I64, _PI64, _PD, _PD, _PD, _PD, _PU8, _PU64],
    "mcs_swarm_batch_config": [_P, _I32, _I32],
    "mcs_swarm_batch_majorant": [_P, _I32],
    "mcs_table_majorants": [_P, _I32, _PI32, _PD, _PD],
    "mcs_pop_init_thermal": [_P, _I64, _U64, _U32, _D],
    "mcs_cgrid_create": [_P, _D, _D, _I32, _I32, _PP],
    "mcs_cgrid_destroy": [_P],
    "mcs_cgrid_steps": [_P, _PD],
    "mcs_cgrid_set": [_P, _I32, _PD],
    "mcs_cgrid_get": [_P, _I32, _PD],
    "mcs_cgrid_zero": [_P, _I32],
    "mcs_cyl_density": [_P, _P, _I32, _D],
    "mcs_cyl_poisson": [_P, _P, _D, _I32, _PI32, _PD],
    "mcs_cgrid_freebc_coeffs": [_P, _PD],
    "mcs_cyl_sort_by_cell": [_P, _P],
    "mcs_cyl_resample": [_P, _P, _I32, _I32, _U64, _U32, _U32, ctypes.POINTER(ResampleStats)],
    "mcs_cgrid_resample_fields": [_P, _PI64, _PD],
    "mcs_cyl_efield_faces": [_P, _D],
    "mcs_advance_cyl": [_P, _P, _P, _D, _I32, _U64, _U32, _U32, ctypes.POINTER(StepStats)],
    "mcs_table_set_photoemission": [_P, _I32, _D, _D, _D, _D],
    "mcs_pop_create_photons": [_P, _I64, _PP],
    "mcs_photons_upload": [_P, _I64, _PD, _PD, _PD, _PD, _PD, _PU8, _PU64, _U64],
    "mcs_photons_download": [_P, _I64, _PI64, _PD, _PD, _PD, _PD, _PD, _PU8, _PU64, _PU64],
    "mcs_advance_mixed": [_P, _P, _P, _D, ctypes.POINTER(Field), _D, _P, _U64, _U32, _U32,
                          ctypes.POINTER(StepStats), ctypes.POINTER(StepStats)],
    "mcs_advance_cyl_mixed": [_P, _P, _P, _D, _P, _D, _I32, _U64, _U32, _U32, ctypes.POINTER(StepStats),
                              ctypes.POINTER(StepStats)],
}

_lib = None


def lib():
    """Load libmcs.so once; raise if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m mcstreamer_amd.build` "
            "(there is no CPU fallback for the particle step)"
        )
    L = ctypes.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int32
    L.mcs_last_error.argtypes = [_P]
    L.mcs_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(ctx_handle, rc):
    if rc != MCS_OK:
        msg = lib().mcs_last_error(ctx_handle) if ctx_handle else b""
        raise MCSError(rc, (msg or b"").decode(errors="replace"))
    return rc


def dptr(a):
    return a.ctypes.data_as(_PD)


def i32ptr(a):
    return a.ctypes.data_as(_PI32)


def u8ptr(a):
    return a.ctypes.data_as(_PU8)


def u64ptr(a):
    return a.ctypes.data_as(_PU64)
