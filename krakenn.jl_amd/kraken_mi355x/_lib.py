"""ctypes binding of libkraken_mi355x.so (the C ABI in include/kraken_mi355x.h).

The library is loaded from this directory only.  If it is missing the import of a
solver entry point raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint8,
                    c_uint64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkraken_mi355x.so")

# --- enums (mirror include/kraken_mi355x.h) ---------------------------------------------
KMI_OK, KMI_ERR_ARG, KMI_ERR_DEVICE, KMI_ERR_STATE, KMI_ERR_NOMEM, KMI_ERR_UNSUPPORTED = range(6)
D2Q9, D3Q19 = 0, 1
STORE_F64, STORE_F32, STORE_F32S, STORE_F16S = 0, 1, 2, 3
BGK, TRT, MRT = 0, 1, 2
(FACE_PERIODIC, FACE_WALL, FACE_MOVING_WALL, FACE_ZOUHE_VELOCITY, FACE_ZOUHE_PRESSURE,
 FACE_ZOUHE_VELOCITY_PROFILE, FACE_ZOUHE_PRESSURE_PROFILE, FACE_ZOUHE_VELOCITY_REBUILD,
 FACE_ZOUHE_PRESSURE_REBUILD) = range(9)
FORCE_NONE, FORCE_GUO, FORCE_GUO_PLAIN_MOMENTS, FORCE_GUO_FIELD = 0, 1, 2, 3
WEST, EAST, SOUTH, NORTH, BOTTOM, TOP = range(6)
THERMAL_ADIABATIC, THERMAL_DIRICHLET = 0, 1
COUPLING_SPLIT, COUPLING_FUSED = 0, 1
FIELD_F, FIELD_G = 0, 1
HOST_F64, HOST_F32, HOST_STORAGE = 0, 1, 2
(RED_MASS, RED_MOMENTUM, RED_KINETIC_ENERGY, RED_THERMAL, RED_NUSSELT, RED_MAX_SPEED, RED_DRAG_MEA,
 RED_DRAG_HALFWAY) = range(8)
SOLID_BOUNCE_BACK, SOLID_INERT = 0, 1

STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f32s": STORE_F32S, "f16s": STORE_F16S}
COLLISION_NAMES = {"bgk": BGK, "trt": TRT, "mrt": MRT}
FACE_NAMES = {"west": WEST, "east": EAST, "south": SOUTH, "north": NORTH, "bottom": BOTTOM, "top": TOP}


class KmiFace(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32),
        ("thermal", c_int32),
        ("u", c_double * 3),
        ("rho", c_double),
        ("T", c_double),
    ]


class KmiDesc(ctypes.Structure):
    _fields_ = [
        ("stencil", c_int32),
        ("storage", c_int32),
        ("collision", c_int32),
        ("thermal", c_int32),
        ("nx", c_int64),
        ("ny", c_int64),
        ("nz", c_int64),
        ("omega", c_double),
        ("s_plus", c_double),
        ("s_minus", c_double),
        ("s_e", c_double),
        ("s_eps", c_double),
        ("s_q", c_double),
        ("s_nu", c_double),
        ("omega_T", c_double),
        ("beta_g", c_double),
        ("T_ref", c_double),
        ("coupling", c_int32),
        ("n_bc_order", c_int32),
        ("bc_order", c_int32 * 6),
        ("faces", KmiFace * 6),
        ("forcing", c_int32),
        ("force", c_double * 3),
        ("solid_mode", c_int32),
    ]


class KmiBatchDesc(ctypes.Structure):
    _fields_ = [
        ("storage", c_int32),
        ("collision", c_int32),
        ("lid", c_int32),
        ("pad", c_int32),
        ("u_lid", c_double),
        ("s_e", c_double),
        ("s_eps", c_double),
        ("s_q", c_double),
        ("Lambda", c_double),
    ]


CAV_RUNNING, CAV_CONVERGED, CAV_DIVERGED, CAV_MAX_STEPS = 0, 1, 2, 3

# name -> (restype, argtypes); every symbol declared in include/kraken_mi355x.h
SIGNATURES = {
    "kmi_abi_version": (c_int, []),
    "kmi_device_count": (c_int, [POINTER(c_int)]),
    "kmi_last_error": (c_char_p, [c_void_p]),
    "kmi_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "kmi_ctx_destroy": (c_int, [c_void_p]),
    "kmi_ctx_synchronize": (c_int, [c_void_p]),
    "kmi_trt_rates": (None, [c_double, c_double, POINTER(c_double), POINTER(c_double)]),
    "kmi_desc_init": (None, [POINTER(KmiDesc), c_int32, c_int64, c_int64, c_int64, c_double]),
    "kmi_desc_validate": (c_int, [POINTER(KmiDesc)]),
    "kmi_desc_sizeof": (c_size_t, []),
    "kmi_lattice_create": (c_int, [c_void_p, POINTER(KmiDesc), POINTER(c_void_p)]),
    "kmi_lattice_destroy": (c_int, [c_void_p]),
    "kmi_set_mask": (c_int, [c_void_p, POINTER(c_uint8)]),
    "kmi_set_face_profile": (c_int, [c_void_p, c_int32, POINTER(c_double), POINTER(c_double)]),
    "kmi_set_face": (c_int, [c_void_p, c_int32, POINTER(KmiFace)]),
    "kmi_set_cut_links": (c_int, [c_void_p, c_int64, POINTER(c_int32), POINTER(c_int32), POINTER(c_double),
                                  POINTER(c_double)]),
    "kmi_set_force_field": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    "kmi_upload": (c_int, [c_void_p, c_int32, c_void_p, c_int32]),
    "kmi_download": (c_int, [c_void_p, c_int32, c_void_p, c_int32]),
    "kmi_init_equilibrium": (c_int, [c_void_p] + [POINTER(c_double)] * 4),
    "kmi_init_temperature": (c_int, [c_void_p, POINTER(c_double)]),
    "kmi_step": (c_int, [c_void_p, c_int64]),
    "kmi_step_residual": (c_int, [c_void_p, c_int64, POINTER(c_double)]),
    "kmi_step_residual_window": (c_int, [c_void_p, c_int64, c_int64, POINTER(c_double)]),
    "kmi_step_timed": (c_int, [c_void_p, c_int64, POINTER(c_double)]),
    "kmi_moments": (c_int, [c_void_p] + [POINTER(c_double)] * 5),
    "kmi_reduce": (c_int, [c_void_p, c_int32, c_int32, POINTER(c_double)]),
    "kmi_lattice_steps": (c_int, [c_void_p, POINTER(c_int64)]),
    "kmi_lattice_set_steps": (c_int, [c_void_p, c_int64]),
    "kmi_lattice_device_bytes": (c_int, [c_void_p, POINTER(c_int64)]),
    "kmi_lattice_set_launch": (c_int, [c_void_p, c_int32, c_int32, c_int32]),
    "kmi_lattice_get_launch": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    "kmi_lattice_set_temporal": (c_int, [c_void_p, c_int32, c_int32]),
    "kmi_lattice_get_temporal": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    "kmi_lattice_autotune": (c_int, [c_void_p, c_int64, POINTER(c_double)]),
    "kmi_measure_copy_bandwidth": (c_int, [c_void_p, c_int64, c_int32, POINTER(c_double)]),
    "kmi_batch_create": (c_int, [c_void_p, POINTER(KmiBatchDesc), c_int32, POINTER(c_int32),
                                 POINTER(c_double), POINTER(c_void_p)]),
    "kmi_batch_destroy": (c_int, [c_void_p]),
    "kmi_batch_upload": (c_int, [c_void_p, c_int32, c_void_p, c_int32]),
    "kmi_batch_download": (c_int, [c_void_p, c_int32, c_void_p, c_int32]),
    "kmi_batch_init_equilibrium": (c_int, [c_void_p, c_int32] + [POINTER(c_double)] * 3),
    "kmi_batch_moments": (c_int, [c_void_p, c_int32] + [POINTER(c_double)] * 3),
    "kmi_batch_reset_status": (c_int, [c_void_p]),
    "kmi_batch_resume": (c_int, [c_void_p]),
    "kmi_batch_step": (c_int, [c_void_p, c_int64]),
    "kmi_batch_consistent_init": (c_int, [c_void_p, c_void_p, c_int32]),
    "kmi_batch_run": (c_int, [c_void_p, c_double, c_int64, c_int64, c_int64, POINTER(c_double)]),
    "kmi_batch_status": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int64), POINTER(c_double)]),
    "kmi_batch_set_launch": (c_int, [c_void_p, c_int32, c_int32, c_int32, c_int32]),
    "kmi_batch_get_launch": (c_int, [c_void_p] + [POINTER(c_int32)] * 4),
    "kmi_batch_set_temporal": (c_int, [c_void_p, c_int32]),
    "kmi_batch_get_temporal": (c_int, [c_void_p, POINTER(c_int32)]),
    "kmi_batch_autotune": (c_int, [c_void_p, c_int64, POINTER(c_double)]),
    "kmi_batch_active_cells": (c_int, [c_void_p, POINTER(c_int64)]),
    "kmi_nn_desc_init": (None, [c_void_p]),
    "kmi_nn_create": (c_int, [c_void_p, c_void_p, POINTER(c_void_p)]),
    "kmi_nn_destroy": (c_int, [c_void_p]),
    "kmi_nn_param_count": (c_int, [c_void_p, POINTER(c_int64)]),
    "kmi_nn_get_params": (c_int, [c_void_p, POINTER(c_float)]),
    "kmi_nn_set_params": (c_int, [c_void_p, POINTER(c_float)]),
    "kmi_nn_forward": (c_int, [c_void_p, c_int32, POINTER(c_float), POINTER(c_float)]),
    "kmi_nn_loss_grad": (c_int, [c_void_p, c_int32, POINTER(c_float), POINTER(c_float), POINTER(c_double),
                                 POINTER(c_float)]),
    "kmi_nn_train_step": (c_int, [c_void_p, c_int32, POINTER(c_float), POINTER(c_float), c_float,
                                  POINTER(c_double)]),
    "kmi_nn_set_dataset": (c_int, [c_void_p, c_int32, POINTER(c_float), POINTER(c_float)]),
    "kmi_nn_train_epochs": (c_int, [c_void_p, c_int32, c_int32, c_float, c_uint64, POINTER(c_double)]),
    "kmi_gemm_bf16": (c_int, [c_void_p, c_int32, c_int32, c_int32, POINTER(c_float), c_int64, c_int64,
                              POINTER(c_float), c_int64, c_int64, c_int32, POINTER(c_float)]),
    "kmi_gemm_bf16_noalloc": (c_int, [c_void_p, c_int32, c_int32, c_int32, POINTER(c_float), c_int64, c_int64,
                              POINTER(c_float), c_int64, c_int64, c_int32, POINTER(c_float)]),
}

_lib = None


class KrakenError(RuntimeError):
    """A kmi_* call failed; `status` is the kmi_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[kmi status {status}] {message}")
        self.status = status


def load() -> ctypes.CDLL:
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
            "or `make -C krakenn.jl_amd`).  There is no CPU fallback."
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.kmi_desc_sizeof() != ctypes.sizeof(KmiDesc):
        raise ImportError("kmi_desc layout mismatch between library and binding")
    _lib = lib
    return lib


def check(rc: int, ctx=None) -> None:
    if rc != KMI_OK:
        msg = load().kmi_last_error(ctx)
        raise KrakenError(rc, msg.decode() if msg else "unknown error")


def desc_init(stencil: int, nx: int, ny: int, nz: int, nu: float) -> KmiDesc:
    d = KmiDesc()
    load().kmi_desc_init(ctypes.byref(d), stencil, nx, ny, nz, nu)
    return d


def trt_rates(nu: float, Lambda: float = 3.0 / 16.0):
    """(s_plus, s_minus) of src/kernels/fused_trt_2d.jl:129-133, computed by the library."""
    sp, sm = c_double(), c_double()
    load().kmi_trt_rates(nu, Lambda, ctypes.byref(sp), ctypes.byref(sm))
    return sp.value, sm.value
