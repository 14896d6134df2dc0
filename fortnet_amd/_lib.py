"""ctypes binding of libfnet_hip.so (the C-ABI declared in include/fnet_hip.h).

The library is the only compute path: if it is missing, or no GPU is visible, every entry point
raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfnet_hip.so")

FNET_FP64 = 0
FNET_FP32 = 1
FNET_FP32_NET64 = 2
FNET_BF16 = 3        # hidden-to-hidden layers on BF16 MFMA, split hi + lo operands
FNET_BF16_FAST = 4   # hidden-to-hidden layers on BF16 MFMA, one pass
FNET_BATCHED = 0
FNET_PER_STRUCT = 1
LOSSES = {"mse": 0, "mae": 1, "mape": 2, "rms": 3}  # fnet_loss
OPTIMIZERS = {"sd": 0, "cg": 1, "lbfgs": 2, "fire": 3}  # fnet_optimizer

EXPORTS = [
    "fnet_ctx_create", "fnet_ctx_destroy", "fnet_last_error", "fnet_params_per_species",
    "fnet_model_upload", "fnet_params_upload", "fnet_batch_upload", "fnet_batch_destroy",
    "fnet_batch_update_coords", "fnet_batch_set_mode", "fnet_batch_n_atoms",
    "fnet_neighbour_list", "fnet_get_neighbours", "fnet_acsf_forward", "fnet_energy",
    "fnet_forces", "fnet_run", "fnet_train_gradients", "fnet_optimizer_step_sd", "fnet_sync",
    "fnet_stage_times", "fnet_batch_reload", "fnet_query", "fnet_set_loss",
    "fnet_set_regularization", "fnet_set_graph", "fnet_batch_sync", "fnet_step_time",
    "fnet_batch_step_info", "fnet_optimizer_init", "fnet_optimizer_step", "fnet_device_memory",
    "fnet_forces_noalloc", "fnet_set_tune", "fnet_get_tune", "fnet_autotune",
]


class FnetError(RuntimeError):
    pass


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("n_g", ctypes.c_int), ("g_type", ctypes.c_void_p), ("g_rcut", ctypes.c_void_p),
        ("g_eta", ctypes.c_void_p), ("g_rs", ctypes.c_void_p), ("g_kappa", ctypes.c_void_p),
        ("g_lambda", ctypes.c_void_p), ("g_xi", ctypes.c_void_p), ("g_atomid", ctypes.c_void_p),
        ("g_z1", ctypes.c_void_p), ("g_z2", ctypes.c_void_p), ("z_mean", ctypes.c_void_p),
        ("z_std", ctypes.c_void_p), ("n_ext_in", ctypes.c_int), ("ext_in_idx", ctypes.c_void_p),
        ("n_species", ctypes.c_int), ("species_z", ctypes.c_void_p), ("n_layers", ctypes.c_int),
        ("dims", ctypes.c_void_p), ("activation", ctypes.c_void_p), ("params", ctypes.c_void_p),
    ]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("n_struct", ctypes.c_int), ("offsets", ctypes.c_void_p), ("coords", ctypes.c_void_p),
        ("latvecs", ctypes.c_void_p), ("periodic", ctypes.c_void_p), ("species", ctypes.c_void_p),
        ("atnum", ctypes.c_void_p), ("n_ext", ctypes.c_int), ("ext", ctypes.c_void_p),
    ]


_lib = None


def load():
    """Loads libfnet_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FnetError("libfnet_hip.so is not built (run `python -m fortnet_amd.build`)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    sig = {
        "fnet_ctx_create": (i32, [i32, ctypes.POINTER(vp)]),
        "fnet_ctx_destroy": (None, [vp]),
        "fnet_last_error": (i32, [vp, ctypes.c_char_p, i32]),
        "fnet_params_per_species": (i64, [i32, vp]),
        "fnet_model_upload": (i32, [vp, ctypes.POINTER(ModelDesc)]),
        "fnet_params_upload": (i32, [vp, vp]),
        "fnet_batch_upload": (i32, [vp, ctypes.POINTER(BatchDesc), ctypes.POINTER(vp)]),
        "fnet_batch_destroy": (None, [vp]),
        "fnet_batch_update_coords": (i32, [vp, vp, vp]),
        "fnet_batch_set_mode": (i32, [vp, i32, i32]),
        "fnet_batch_n_atoms": (i64, [vp]),
        "fnet_neighbour_list": (i32, [vp, vp, ctypes.POINTER(i64)]),
        "fnet_get_neighbours": (i32, [vp, vp, vp, vp, vp]),
        "fnet_acsf_forward": (i32, [vp, vp, vp]),
        "fnet_energy": (i32, [vp, vp, vp, vp]),
        "fnet_forces": (i32, [vp, vp, vp]),
        "fnet_run": (i32, [vp, vp]),
        "fnet_train_gradients": (i32, [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp]),
        "fnet_optimizer_step_sd": (i32, [vp, dbl, dbl, dbl, vp, ctypes.POINTER(i32)]),
        "fnet_sync": (i32, [vp]),
        "fnet_stage_times": (i32, [vp, vp, i32]),
        "fnet_set_loss": (i32, [vp, i32]),
        "fnet_set_regularization": (i32, [vp, dbl, dbl]),
        "fnet_batch_reload": (i32, [vp, ctypes.POINTER(BatchDesc)]),
        "fnet_query": (i32, [vp]),
        "fnet_set_graph": (i32, [vp, i32]),
        "fnet_batch_sync": (i32, [vp]),
        "fnet_step_time": (i32, [vp, ctypes.POINTER(dbl)]),
        "fnet_batch_step_info": (i32, [vp, ctypes.POINTER(i32)]),
        "fnet_optimizer_init": (i32, [vp, i32, dbl, dbl, dbl, dbl, i32, i32]),
        "fnet_optimizer_step": (i32, [vp, i32, vp, ctypes.POINTER(i32)]),
        "fnet_device_memory": (i32, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "fnet_forces_noalloc": (i32, [vp, vp]),
        "fnet_set_tune": (i32, [vp, vp, vp]),
        "fnet_get_tune": (i32, [vp, vp, vp]),
        "fnet_autotune": (i32, [vp, vp, i32, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def check(ctx, rc, what):
    if rc != 0:
        msg = ""
        if ctx:
            buf = ctypes.create_string_buffer(1024)
            load().fnet_last_error(ctx, buf, 1024)
            msg = buf.value.decode(errors="replace")
        raise FnetError("%s failed (status %d): %s" % (what, rc, msg))
