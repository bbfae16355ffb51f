"""ves3d_amd -- MI355X (gfx950) implementation of the Ves3D per-time-step hot path.

The compute lives in ``lib/libves3d_hip.so`` (hand-written HIP kernels behind
the C ABI declared in ``include/ves3d_hip.h``).  This module only loads that
library and gives Python callers typed access to it; device buffers are
passed as raw pointers (torch tensors on ``cuda`` are accepted and their
``data_ptr()`` is used).  There is no CPU fallback: if the library is missing
or fails to load, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libves3d_hip.so")

c_int = ctypes.c_int
c_size = ctypes.c_size_t
c_ptr = ctypes.c_void_p
c_double = ctypes.c_double

ERRORS = {
    0: "Success", 1: "UnknownError", 2: "ArithmeticError", 5: "MemoryError",
    7: "NotImplementedError", 10: "InvalidParameterError", 20: "InvalidDeviceError",
    30: "AccuracyError", 31: "DivergenceError", 50: "IOError", 51: "IOBadStream",
}


class Ves3dError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__("%s returned %d (%s)" % (fn, code, ERRORS.get(code, "?")))
        self.code = code


# name -> argtypes ; all return int (Error_t)
_SIGS = {}


def _sig(name, *args):
    _SIGS[name] = args


P = c_ptr
S = c_size
_sig("ves3d_hip_init", c_int)
_sig("ves3d_hip_device_info", ctypes.c_char_p, S, ctypes.POINTER(c_int), ctypes.POINTER(c_int))
_sig("ves3d_hip_malloc", ctypes.POINTER(c_ptr), S)
_sig("ves3d_hip_free", P)
_sig("ves3d_hip_memcpy", P, P, S, c_int)
_sig("ves3d_hip_memset", P, c_int, S)
_sig("ves3d_hip_synchronize")
for _t in ("f64", "f32"):
    _sig("ves3d_hip_direct_stokes_" + _t, P, P, P, S, S, S, P, S, S, P, P)
    _sig("ves3d_hip_direct_stokes_dl_" + _t, P, P, P, P, S, S, S, P, S, S, P, P)
for _t, _real in (("f64", c_double), ("f32", ctypes.c_float)):
    _sig("ves3d_hip_dot_product_" + _t, P, P, S, S, P)
    _sig("ves3d_hip_cross_product_" + _t, P, P, S, S, P)
    _sig("ves3d_hip_sqrt_" + _t, P, S, P)
    _sig("ves3d_hip_ax_" + _t, P, P, S, S, P)
    _sig("ves3d_hip_xy_" + _t, P, P, S, P)
    _sig("ves3d_hip_xyinv_" + _t, P, P, S, P)
    _sig("ves3d_hip_uyinv_" + _t, P, P, S, S, P)
    _sig("ves3d_hip_axpy_" + _t, _real, P, P, S, P)
    _sig("ves3d_hip_apx_" + _t, P, P, S, S, P)
    _sig("ves3d_hip_avpw_" + _t, P, P, P, S, S, P)
    _sig("ves3d_hip_xvpw_" + _t, P, P, P, S, S, P)
    _sig("ves3d_hip_reduce_" + _t, P, c_int, P, P, S, S, P)
    _sig("ves3d_hip_maxabs_" + _t, P, S, P)
    _sig("ves3d_hip_algebraic_dot_" + _t, P, P, S, P)
    _sig("ves3d_hip_isnumeric_" + _t, P, S, P)
    _sig("ves3d_hip_transpose_" + _t, P, S, S, P)
    _sig("ves3d_hip_gemm_" + _t, ctypes.c_char, ctypes.c_char, c_int, c_int, c_int, _real, P, c_int,
         P, c_int, _real, P, c_int)
_sig("ves3d_hip_set_gemm_kernel", c_int)
_sig("ves3d_hip_set_gemm_tile", c_int)
_sig("ves3d_hip_set_far_kernel", c_int)
_sig("ves3d_hip_set_pole_kernel", c_int)
_sig("ves3d_hip_pole_kernel_shape", c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
     ctypes.POINTER(c_int))
_sig("ves3d_hip_valu_probe", c_int, P)
_sig("ves3d_hip_stokes_all2all_f64", P, P, S, P, P)
_sig("ves3d_hip_stokes_all2all_f32c", P, P, S, P, P)
_sig("ves3d_hip_direct_eval_f64", P, P, P, S, P, P, S, S, S, P)
_sig("ves3d_hip_direct_eval_f32c", P, P, P, S, P, P, S, S, S, P)
_sig("ves3d_hip_direct_eval_f64_noalloc", P, P, P, S, P, P, S, S, S, P)
for _t in ("f64", "f32"):
    _sig("ves3d_hip_shc2grid_" + _t, P, S, c_int, c_int, P, P, P, P)
    _sig("ves3d_hip_grid2shc_" + _t, P, S, c_int, c_int, P, P)
    _sig("ves3d_hip_shc2grid_transpose_" + _t, P, S, c_int, c_int, P, P)
    _sig("ves3d_hip_shc2pole_" + _t, P, S, c_int, P, P)
    _sig("ves3d_hip_singular_build_" + _t, P, S, c_int, c_int, P, P, P)
    _sig("ves3d_hip_self_apply_" + _t, P, P, P, P, S, c_int, P, P)
_sig("ves3d_hip_shtrans_forward", c_int, S, P, P, P)
_sig("ves3d_hip_shtrans_backward", c_int, S, P, c_int, c_int, P, P)
_sig("ves3d_hip_shtrans_scale_freq", c_int, S, P, P, P, P)
_sig("ves3d_hip_shtrans_low_pass", c_int, S, c_int, P, P, P)
_sig("ves3d_hip_resample", c_int, c_int, S, P, P, P)
_sig("ves3d_hip_rotate_all_f64", P, S, c_int, c_int, P, P)
_sig("ves3d_hip_rotate_transpose_f64", P, S, c_int, c_int, P, P)
_sig("ves3d_hip_bicgstab", S, P, P, P, P, P, P, P, P, P, P)
_sig("ves3d_hip_fgmres", S, P, P, P, P, P, P, P, P, P, c_int, P)
_sig("ves3d_hip_stokes_create", c_int, c_int, S, c_double, P)
_sig("ves3d_hip_stokes_destroy", P)
_sig("ves3d_hip_stokes_set_src", P, P, P)
_sig("ves3d_hip_stokes_eval", P, P, P, P, S, P, P)
_sig("ves3d_hip_stepper_create", P, P)
_sig("ves3d_hip_stepper_destroy", P)
_sig("ves3d_hip_stepper_size", P, P)
_sig("ves3d_hip_stepper_set_state", P, P, P, P)
_sig("ves3d_hip_stepper_get_state", P, P, P, P)
_sig("ves3d_hip_stepper_prepare", P)
_sig("ves3d_hip_stepper_rhs", P, P)
_sig("ves3d_hip_stepper_matvec", P, P, P)
_sig("ves3d_hip_stepper_precond", P, P, P)
_sig("ves3d_hip_stepper_matvec_local", P, P, P)
_sig("ves3d_hip_stepper_far_pairs", P, P)
_sig("ves3d_hip_stepper_residual", P, P)
_sig("ves3d_hip_stepper_rhs_matvec", P, P, P, P)
_sig("ves3d_hip_stepper_precond_learned", P, P, P)
_sig("ves3d_hip_stepper_precond_inner", P, P, P, c_int, P)
_sig("ves3d_hip_stepper_frames", P, P)
_sig("ves3d_hip_stepper_step", P, P)
_sig("ves3d_hip_stepper_attach_pc", P, P, c_int)
_sig("ves3d_hip_stepper_set_vesicle_props", P, P, P, P, P)
_sig("ves3d_hip_stepper_reparam", P, P)
_sig("ves3d_hip_stepper_area_volume", P, P)
_sig("ves3d_hip_stepper_set_area_volume", P, P, P)
_sig("ves3d_read_ascii", ctypes.c_char_p, P, S, P)
_sig("ves3d_initialize_shapes", c_int, S, P, S, P, P)
_sig("ves3d_read_vesicle_props", ctypes.c_char_p, S, P)
_sig("ves3d_checkpoint_defaults", c_int, S, P)
_sig("ves3d_checkpoint_write", ctypes.c_char_p, P, S, P, P, P, P)
_sig("ves3d_checkpoint_read", ctypes.c_char_p, P, P, P, S, P, P, P, S)
_sig("ves3d_hip_movepole_create", c_int, c_int, P, P, P)
_sig("ves3d_hip_movepole_destroy", P)
_sig("ves3d_hip_movepole_set_operands", P, P, P, c_int, P)
_sig("ves3d_hip_movepole_apply", P, c_int, c_int, P)
_sig("ves3d_hip_movepole_matrices", P, P)
_sig("ves3d_hip_aggregate_rotation_f64", c_int, c_int, c_int, c_int, P, P, P, P, P)
_sig("ves3d_hip_near_create", c_int, c_int, S, c_double, c_double, P)
_sig("ves3d_hip_near_destroy", P)
_sig("ves3d_hip_near_set_coord", P, P, P, P)
_sig("ves3d_hip_near_info", P, P, P)
_sig("ves3d_hip_near_force_repul", P, P, P)
_sig("ves3d_hip_near_apply", P, P, P, P, P, P)
_sig("ves3d_hip_lpc_create", c_int, c_int, S, ctypes.c_ulonglong, P)
_sig("ves3d_hip_lpc_destroy", P)
_sig("ves3d_hip_lpc_num_params", P, P)
_sig("ves3d_hip_lpc_samples", P, P)
_sig("ves3d_hip_lpc_set_params", P, P)
_sig("ves3d_hip_lpc_get_params", P, P)
_sig("ves3d_hip_lpc_set_activations", P, c_int)
_sig("ves3d_hip_lpc_forward", P, P, S, P, P)
_sig("ves3d_hip_lpc_loss_grad", P, P, P, S, P, P, P)
_sig("ves3d_hip_lpc_adam", P, ctypes.c_float, P)
_sig("ves3d_hip_lpc_train", P, c_int, c_int, ctypes.c_float, P)
_sig("ves3d_hip_lpc_set_weight_decay", P, ctypes.c_float)
_sig("ves3d_hip_lpc_dataset_loss", P, P, P)

_lib = None


def load():
    """Loads the HIP library (once) and attaches the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("ves3d_amd: %s is not built (run python -m ves3d_amd.build)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = list(args)
        fn.restype = c_int
    _lib = lib
    return lib


def exported_symbols():
    """Names of the C entry points this module binds."""
    return sorted(_SIGS)


def ptr(x):
    """Raw pointer of a torch tensor / int / None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def stream_ptr(stream=None):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def call(name, *args):
    """Calls a C entry point and raises Ves3dError on a non-zero status."""
    fn = getattr(load(), name)
    rc = fn(*args)
    if rc != 0:
        raise Ves3dError(name, rc)
    return rc


def init(device=0):
    call("ves3d_hip_init", device)


def device_info():
    name = ctypes.create_string_buffer(256)
    cu, clk = c_int(0), c_int(0)
    call("ves3d_hip_device_info", name, 256, ctypes.byref(cu), ctypes.byref(clk))
    return name.value.decode(), cu.value, clk.value


from . import implicit, sph, stokes  # noqa: E402,F401
