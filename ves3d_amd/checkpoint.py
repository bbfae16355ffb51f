"""The reference's ASCII checkpoint (ves3d_checkpoint_*): Parameters::pack
followed by EvolveSurface::pack (src/Monitor.cc:58-78), read back as
Simulation does to restart (src/ves3d_simulation.cc:119-123)."""
import ctypes

import numpy as np

from . import call

_C32 = ctypes.c_char * 32


class CheckpointParams(ctypes.Structure):
    """Mirror of ves3d_checkpoint_params (include/ves3d_hip.h)."""
    _fields_ = [
        ("version", ctypes.c_int),
        ("n_surfs", ctypes.c_int), ("sh_order", ctypes.c_int), ("filter_freq", ctypes.c_int),
        ("upsample_freq", ctypes.c_int),
        ("bending_modulus", ctypes.c_double), ("viscosity_contrast", ctypes.c_double),
        ("time_horizon", ctypes.c_double), ("ts", ctypes.c_double), ("time_tol", ctypes.c_double),
        ("time_iter_max", ctypes.c_int), ("time_adaptive", ctypes.c_int),
        ("solve_for_velocity", ctypes.c_int), ("pseudospectral", ctypes.c_int),
        ("scheme", _C32), ("time_precond", _C32), ("bg_flow", _C32), ("singular_stokes", _C32),
        ("rep_type", _C32),
        ("rep_maxit", ctypes.c_int), ("rep_upsample", ctypes.c_int), ("rep_filter_freq", ctypes.c_int),
        ("rep_ts", ctypes.c_double), ("rep_tol", ctypes.c_double), ("rep_exponent", ctypes.c_double),
        ("repul_dist", ctypes.c_double), ("bg_flow_param", ctypes.c_double),
        ("periodic_length", ctypes.c_double),
        ("interaction_upsample", ctypes.c_int), ("checkpoint", ctypes.c_int),
        ("checkpoint_stride", ctypes.c_double), ("error_factor", ctypes.c_double),
        ("num_threads", ctypes.c_int),
        ("excess_density", ctypes.c_double), ("gravity_field", ctypes.c_double * 3),
        ("sht_resample_order", ctypes.c_int), ("diff_filter_freq", ctypes.c_int),
        ("reparam_filter_freq", ctypes.c_int),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, bytes):
                v = v.decode()
            elif name == "gravity_field":
                v = [v[0], v[1], v[2]]
            out[name] = v
        return out


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def defaults(sh_order, nves):
    prm = CheckpointParams()
    call("ves3d_checkpoint_defaults", sh_order, nves, ctypes.byref(prm))
    return prm


def write(path, X, params=None, bending_modulus=None, viscosity_contrast=None,
          excess_density=None):
    """X: positions [nves][3][(p+1)2p]; property arrays default to the
    parameters' values."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    nv = X.shape[0]
    p = int(round((-1 + np.sqrt(1 + 2 * X.shape[2])) / 2))
    prm = params if params is not None else defaults(p, nv)
    fill = lambda a, v: np.ascontiguousarray(np.full(nv, v) if a is None else a, dtype=np.float64)
    bm = fill(bending_modulus, prm.bending_modulus)
    vc = fill(viscosity_contrast, prm.viscosity_contrast)
    ed = fill(excess_density, prm.excess_density)
    call("ves3d_checkpoint_write", str(path).encode(), ctypes.byref(prm), nv, _p(X), _p(bm),
         _p(vc), _p(ed))


def read(path):
    """Returns (params, X [nves][3][(p+1)2p], bending_modulus, viscosity_contrast,
    excess_density)."""
    prm = CheckpointParams()
    n = ctypes.c_size_t()
    call("ves3d_checkpoint_read", str(path).encode(), ctypes.byref(prm), ctypes.byref(n), None, 0,
         None, None, None, 0)
    nv = n.value
    ng = 2 * prm.sh_order * (prm.sh_order + 1)
    X = np.zeros((nv, 3, ng))
    bm, vc, ed = np.zeros(nv), np.zeros(nv), np.zeros(nv)
    call("ves3d_checkpoint_read", str(path).encode(), ctypes.byref(prm), ctypes.byref(n), _p(X),
         X.size, _p(bm), _p(vc), _p(ed), nv)
    return prm, X, bm, vc, ed
