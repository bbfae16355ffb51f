"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference's own CPU kernels (oracle/_ref).  Test-only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_cpukernels.so")

_orc = None
_ref = None


def p(a):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def load():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE):
            raise ImportError("oracle not built: run make -C oracle")
        _orc = ctypes.CDLL(ORACLE)
        for f in dir(_orc):
            pass
        for name in ("orc_direct_stokes", "orc_direct_stokes_dl", "orc_stokes_all2all",
                     "orc_stokes_sl_dl_eval", "orc_stokes_all2all_f32"):
            getattr(_orc, name).restype = None
    return _orc


def load_ref():
    """The reference's src/CPUKernels.cc compiled by oracle/build_ref.sh, or None."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF):
            return None
        os.environ.setdefault("VES3D_DIR", "/tmp")
        _ref = ctypes.CDLL(REF, mode=os.RTLD_LAZY)
    return _ref


def ci(x):
    return ctypes.c_int(int(x))


def cs(x):
    return ctypes.c_size_t(int(x))
