"""Python access to the implicit step (ves3d_hip_stepper_*) and the device
BiCGStab (ves3d_hip_bicgstab)."""
import ctypes

import numpy as np

import torch

from . import Ves3dError, call, load, ptr, stream_ptr

DIVERGENCE_ERROR = 31  # ErrorEvent::DivergenceError (include/Error.h:46-67)


class StepDivergence(Ves3dError):
    """The step's solve failed or its true residual missed the tolerance;
    .stats holds the step statistics."""

    def __init__(self, stats):
        super().__init__("ves3d_hip_stepper_step", DIVERGENCE_ERROR)
        self.stats = stats


class StepParams(ctypes.Structure):
    _fields_ = [("nves", ctypes.c_size_t), ("p", ctypes.c_int), ("p_up", ctypes.c_int),
                ("filter_freq", ctypes.c_int), ("dt", ctypes.c_double),
                ("shear_rate", ctypes.c_double), ("bending_modulus", ctypes.c_double),
                ("tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("precision", ctypes.c_int), ("precond", ctypes.c_int),
                ("near_off", ctypes.c_int), ("repul_dist", ctypes.c_double),
                ("r_near", ctypes.c_double), ("post_off", ctypes.c_int),
                ("inner_max_iter", ctypes.c_int), ("inner_tol", ctypes.c_double),
                ("extrapolate", ctypes.c_int), ("far_coarse", ctypes.c_double),
                ("far_coarse2", ctypes.c_double), ("fallback", ctypes.c_int),
                ("solver", ctypes.c_int), ("restart", ctypes.c_int)]


class StepStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("matvecs", ctypes.c_int), ("status", ctypes.c_int),
                ("fell_back", ctypes.c_int), ("relres", ctypes.c_double),
                ("true_relres", ctypes.c_double), ("ms_setup", ctypes.c_double),
                ("ms_rhs", ctypes.c_double), ("ms_solve", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("ms_far", ctypes.c_double),
                ("far_calls", ctypes.c_int), ("near_pairs", ctypes.c_int),
                ("r_near", ctypes.c_double), ("reparam_iters", ctypes.c_int),
                ("av_iters", ctypes.c_int), ("ms_post", ctypes.c_double),
                ("ms_far2", ctypes.c_double), ("far2_calls", ctypes.c_int),
                ("inner_iters", ctypes.c_int), ("local_matvecs", ctypes.c_int),
                ("far_pairs", ctypes.c_double), ("ms_near", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)


class Stepper:
    """One globally implicit interfacial-velocity step per call (velocity
    form, Galerkin packing, shear flow)."""

    def __init__(self, nves, p, dt=1e-3, shear_rate=0.1, bending_modulus=1e-2, tol=1e-6,
                 max_iter=200, precision=0, precond=1, p_up=0, filter_freq=0, near=True,
                 repul_dist=0.0, r_near=0.0, post=True, inner_max_iter=0, inner_tol=0.0,
                 extrapolate=False, far_coarse=0.0, far_coarse2=0.0, fallback="diagonal",
                 solver="bicgstab", restart=0):
        self.params = StepParams(nves, p, p_up, filter_freq, dt, shear_rate, bending_modulus, tol,
                                 max_iter, precision, precond, 0 if near else 1, repul_dist, r_near,
                                 0 if post else 1, inner_max_iter, inner_tol,
                                 1 if extrapolate else 0, far_coarse, far_coarse2,
                                 {"diagonal": 0, "none": 1}[fallback],
                                 {"bicgstab": 0, "fgmres": 1}[solver], restart)
        self.nves, self.p = nves, p
        self.ng = 2 * p * (p + 1)
        h = ctypes.c_void_p()
        call("ves3d_hip_stepper_create", ctypes.byref(self.params), ctypes.byref(h))
        self.h = h
        n = ctypes.c_size_t()
        call("ves3d_hip_stepper_size", self.h, ctypes.byref(n))
        self.n = n.value

    def close(self):
        if self.h:
            call("ves3d_hip_stepper_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, t, numel):
        if t is None:
            return
        if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous() or t.numel() != numel:
            raise ValueError("expected a contiguous float64 CUDA tensor of %d elements" % numel)

    def set_state(self, X, U=None, sigma=None):
        self._chk(X, 3 * self.nves * self.ng)
        self._chk(U, 3 * self.nves * self.ng)
        self._chk(sigma, self.nves * self.ng)
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_set_state", self.h, ptr(X), ptr(U), ptr(sigma))

    def get_state(self):
        X = torch.empty((self.nves, 3, self.ng), dtype=torch.float64, device="cuda")
        U = torch.empty_like(X)
        s = torch.empty((self.nves, self.ng), dtype=torch.float64, device="cuda")
        call("ves3d_hip_stepper_get_state", self.h, ptr(X), ptr(U), ptr(s))
        return X, U, s

    def prepare(self):
        call("ves3d_hip_stepper_prepare", self.h)

    def far_pairs(self):
        """Source-target pairs of one far-field evaluation (after prepare)."""
        v = ctypes.c_double()
        call("ves3d_hip_stepper_far_pairs", self.h, ctypes.byref(v))
        return v.value

    def rhs(self):
        b = torch.empty(self.n, dtype=torch.float64, device="cuda")
        call("ves3d_hip_stepper_rhs", self.h, ptr(b))
        return b

    def matvec(self, x):
        self._chk(x, self.n)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_matvec", self.h, ptr(x), ptr(y))
        return y

    def precond(self, x):
        self._chk(x, self.n)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_precond", self.h, ptr(x), ptr(y))
        return y

    def rhs_matvec(self, x):
        """(b, A x) with one far-field pass for both."""
        self._chk(x, self.n)
        b = torch.empty_like(x)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_rhs_matvec", self.h, ptr(x), ptr(b), ptr(y))
        return b, y

    def residual(self):
        """|b - A x| / |b| of the current (positions, velocity, tension)."""
        r = ctypes.c_double()
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_residual", self.h, ctypes.byref(r))
        return r.value

    def matvec_local(self, x):
        """Every vesicle's own operator (no far field, no near correction)."""
        self._chk(x, self.n)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_matvec_local", self.h, ptr(x), ptr(y))
        return y

    def precond_inner(self, x, learned=False):
        """(y, iterations): y ~ A_loc^-1 x by the inner BiCGStab of the
        inner-solve preconditioner (diagonal or learned inner preconditioner)."""
        self._chk(x, self.n)
        y = torch.empty_like(x)
        it = ctypes.c_int()
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_precond_inner", self.h, ptr(x), ptr(y), 1 if learned else 0,
             ctypes.byref(it))
        return y, it.value

    def precond_learned(self, x):
        self._chk(x, self.n)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_precond_learned", self.h, ptr(x), ptr(y))
        return y

    def frames(self):
        """Canonical frames [nv][3][3] of the last prepare."""
        F = torch.empty((self.nves, 3, 3), dtype=torch.float64, device="cuda")
        call("ves3d_hip_stepper_frames", self.h, ptr(F))
        return F

    def set_vesicle_props(self, bending_modulus=None, viscosity_contrast=None,
                          excess_density=None, gravity=None):
        """Per-vesicle properties ([nves] each) and the gravity field (3);
        None leaves an entry unchanged."""
        def arr(a, n):
            if a is None:
                return None, None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
            return a, a.ctypes.data_as(ctypes.c_void_p)
        kb, pkb = arr(bending_modulus, self.nves)
        vc, pvc = arr(viscosity_contrast, self.nves)
        rho, prho = arr(excess_density, self.nves)
        g, pg = arr(gravity, 3)
        call("ves3d_hip_stepper_set_vesicle_props", self.h, pkb, pvc, prho, pg)

    def step(self, check=True):
        """One implicit step; returns its statistics.  A solve that fails or
        whose true residual misses the tolerance is a DivergenceError
        (src/InterfacialVelocity.cc:176-188): the state is left as it was,
        and the call raises StepDivergence (carrying the statistics) or, with
        check=False, returns them with "diverged" set."""
        st = StepStats()
        torch.cuda.synchronize()
        rc = load().ves3d_hip_stepper_step(self.h, ctypes.byref(st))
        d = st.as_dict()
        d["diverged"] = rc == DIVERGENCE_ERROR
        if rc == DIVERGENCE_ERROR and check:
            raise StepDivergence(d)
        if rc not in (0, DIVERGENCE_ERROR):
            raise Ves3dError("ves3d_hip_stepper_step", rc)
        return d

    def reparam(self):
        """Reparametrization of the current positions; returns the iterations."""
        it = ctypes.c_int()
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_reparam", self.h, ctypes.byref(it))
        return it.value

    def area_volume(self):
        """Area / volume correction of the current positions; returns the
        iterations."""
        it = ctypes.c_int()
        torch.cuda.synchronize()
        call("ves3d_hip_stepper_area_volume", self.h, ctypes.byref(it))
        return it.value

    def set_area_volume(self, area, vol):
        a = np.ascontiguousarray(area, dtype=np.float64)
        v = np.ascontiguousarray(vol, dtype=np.float64)
        call("ves3d_hip_stepper_set_area_volume", self.h, a.ctypes.data_as(ctypes.c_void_p),
             v.ctypes.data_as(ctypes.c_void_p))


class LearnedPC:
    """The learned SH-space preconditioner (ves3d_hip_lpc_*)."""

    def __init__(self, nin, width=256, capacity=0, seed=1):
        h = ctypes.c_void_p()
        call("ves3d_hip_lpc_create", nin, width, capacity, seed, ctypes.byref(h))
        self.h = h
        self.nin, self.width = nin, width
        self.activations = "fp32"
        n = ctypes.c_size_t()
        call("ves3d_hip_lpc_num_params", self.h, ctypes.byref(n))
        self.num_params = n.value

    def close(self):
        if self.h:
            call("ves3d_hip_lpc_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layout(self):
        """Offsets of W1, b1, W2, b2, W3, b3 in the parameter vector."""
        w, n = self.width, self.nin
        sizes = [w * n, w, w * w, w, n * w, n]
        shapes = [(w, n), (w,), (w, w), (w,), (n, w), (n,)]
        offs = [0]
        for s in sizes[:-1]:
            offs.append(offs[-1] + s)
        return list(zip(offs, sizes, shapes))

    def get_params(self):
        out = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        call("ves3d_hip_lpc_get_params", self.h, ptr(out))
        return out

    def set_params(self, p):
        torch.cuda.synchronize()
        call("ves3d_hip_lpc_set_params", self.h, ptr(p))

    def set_activations(self, kind):
        """"fp32" (FP32 MFMA tiles) or "bf16" (BF16 operands, FP32
        accumulation, FP32 master weights)."""
        if kind not in ("fp32", "bf16"):
            raise ValueError(kind)
        call("ves3d_hip_lpc_set_activations", self.h, 1 if kind == "bf16" else 0)
        self.activations = kind

    def samples(self):
        n = ctypes.c_size_t()
        call("ves3d_hip_lpc_samples", self.h, ctypes.byref(n))
        return n.value

    def forward(self, Z):
        Y = torch.empty_like(Z)
        torch.cuda.synchronize()
        call("ves3d_hip_lpc_forward", self.h, ptr(Z), Z.shape[0], ptr(Y), stream_ptr())
        torch.cuda.synchronize()
        return Y

    def loss_grad(self, Z, T):
        g = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        loss = ctypes.c_double()
        torch.cuda.synchronize()
        call("ves3d_hip_lpc_loss_grad", self.h, ptr(Z), ptr(T), Z.shape[0], ctypes.byref(loss),
             ptr(g), stream_ptr())
        return loss.value, g

    def adam(self, lr):
        call("ves3d_hip_lpc_adam", self.h, ctypes.c_float(lr), stream_ptr())
        torch.cuda.synchronize()

    def train(self, epochs, batch, lr):
        loss = ctypes.c_double()
        torch.cuda.synchronize()
        call("ves3d_hip_lpc_train", self.h, epochs, batch, ctypes.c_float(lr), ctypes.byref(loss))
        return loss.value

    def set_weight_decay(self, wd):
        call("ves3d_hip_lpc_set_weight_decay", self.h, ctypes.c_float(wd))

    def dataset_loss(self):
        """(loss over all recorded samples, mean square of their targets)."""
        loss, tms = ctypes.c_double(), ctypes.c_double()
        torch.cuda.synchronize()
        call("ves3d_hip_lpc_dataset_loss", self.h, ctypes.byref(loss), ctypes.byref(tms))
        return loss.value, tms.value


def attach(stepper, pc, collect):
    call("ves3d_hip_stepper_attach_pc", stepper.h, pc.h if pc is not None else None,
         1 if collect else 0)


def fgmres(apply_a, x, b, tol, max_iter, restart=30, apply_p=None, stream=None):
    """Device flexible GMRES with Python-side operator callbacks receiving raw
    device pointers (ctx, x, y).  Returns (iterations, relres, status)."""
    a_cb = APPLY_FN(apply_a)
    p_cb = APPLY_FN(apply_p) if apply_p is not None else None
    it = ctypes.c_int(max_iter)
    tl = ctypes.c_double(tol)
    status = ctypes.c_int(0)
    torch.cuda.synchronize()
    fa = ctypes.cast(a_cb, ctypes.c_void_p)
    fp = ctypes.cast(p_cb, ctypes.c_void_p) if p_cb is not None else None
    call("ves3d_hip_fgmres", x.numel(), fa, None, fp, None, ptr(x), ptr(b),
         ctypes.byref(it), ctypes.byref(tl), ctypes.byref(status), restart, stream_ptr(stream))
    return it.value, tl.value, status.value


def bicgstab(apply_a, x, b, tol, max_iter, apply_p=None, stream=None):
    """Device BiCGStab with Python-side operator callbacks receiving raw
    device pointers (ctx, x, y).  Returns (iterations, relres, status)."""
    a_cb = APPLY_FN(apply_a)
    p_cb = APPLY_FN(apply_p) if apply_p is not None else None
    it = ctypes.c_int(max_iter)
    tl = ctypes.c_double(tol)
    status = ctypes.c_int(0)
    torch.cuda.synchronize()
    fa = ctypes.cast(a_cb, ctypes.c_void_p)
    fp = ctypes.cast(p_cb, ctypes.c_void_p) if p_cb is not None else None
    call("ves3d_hip_bicgstab", x.numel(), fa, None, fp, None, ptr(x), ptr(b),
         ctypes.byref(it), ctypes.byref(tl), ctypes.byref(status), stream_ptr(stream))
    return it.value, tl.value, status.value
