"""Host-loop restatement of the Device<CPU> elementwise / reduction /
transpose / gemm members (src/DeviceCPU.cc:64-640), vectorised with numpy.

TEST INFRASTRUCTURE ONLY (the checker of ves3d_amd/csrc/device_ops.hip).
Vector fields are AxisMajor [n][3][stride]; scalar fields [n][stride].
"""
import numpy as np


def dot_product(u, v, stride, n):                    # :64-83
    return np.sum(u.reshape(n, 3, stride) * v.reshape(n, 3, stride), axis=1)


def cross_product(u, v, stride, n):                  # :85-113
    return np.cross(u.reshape(n, 3, stride), v.reshape(n, 3, stride), axis=1)


def ax(a, x, stride, n):                             # :128-136
    return a[None, :] * x.reshape(n, stride)


def xyinv(x, y):                                     # :150-176
    return (1.0 if x is None else x) / y


def uyinv(u, y, stride, n):                          # :178-198
    return u.reshape(n, 3, stride) / y.reshape(n, 1, stride)


def axpy(a, x, y):                                   # :200-223
    return a * x + (0.0 if y is None else y)


def apx(a, x, stride, ns):                           # :225-236
    return a[:, None] + x.reshape(ns, stride)


def avpw(a, v, w, stride, n):                        # :238-283
    r = a[:, None, None] * v.reshape(n, 3, stride)
    return r + (0.0 if w is None else w.reshape(n, 3, stride))


def xvpw(x, v, w, stride, n):                        # :285-336
    r = x.reshape(n, 1, stride) * v.reshape(n, 3, stride)
    return r + (0.0 if w is None else w.reshape(n, 3, stride))


def reduce(x, x_dim, w, qw, stride, ns):             # :338-380
    wq = w.reshape(ns, stride) * qw[None, :]
    if x is None:
        return wq.sum(axis=1)
    return np.einsum("ikj,ij->ik", x.reshape(ns, x_dim, stride), wq)


def maxabs(x):                                       # :557-580
    return 0.0 if x.size == 0 else float(np.max(np.abs(x)))


def transpose(a, h, w):                              # :582-592
    return a.reshape(h, w).T.copy()


def gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """BLAS column-major dgemm semantics (include/VesBlas.h:15-40)."""
    Am = A.reshape(-1, lda).T          # column-major view: Am[r][c] = A[r + c lda]
    Bm = B.reshape(-1, ldb).T
    Cm = C.reshape(-1, ldc).T.copy()
    opA = Am[:k, :m].T if ta in "Tt" else Am[:m, :k]
    opB = Bm[:n, :k].T if tb in "Tt" else Bm[:k, :n]
    Cm[:m, :n] = alpha * (opA @ opB) + beta * Cm[:m, :n]
    return Cm.T.reshape(-1)
