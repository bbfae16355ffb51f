"""GPU parity of the Device<> elementwise / reduction / transpose / gemm
members against the host-loop oracle (oracle/device.py), in the manner of the
reference's test/DeviceTest.h (eps = 100 machine epsilon, :45-53)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import device as D
from ves3d_amd import call, ptr

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(41)


@pytest.fixture(params=[("f64", torch.float64, np.float64), ("f32", torch.float32, np.float32)])
def prec(request):
    return request.param


def cu(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)


def close(g, ref, npdt):
    eps = 100 * np.finfo(npdt).eps
    g = g.detach().cpu().numpy().astype(np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert np.max(np.abs(g.reshape(-1) - ref.reshape(-1))) <= eps * max(1.0, np.max(np.abs(ref))) * 10


def test_elementwise(hip, prec):
    s, tdt, ndt = prec
    n, stride = 5, 61
    u = RNG.standard_normal((n, 3, stride)).astype(ndt)
    v = RNG.standard_normal((n, 3, stride)).astype(ndt)
    w = RNG.standard_normal((n, 3, stride)).astype(ndt)
    y = RNG.uniform(0.5, 2, (n, stride)).astype(ndt)
    a = RNG.standard_normal(stride).astype(ndt)
    an = RNG.standard_normal(n).astype(ndt)
    gu, gv, gw, gy = cu(u, tdt), cu(v, tdt), cu(w, tdt), cu(y, tdt)
    out_s = torch.empty((n, stride), dtype=tdt, device="cuda")
    out_v = torch.empty((n, 3, stride), dtype=tdt, device="cuda")
    call("ves3d_hip_dot_product_" + s, ptr(gu), ptr(gv), stride, n, ptr(out_s)); torch.cuda.synchronize()
    close(out_s, D.dot_product(u, v, stride, n), ndt)
    call("ves3d_hip_cross_product_" + s, ptr(gu), ptr(gv), stride, n, ptr(out_v)); torch.cuda.synchronize()
    close(out_v, D.cross_product(u, v, stride, n), ndt)
    call("ves3d_hip_sqrt_" + s, ptr(gy), n * stride, ptr(out_s)); torch.cuda.synchronize()
    close(out_s, np.sqrt(y), ndt)
    call("ves3d_hip_ax_" + s, ptr(cu(a, tdt)), ptr(gy), stride, n, ptr(out_s)); torch.cuda.synchronize()
    close(out_s, D.ax(a, y, stride, n), ndt)
    call("ves3d_hip_xy_" + s, ptr(gy), ptr(gy), n * stride, ptr(out_s)); torch.cuda.synchronize()
    close(out_s, y * y, ndt)
    for x in (None, y * 0.5):
        call("ves3d_hip_xyinv_" + s, ptr(None if x is None else cu(x, tdt)), ptr(gy), n * stride, ptr(out_s))
        torch.cuda.synchronize()
        close(out_s, D.xyinv(x, y), ndt)
    call("ves3d_hip_uyinv_" + s, ptr(gu), ptr(gy), stride, n, ptr(out_v)); torch.cuda.synchronize()
    close(out_v, D.uyinv(u, y, stride, n), ndt)
    for yy in (None, w):
        call("ves3d_hip_axpy_" + s, 0.7, ptr(gu), ptr(None if yy is None else gw), 3 * n * stride, ptr(out_v))
        torch.cuda.synchronize()
        close(out_v, D.axpy(ndt(0.7), u, yy), ndt)
    call("ves3d_hip_apx_" + s, ptr(cu(an, tdt)), ptr(gy), stride, n, ptr(out_s)); torch.cuda.synchronize()
    close(out_s, D.apx(an, y, stride, n), ndt)
    for ww in (None, w):
        call("ves3d_hip_avpw_" + s, ptr(cu(an, tdt)), ptr(gv), ptr(None if ww is None else gw), stride, n,
             ptr(out_v)); torch.cuda.synchronize()
        close(out_v, D.avpw(an, v, ww, stride, n), ndt)
        call("ves3d_hip_xvpw_" + s, ptr(gy), ptr(gv), ptr(None if ww is None else gw), stride, n, ptr(out_v))
        torch.cuda.synchronize()
        close(out_v, D.xvpw(y, v, ww, stride, n), ndt)


def test_reductions_transpose_gemm(hip, prec):
    s, tdt, ndt = prec
    ns, xd, stride = 4, 3, 131
    x = RNG.standard_normal((ns, xd, stride)).astype(ndt)
    w = RNG.uniform(0.1, 1, (ns, stride)).astype(ndt)
    qw = RNG.uniform(0.1, 1, stride).astype(ndt)
    out = torch.empty(ns * xd, dtype=tdt, device="cuda")
    # keep every device input referenced until the kernels have run: a freed
    # temporary's block would be reused by the next upload
    gx, gw, gq = cu(x, tdt), cu(w, tdt), cu(qw, tdt)
    call("ves3d_hip_reduce_" + s, ptr(gx), xd, ptr(gw), ptr(gq), stride, ns, ptr(out))
    torch.cuda.synchronize()
    close(out, D.reduce(x, xd, w, qw, stride, ns), ndt)
    call("ves3d_hip_reduce_" + s, None, xd, ptr(gw), ptr(gq), stride, ns, ptr(out))
    torch.cuda.synchronize()
    close(out[:ns], D.reduce(None, 1, w, qw, stride, ns), ndt)
    ctype = ctypes.c_double if s == "f64" else ctypes.c_float
    big = RNG.standard_normal(100003).astype(ndt)
    r = ctype()
    call("ves3d_hip_maxabs_" + s, ptr(cu(big, tdt)), big.size, ctypes.byref(r))
    assert r.value == pytest.approx(D.maxabs(big))
    gb = cu(big, tdt)
    call("ves3d_hip_algebraic_dot_" + s, ptr(gb), ptr(gb), big.size, ctypes.byref(r))
    assert r.value == pytest.approx(float(np.dot(big.astype(np.float64), big)), rel=1e-6)
    ok = ctypes.c_int()
    call("ves3d_hip_isnumeric_" + s, ptr(cu(big, tdt)), big.size, ctypes.byref(ok))
    assert ok.value == 1
    big[777] = np.nan
    call("ves3d_hip_isnumeric_" + s, ptr(cu(big, tdt)), big.size, ctypes.byref(ok))
    assert ok.value == 0
    h, wd = 37, 70
    a = RNG.standard_normal((h, wd)).astype(ndt)
    t = torch.empty(h * wd, dtype=tdt, device="cuda")
    call("ves3d_hip_transpose_" + s, ptr(cu(a, tdt)), h, wd, ptr(t)); torch.cuda.synchronize()
    close(t, D.transpose(a, h, wd), ndt)
    m, n, k = 13, 9, 17
    for ta, tb in (("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")):
        lda = (m if ta == "N" else k) + 2
        ldb = (k if tb == "N" else n) + 1
        ldc = m + 3
        A = RNG.standard_normal(lda * (k if ta == "N" else m)).astype(ndt)
        B = RNG.standard_normal(ldb * (n if tb == "N" else k)).astype(ndt)
        C = RNG.standard_normal(ldc * n).astype(ndt)
        gA, gB, gC = cu(A, tdt), cu(B, tdt), cu(C, tdt)
        call("ves3d_hip_gemm_" + s, ta.encode(), tb.encode(), m, n, k, 1.5, ptr(gA), lda,
             ptr(gB), ldb, 0.5, ptr(gC), ldc)
        torch.cuda.synchronize()
        close(gC, D.gemm(ta, tb, m, n, k, 1.5, A.astype(np.float64), lda, B.astype(np.float64), ldb, 0.5,
                         C.astype(np.float64), ldc), ndt)
