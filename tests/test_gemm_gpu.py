"""GPU parity of the MFMA grouped-GEMM kernel against the VALU kernel and
numpy, through the BLAS-style gemm entry and through the SH transforms and
the singular-operator build that run on it.

Tolerances: FP64 1e-13 relative (both kernels accumulate K in FP64, in a
different order); FP32 against an FP64 numpy product: 3e-6 relative to the
largest |A||B| row-column sum.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import sph
from ves3d_amd import call, ptr
from ves3d_amd import sph as S

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(123)


def cu(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)


GEMM_KINDS = [(1, -1), (0, -1)] + [(1, t) for t in range(9)]


@pytest.fixture(params=GEMM_KINDS,
                ids=["mfma", "valu"] + ["mfma-tile%d" % t for t in range(9)])
def mfma(request, hip):
    """Both kernels behind the grouped GEMM and every MFMA tile shape; the
    defaults (MFMA, automatic tile) are restored."""
    kind, tile = request.param
    call("ves3d_hip_set_gemm_kernel", kind)
    call("ves3d_hip_set_gemm_tile", tile)
    yield kind
    call("ves3d_hip_set_gemm_kernel", 1)
    call("ves3d_hip_set_gemm_tile", -1)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (13, 9, 17), (64, 64, 16), (70, 130, 33),
                                   (257, 3, 129)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_blas_gemm_mfma(mfma, m, n, k, ta, tb):
    for s, tdt, ndt, tol in (("f64", torch.float64, np.float64, 1e-13),
                             ("f32", torch.float32, np.float32, 3e-6)):
        lda = (m if ta == "N" else k) + 1
        ldb = (k if tb == "N" else n) + 2
        ldc = m + 1
        A = RNG.standard_normal(lda * (k if ta == "N" else m)).astype(ndt)
        B = RNG.standard_normal(ldb * (n if tb == "N" else k)).astype(ndt)
        C = RNG.standard_normal(ldc * n).astype(ndt)
        gA, gB, gC = cu(A, tdt), cu(B, tdt), cu(C, tdt)
        print("gemm %s %s%s m=%d n=%d k=%d" % (s, ta, tb, m, n, k), flush=True)
        call("ves3d_hip_gemm_" + s, ta.encode(), tb.encode(), m, n, k, 1.25, ptr(gA), lda, ptr(gB),
             ldb, -0.5, ptr(gC), ldc)
        torch.cuda.synchronize()
        Am = A.astype(np.float64).reshape(-1, lda).T
        Bm = B.astype(np.float64).reshape(-1, ldb).T
        opA = Am[:k, :m].T if ta == "T" else Am[:m, :k]
        opB = Bm[:n, :k].T if tb == "T" else Bm[:k, :n]
        want = C.astype(np.float64).reshape(-1, ldc).T.copy()
        want[:m, :n] = 1.25 * (opA @ opB) - 0.5 * want[:m, :n]
        got = gC.cpu().numpy().astype(np.float64).reshape(-1, ldc).T
        scale = 1.25 * (np.abs(opA) @ np.abs(opB)) + 0.5 * np.abs(want[:m, :n]) + 1.0
        assert np.max(np.abs(got - want) / np.pad(scale, ((0, ldc - m), (0, 0)),
                                                    constant_values=1.0)) < tol


@pytest.mark.parametrize("m,n,k", [(64, 32, 64), (128, 48, 32), (256, 16, 16), (96, 20, 36)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_blas_gemm_vector_loads(mfma, m, n, k, ta, tb):
    """Leading dimensions equal to the contiguous extents (multiples of 4), so
    the A operand qualifies for 16-byte vector loads in both precisions."""
    for s, tdt, ndt, tol in (("f64", torch.float64, np.float64, 1e-13),
                             ("f32", torch.float32, np.float32, 3e-6)):
        lda = m if ta == "N" else k
        ldb = k if tb == "N" else n
        ldc = m
        A = RNG.standard_normal(lda * (k if ta == "N" else m)).astype(ndt)
        B = RNG.standard_normal(ldb * (n if tb == "N" else k)).astype(ndt)
        C = RNG.standard_normal(ldc * n).astype(ndt)
        gA, gB, gC = cu(A, tdt), cu(B, tdt), cu(C, tdt)
        print("gemm vec %s %s%s m=%d n=%d k=%d" % (s, ta, tb, m, n, k), flush=True)
        call("ves3d_hip_gemm_" + s, ta.encode(), tb.encode(), m, n, k, 0.75, ptr(gA), lda, ptr(gB),
             ldb, 0.25, ptr(gC), ldc)
        torch.cuda.synchronize()
        Am = A.astype(np.float64).reshape(-1, lda).T
        Bm = B.astype(np.float64).reshape(-1, ldb).T
        opA = Am[:k, :m].T if ta == "T" else Am[:m, :k]
        opB = Bm[:n, :k].T if tb == "T" else Bm[:k, :n]
        Cm = C.astype(np.float64).reshape(-1, ldc).T
        want = 0.75 * (opA @ opB) + 0.25 * Cm
        got = gC.cpu().numpy().astype(np.float64).reshape(-1, ldc).T
        scale = 0.75 * (np.abs(opA) @ np.abs(opB)) + 0.25 * np.abs(Cm) + 1.0
        assert np.max(np.abs(got - want) / scale) < tol


@pytest.mark.parametrize("p0,p1", [(6, 12), (16, 32)])
def test_transforms_mfma_vs_valu(hip, p0, p1):
    nf = 37
    Sh = RNG.standard_normal((nf, p0 * (p0 + 2)))
    outs = []
    gS = cu(Sh)
    for kind, tile in ((0, -1), (1, -1), (1, 0), (1, 3), (1, 7)):
        call("ves3d_hip_set_gemm_kernel", kind)
        call("ves3d_hip_set_gemm_tile", tile)
        try:
            g = torch.empty(nf * 2 * p1 * (p1 + 1), dtype=torch.float64, device="cuda")
            c = torch.empty(nf * p0 * (p0 + 2), dtype=torch.float64, device="cuda")
            S.shc2grid(gS, nf, p0, p1, g)
            S.grid2shc(g, nf, p1, p0, c)
            torch.cuda.synchronize()
            outs.append((g.cpu().numpy(), c.cpu().numpy()))
        finally:
            call("ves3d_hip_set_gemm_kernel", 1)
            call("ves3d_hip_set_gemm_tile", -1)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.max(np.abs(a - b)) <= 1e-13 * max(1.0, np.max(np.abs(a)))
    assert np.max(np.abs(outs[1][0].reshape(nf, -1) - sph.shc2grid(Sh, p0, p1))) < 1e-12 * \
        max(1.0, np.max(np.abs(outs[1][0])))


def test_singular_build_tiles(hip):
    """The singular-operator build (grouped rotation and transform GEMMs of
    every shape) with the automatic MFMA tiles against the 64x64 tile and
    the VALU kernel: FP64 to 1e-12 of the largest entry, FP32 to 1e-5."""
    from ves3d_amd import workloads as W
    nv, p = 3, 8
    X = torch.from_numpy(W.perturbed_shapes(nv, p, amp=0.1, seed=5)).cuda()
    n = 3 * p * (p + 2)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        res = []
        for kind, tile in ((1, -1), (1, 0), (1, 5), (0, -1)):
            call("ves3d_hip_set_gemm_kernel", kind)
            call("ves3d_hip_set_gemm_tile", tile)
            try:
                SL = torch.empty((nv, n, n), dtype=dt, device="cuda")
                DL = torch.empty((nv, n, n), dtype=dt, device="cuda")
                print("singular build %s kind %d tile %d" % (dt, kind, tile), flush=True)
                S.singular_build(X, nv, p, 2 * p, SL, DL)
                torch.cuda.synchronize()
                res.append((SL.double().cpu().numpy(), DL.double().cpu().numpy()))
            finally:
                call("ves3d_hip_set_gemm_kernel", 1)
                call("ves3d_hip_set_gemm_tile", -1)
        for o in res[1:]:
            for a, b in zip(res[0], o):
                assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(a))
