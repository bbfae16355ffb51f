"""Input formats (host code of the library, no GPU): the ASCII reader against
numpy's parser on the reference's own data files (fixtures cut from
precomputed/shape_gallery_6.txt and geometry_spec_sed_rand_512.txt), and
InitializeShapes against its definition (X = Rz(c) Ry(b) Rz(a) (s x) + centre,
src/HelperFuns.cc:314-432): rigid motions keep the shape's pairwise
distances times the scale and move its centroid to R (s c0) + centre."""
import ctypes
import math
import os

import numpy as np

import ves3d_amd
from ves3d_amd import call

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _read(path):
    n = ctypes.c_size_t()
    call("ves3d_read_ascii", path.encode(), None, 0, ctypes.byref(n))
    out = np.zeros(n.value)
    call("ves3d_read_ascii", path.encode(), out.ctypes.data_as(ctypes.c_void_p), out.size,
         ctypes.byref(n))
    return out


def test_read_ascii_matches_numpy():
    for fn in ("shape_gallery_6_head.txt", "geometry_spec_head.txt"):
        path = os.path.join(GOLD, fn)
        got = _read(path)
        want = np.loadtxt(path, comments="#").ravel()
        assert np.array_equal(got, want), fn
    assert _read(os.path.join(GOLD, "shape_gallery_6_head.txt")).size == 2 * 3 * 84


def test_read_ascii_rejects_missing_file_and_small_buffer():
    n = ctypes.c_size_t()
    rc = ves3d_amd.load().ves3d_read_ascii(b"/nonexistent/file.txt", None, 0, ctypes.byref(n))
    assert rc == 10
    buf = np.zeros(3)
    rc = ves3d_amd.load().ves3d_read_ascii(os.path.join(GOLD, "geometry_spec_head.txt").encode(),
                                           buf.ctypes.data_as(ctypes.c_void_p), 3, ctypes.byref(n))
    assert rc == 5


def _rz(a):
    return np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])


def _ry(a):
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def test_initialize_shapes():
    p = 6
    ng = 2 * p * (p + 1)
    gallery = _read(os.path.join(GOLD, "shape_gallery_6_head.txt"))
    spec = _read(os.path.join(GOLD, "geometry_spec_head.txt")).reshape(-1, 8)
    spec[1, 0] = 1.0                                   # use the second shape once
    spec[2, 5:] = [0.3, -1.1, 2.0]                     # distinct z, y, z angles
    nv = spec.shape[0]
    X = np.zeros((nv, 3, ng))
    call("ves3d_initialize_shapes", p, nv, gallery.ctypes.data_as(ctypes.c_void_p), gallery.size,
         np.ascontiguousarray(spec).ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p))
    shapes = gallery.reshape(-1, 3, ng)
    for v in range(nv):
        s = shapes[int(spec[v, 0])]
        R = _rz(spec[v, 7]) @ _ry(spec[v, 6]) @ _rz(spec[v, 5])
        want = R @ (spec[v, 1] * s) + spec[v, 2:5, None]
        assert np.allclose(X[v], want, atol=1e-13), v
        d0 = np.linalg.norm(s[:, :1] - s, axis=0) * spec[v, 1]
        d1 = np.linalg.norm(X[v][:, :1] - X[v], axis=0)
        assert np.allclose(d0, d1, atol=1e-12)
    bad = spec.copy()
    bad[0, 0] = 7.0                                    # no such shape
    rc = ves3d_amd.load().ves3d_initialize_shapes(p, nv, gallery.ctypes.data_as(ctypes.c_void_p),
                                                  gallery.size, bad.ctypes.data_as(ctypes.c_void_p),
                                                  X.ctypes.data_as(ctypes.c_void_p))
    assert rc == 10


def test_read_vesicle_props():
    """The reference's vesicle-properties file (precomputed/
    vesprops_rand_polysed_512.txt): one row (bending modulus, viscosity
    contrast, excess density) per vesicle, returned by property."""
    from ves3d_amd import workloads as W
    path = os.path.join(GOLD, "vesprops_rand_polysed_512.txt")
    want = np.loadtxt(path)
    for nv in (512, 17):
        kb, vc, rho = W.read_vesicle_props(path, nv)
        assert np.array_equal(kb, want[:nv, 0])
        assert np.array_equal(vc, want[:nv, 1])
        assert np.array_equal(rho, want[:nv, 2])
    out = np.zeros(3 * 513)
    rc = ves3d_amd.load().ves3d_read_vesicle_props(path.encode(), 513,
                                                   out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 51  # IOBadStream: fewer than 3 nves numbers


def test_polydisperse_suspension_inputs():
    """The reference's polydisperse sedimentation inputs
    (experiment/polydisperse_config.in): 512 copies of the p = 16 gallery's
    first shape, scaled 0.65 or 1.05, whose positions InitializeShapes
    builds; the vesicles are disjoint (every pair's closest points are
    farther apart than 0.01)."""
    p = 16
    gallery = _read(os.path.join(GOLD, "shape_gallery_16_shape0.txt"))
    spec = _read(os.path.join(GOLD, "geometry_spec_polysed_512.txt"))
    assert gallery.size == 3 * 2 * p * (p + 1) and spec.size == 8 * 512
    from ves3d_amd import workloads as W
    X = W.suspension_from_files(os.path.join(GOLD, "shape_gallery_16_shape0.txt"),
                                os.path.join(GOLD, "geometry_spec_polysed_512.txt"), p)
    assert X.shape == (512, 3, 2 * p * (p + 1))
    c = X.mean(axis=2)
    r = np.sqrt(np.max(np.sum((X - c[:, :, None]) ** 2, axis=1), axis=1))
    d = np.linalg.norm(c[:, None] - c[None], axis=2)
    close = np.argwhere((d < r[:, None] + r[None, :]) & (np.arange(512)[:, None] < np.arange(512)))
    for a, b in close:
        dmin = np.min(np.linalg.norm(X[a][:, :, None] - X[b][:, None, :], axis=0))
        assert dmin > 0.01, (a, b, dmin)
