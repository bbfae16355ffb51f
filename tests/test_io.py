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
