"""Synthetic vesicle suspensions (host-side input generation, numpy only).

Grids follow the reference's SphericalHarmonics convention: latitudes are the
p+1 Gauss-Legendre nodes x_i in ascending order, longitudes phi_j = pi j / p
(j < 2p), point index i*2p + j.  Shapes are embedded as the reference's
shape files are (precomputed/shape_gallery_*.txt): the point (i, j) of the
unit sphere is (sqrt(1-x_i^2) cos phi_j, sqrt(1-x_i^2) sin phi_j, x_i), so the
polar axis is z and the grid's x_u x x_v normal points outward.
"""
import numpy as np

SEED_BASE = 20261202


def gl_nodes(p):
    """Gauss-Legendre nodes (ascending) and weights with p+1 points."""
    x, w = np.polynomial.legendre.leggauss(p + 1)
    return x, w


def sphere_grid(p):
    """Unit-sphere grid [3][2p(p+1)] and the smooth quadrature weight per
    point (w_i * pi/p, which integrates dS exactly for polynomials).

    Point (i, j) is (rho cos phi_j, rho sin phi_j, x_i) with x_i the ascending
    Gauss-Legendre nodes and phi_j = pi j / p: the orientation of the
    reference's shape files (precomputed/shape_gallery_*.txt), for which the
    grid's x_u x x_v normal points outward and a unit sphere has H = -1."""
    x, w = gl_nodes(p)
    phi = np.pi * np.arange(2 * p) / p
    s = np.sqrt(1.0 - x * x)
    X = np.empty((3, p + 1, 2 * p))
    X[0] = s[:, None] * np.cos(phi)[None, :]
    X[1] = s[:, None] * np.sin(phi)[None, :]
    X[2] = x[:, None]
    qw = np.repeat(w * np.pi / p, 2 * p)
    return X.reshape(3, -1), qw


def random_centres(nv, volfrac, radius=1.0, seed=SEED_BASE, gap=0.0):
    """Centres uniform in a cube sized for `volfrac` (sphere volume fraction),
    overlapping spheres rejected (centre distance < 2 radius + gap)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    side = (nv * 4.0 / 3.0 * np.pi * radius ** 3 / volfrac) ** (1.0 / 3.0)
    out = np.empty((nv, 3))
    k = 0
    dmin2 = (2 * radius + gap) ** 2
    tries = 0
    while k < nv:
        c = rng.uniform(radius, side - radius, size=3)
        if k == 0 or np.min(np.sum((out[:k] - c) ** 2, axis=1)) >= dmin2:
            out[k] = c
            k += 1
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place %d spheres at volume fraction %g" % (nv, volfrac))
    return out, side


def _radial_perturbation(p, amp, rng, lmax=None):
    """Band-limited random field on the grid: sum over l=1..lmax, m=0..l of
    (a cos m phi + b sin m phi) Pbar_l^m(x), a, b ~ N(0, (amp/(l+1)^2)^2),
    Pbar orthonormal (scipy lpmv with the normalisation factor)."""
    from scipy.special import lpmv, factorial
    lmax = p if lmax is None else lmax
    x, _ = gl_nodes(p)
    phi = np.pi * np.arange(2 * p) / p
    r = np.zeros((p + 1, 2 * p))
    for l in range(1, lmax + 1):
        sd = amp / (l + 1) ** 2
        for m in range(0, l + 1):
            nrm = np.sqrt((2 * l + 1) / (4 * np.pi) * factorial(l - m) / factorial(l + m))
            P = nrm * lpmv(m, l, x)
            a, b = rng.normal(0, sd, 2)
            r += P[:, None] * (a * np.cos(m * phi) + (b * np.sin(m * phi) if m > 0 else 0.0))[None, :]
    return r.reshape(-1)


def perturbed_shapes(nves, p, amp=0.05, seed=SEED_BASE, radius=1.0):
    """Unit spheres with a seeded radial perturbation (coefficients
    N(0, (amp/(l+1)^2)^2), l = 1..p).  Returns [nves][3][2p(p+1)] centred at 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    X, _ = sphere_grid(p)
    out = np.empty((nves, 3, X.shape[1]))
    for v in range(nves):
        rr = radius * (1.0 + _radial_perturbation(p, amp, rng))
        out[v] = X * rr[None, :]
    return out


def ellipsoid_shapes(nves, p, seed=SEED_BASE, c_mean=0.6, c_jitter=0.05, amp=0.0):
    """Ellipsoids with semi-axes (1, 1, c), c = c_mean +- c_jitter uniform,
    the short axis along the grid's polar axis (z).  Returns the shapes
    [nves][3][2p(p+1)] and their volumes (4/3 pi c)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    X, _ = sphere_grid(p)
    out = np.empty((nves, 3, X.shape[1]))
    vol = np.empty(nves)
    for v in range(nves):
        c = c_mean + rng.uniform(-c_jitter, c_jitter)
        rr = 1.0 + (_radial_perturbation(p, amp, rng) if amp > 0 else np.zeros(X.shape[1]))
        out[v] = X * rr[None, :]
        out[v, 2] *= c
        vol[v] = 4.0 / 3.0 * np.pi * c
    return out, vol


def random_rotation(rng):
    """Uniformly distributed rotation matrix (QR of a Gaussian matrix with
    the sign convention that makes the factorisation unique)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def suspension(shapes, volfrac, seed=SEED_BASE, gap=0.05, volumes=None, rotate=False):
    """Places the vesicle shapes [nv][3][np] (centred at 0) at rejection-
    sampled centres in a cube sized so that the shapes fill `volfrac` of it
    (volumes: per-shape volumes, default that of the bounding sphere); the
    bounding spheres plus `gap` may not overlap.  rotate: apply a seeded
    random rotation to every shape first.  Returns positions [nv][3][np]."""
    nv = shapes.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed + 77))
    if rotate:
        shapes = np.stack([random_rotation(rng) @ s for s in shapes])
    rad = float(np.max(np.sqrt(np.sum(shapes ** 2, axis=1))))
    if volumes is None:
        volumes = np.full(nv, 4.0 / 3.0 * np.pi * rad ** 3)
    eq = float(np.sum(volumes)) / nv / (4.0 / 3.0 * np.pi * rad ** 3)  # per-shape fill of its ball
    ball_frac = volfrac / eq
    if ball_frac <= 0.25:
        c, _ = random_centres(nv, ball_frac, rad, seed, gap)
    else:
        # dense: random sequential addition of the bounding balls jams near
        # this fraction, so centres go on a jittered cubic lattice instead
        # (jitter bounded so that bounding balls + gap stay disjoint)
        # (a box of a x b x c sites, cubic when that fits, else the smallest
        # site count >= nv, which packs denser when nv is not a cube)
        k = int(np.ceil(nv ** (1.0 / 3.0)))
        total = float(np.sum(volumes)) / volfrac
        dims = (k, k, k)
        h = total ** (1.0 / 3.0) / k
        if h < 2 * rad + gap:
            best = None
            for a in range(1, k + 1):
                for b in range(a, k + 1):
                    c = -(-nv // (a * b))
                    if c < b:
                        continue
                    key = (a * b * c, c - a)
                    if best is None or key < best[0]:
                        best = (key, (c, b, a))
            dims = best[1]
            h = (total / (dims[0] * dims[1] * dims[2])) ** (1.0 / 3.0)
        if h < 2 * rad + gap:
            raise RuntimeError("volume fraction %g too dense for a lattice of bounding balls" % volfrac)
        jit = 0.5 * (h - 2 * rad - gap)
        idx = np.stack(np.meshgrid(np.arange(dims[0]), np.arange(dims[1]), np.arange(dims[2]),
                                   indexing="ij"), -1)
        idx = idx.reshape(-1, 3)[rng.permutation(dims[0] * dims[1] * dims[2])[:nv]]
        c = (idx + 0.5) * h + rng.uniform(-jit, jit, size=(nv, 3))
    return shapes + c[:, :, None]


def reference_cpu_case(p=12, seed=SEED_BASE):
    """The plain method's CPU-runnable case: 2 unit spheres with seeded radial
    perturbations (coefficients N(0, (0.05/(l+1)^2)^2), l = 1..p), centred at
    (-1.25, 0, 0) and (1.25, 0, 0).  Returns positions [2][3][2p(p+1)]; the
    step is bending modulus 0.01, shear u = (0.1 z, 0, 0), dt = 1e-3,
    BiCGStab to 1e-8."""
    shapes = perturbed_shapes(2, p, amp=0.05, seed=seed)
    c = np.array([[-1.25, 0.0, 0.0], [1.25, 0.0, 0.0]])
    return shapes + c[:, :, None]


def implicit_step_case(nv=512, p=16, volfrac=0.20, seed=SEED_BASE + 3):
    """The full implicit-step workload: nv perturbed ellipsoids (semi-axes
    1, 1, 0.6 +- 0.05) with random orientations at the given volume fraction.
    Returns positions [nv][3][2p(p+1)]."""
    shapes, vol = ellipsoid_shapes(nv, p, seed=seed)
    return suspension(shapes, volfrac, seed=seed, gap=0.05, volumes=vol, rotate=True)


def stokes_direct_sum_case(nv=512, p=16, volfrac=0.15, seed=SEED_BASE + 1):
    """The 'Stokes direct summation alone' workload: nv unit spheres at SH
    order p (2p(p+1) points each) with centres uniform in a cube at the given
    volume fraction; single-layer densities U[-1,1]^3 times the quadrature
    weight, double-layer densities the same with unit normals.

    Returns point-major arrays: pos [N,3], sl [N,3], dl [N,6] (=[g, n])."""
    X, qw = sphere_grid(p)
    c, _ = random_centres(nv, volfrac, 1.0, seed)
    npv = X.shape[1]
    pos = (c[:, None, :] + X.T[None, :, :]).reshape(-1, 3)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    w = np.tile(qw, nv)[:, None]
    sl = rng.uniform(-1, 1, size=(nv * npv, 3)) * w
    g = rng.uniform(-1, 1, size=(nv * npv, 3)) * w
    n = np.tile(X.T, (nv, 1))
    dl = np.concatenate([g, n], axis=1)
    return np.ascontiguousarray(pos), np.ascontiguousarray(sl), np.ascontiguousarray(dl), npv


def read_ascii(path):
    """All numbers of an ASCII data file ('#' comments), through the
    library's DataIO::ReadDataStl equivalent."""
    import ctypes
    from . import call
    n = ctypes.c_size_t()
    call("ves3d_read_ascii", path.encode(), None, 0, ctypes.byref(n))
    out = np.zeros(n.value)
    call("ves3d_read_ascii", path.encode(), out.ctypes.data_as(ctypes.c_void_p), out.size,
         ctypes.byref(n))
    return out


def suspension_from_files(gallery_path, spec_path, p, nves=None):
    """Initial positions [nves][3][2p(p+1)] from a shape gallery and a
    geometry specification file in the reference's formats (Simulation::
    setup_from_options, src/ves3d_simulation.cc:126-168)."""
    import ctypes
    from . import call
    gallery = read_ascii(gallery_path)
    spec = read_ascii(spec_path)
    nv = spec.size // 8 if nves is None else nves
    if spec.size < 8 * nv:
        raise ValueError("geometry file holds %d vesicles, %d requested" % (spec.size // 8, nv))
    X = np.zeros((nv, 3, 2 * p * (p + 1)))
    spec = np.ascontiguousarray(spec[:8 * nv])
    call("ves3d_initialize_shapes", p, nv, gallery.ctypes.data_as(ctypes.c_void_p), gallery.size,
         spec.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p))
    return X


def read_vesicle_props(path, nves):
    """The vesicle-properties file (src/ves3d_simulation.cc:174-196):
    (bending modulus, viscosity contrast, excess density), each [nves]."""
    import ctypes
    from . import call
    out = np.zeros(3 * nves)
    call("ves3d_read_vesicle_props", path.encode(), nves, out.ctypes.data_as(ctypes.c_void_p))
    return out[:nves].copy(), out[nves:2 * nves].copy(), out[2 * nves:].copy()
