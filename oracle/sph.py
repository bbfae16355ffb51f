"""FP64 numpy restatement of the reference's SphericalHarmonics<Real>
(include/SphericalHarmonics.h, src/SphericalHarmonics.cc).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product.

Layouts (as in the reference):
  grid  : per function (p+1) x 2p values, latitude-major; latitude i is the
          i-th Gauss-Legendre node x_i (ascending), longitude j is phi = pi j/p.
  coeffs: per function p(p+2) values: for j = 0..2p-1 a block of length
          p+1-(j+1)//2: j=0 -> m=0 (cos), j=2m-1 -> m cos, j=2m -> m sin,
          j=2p-1 -> m=p (cos).  Within a block, degree n = m..p.
          (src/SphericalHarmonics.cc:30-40)
Every matrix builder caches its result like the reference's MatrixStorage.
"""
import functools

import numpy as np

SQRT2PI = np.sqrt(2 * np.pi)


# ---------------------------------------------------------------- quadrature
@functools.lru_cache(maxsize=None)
def legendre_nodes_weights(p1):
    """cgqf(p1+1, 1, 0, 0, -1, 1) (src/legendre_rule.cc:269): Gauss-Legendre
    nodes (ascending) and weights on [-1,1].  Restated with Newton iteration on
    the three-term recurrence (same nodes; the reference uses Golub-Welsch)."""
    n = p1 + 1
    x = np.cos(np.pi * (np.arange(n) + 0.75) / (n + 0.5))
    for _ in range(100):
        p0 = np.ones_like(x)
        pk = x.copy()
        for k in range(2, n + 1):
            p0, pk = pk, ((2 * k - 1) * x * pk - (k - 1) * p0) / k
        if n == 1:
            p0, pk = np.ones_like(x), x
        dp = n * (x * pk - p0) / (x * x - 1)
        dx = pk / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    p0 = np.ones_like(x)
    pk = x.copy()
    for k in range(2, n + 1):
        p0, pk = pk, ((2 * k - 1) * x * pk - (k - 1) * p0) / k
    if n == 1:
        p0, pk = np.ones_like(x), x
    dp = n * (x * pk - p0) / (x * x - 1)
    w = 2.0 / ((1 - x * x) * dp * dp)
    order = np.argsort(x)
    return x[order].copy(), w[order].copy()


def legendre_nodes(p1):
    """LegendreNodes (src/SphericalHarmonics.cc:543-555)."""
    return legendre_nodes_weights(p1)[0]


@functools.lru_cache(maxsize=None)
def legendre_weights(p1):
    """LegendreWeights: w_i * pi/p1 / sqrt(1-x_i^2) (src/SphericalHarmonics.cc:557-571)."""
    x, w = legendre_nodes_weights(p1)
    return w * np.pi / p1 / np.sqrt(1 - x * x)


@functools.lru_cache(maxsize=None)
def singular_weights(p1):
    """SingularWeights (src/SphericalHarmonics.cc:573-604)."""
    x, w = legendre_nodes_weights(p1)
    alp0 = legpoly(np.array([1.0]), p1)      # [(p1+1)(p1+2)/2][1]
    alp = legpoly(x, p1)                     # [(p1+1)(p1+2)/2][p1+1]
    Yf = np.zeros(p1 + 1)
    for j in range(p1 + 1):
        Yf += 4 * np.pi / (2 * j + 1) * alp0[j, 0] * alp[j]
    return (w * np.pi / p1) * Yf / np.cos(np.arccos(x) / 2)


# ---------------------------------------------------------------- Legendre
def legpoly(X, degree):
    """LegPoly (src/SphericalHarmonics.cc:888-928): normalised associated
    Legendre functions, rows ordered (m=0: n=0..deg), (m=1: n=1..deg), ...
    Returns [(deg+1)(deg+2)/2][len(X)]."""
    X = np.asarray(X, dtype=np.float64)
    N = X.size
    out = np.zeros(((degree + 1) * (degree + 2) // 2, N))
    u = np.sqrt(np.maximum(0.0, 1 - X * X))
    u[X * X > 1.0] = 0
    fact = 1.0 / np.sqrt(4 * np.pi)
    starts = [0]
    for m in range(1, degree + 1):
        starts.append(starts[-1] + (degree - m + 2))
    out[0] = fact
    for i in range(1, degree + 1):
        c = np.sqrt(3.0 / 2.0) if i == 1 else 1.0
        if i > 1:
            c *= np.sqrt((2 * i + 1) / (2.0 * i))
        out[starts[i]] = -out[starts[i - 1]] * u * c
    for m in range(degree):
        base = starts[m]
        pmm = np.zeros(N)
        pmmp1 = out[base].copy()
        for ll in range(m + 1, degree + 1):
            a = np.sqrt(((2 * ll - 1) * (2 * ll + 1)) / ((ll - m) * (ll + m)))
            b = np.sqrt(((2 * ll + 1) * (ll + m - 1) * (ll - m - 1)) / ((ll - m) * (ll + m) * (2 * ll - 3)))
            pll = X * a * pmmp1 - b * pmm
            pmm = pmmp1
            pmmp1 = pll
            out[base + ll - m] = pll
    return out


def legpoly_deriv(X, degree):
    """LegPolyDeriv (src/SphericalHarmonics.cc:930-950)."""
    X = np.asarray(X, dtype=np.float64)
    lp = legpoly(X, degree)
    out = np.zeros_like(lp)
    s = np.sqrt(1 - X * X)

    def idx(m, n):
        return ((degree * 2 - m + 1) * m) // 2 + n

    for m in range(degree + 1):
        for n in range(m, degree + 1):
            Pn = lp[idx(m, n)]
            Pn_ = lp[idx(m + 1, n)] if m + 1 <= n else lp[0]
            c1 = float(m)
            c2 = np.sqrt(n + m + 1) * np.sqrt(n - m if n > m else 1) if (m + 1) <= n else 0.0
            out[idx(m, n)] = -(c1 * X * Pn + c2 * s * Pn_) / s
    return out


def legpoly_ladder(X, degree):
    """First and second latitude derivatives of sqrt-normalised Legendre
    functions by the ladder of matlab/Ynm.m:18-40 (the generator of the
    reference's d1legTrans / d2legTrans files):
        DP_m  = c_m P_{m-1} - c_{m+1} P_{m+1},  c_k = sqrt((n+k)(n+1-k))/2,
        D2P_m = c_m DP_{m-1} - c_{m+1} DP_{m+1},
    with P_{-m} = (-1)^m P_m.  Returns (H, W) = (-DP, D2P) in the LegPoly row
    order; H equals -legpoly_deriv."""
    X = np.asarray(X, dtype=np.float64)
    P = legpoly(X, degree)
    H = np.zeros_like(P)
    Wd = np.zeros_like(P)

    def idx(m, n):
        return ((degree * 2 - m + 1) * m) // 2 + n

    for n in range(degree + 1):
        full = {}
        for m in range(-n - 1, n + 2):
            if abs(m) > n:
                full[m] = np.zeros_like(X)
            elif m >= 0:
                full[m] = P[idx(m, n)]
            else:
                full[m] = (-1) ** (-m) * P[idx(-m, n)]
        c = {k: np.sqrt(max((n + k) * (n + 1 - k), 0)) / 2.0 for k in range(-n - 1, n + 3)}
        DP = {}
        for m in range(-n - 1, n + 2):
            DP[m] = np.zeros_like(X) if abs(m) > n else c[m] * full[m - 1] - c[m + 1] * full[m + 1]
        for m in range(0, n + 1):
            H[idx(m, n)] = -DP[m]
            Wd[idx(m, n)] = c[m] * DP[m - 1] - c[m + 1] * DP[m + 1]
    return H, Wd


@functools.lru_cache(maxsize=None)
def mat_legendre_d2(p0, p1):
    """Second latitude-derivative analogue of MatLegendre (the d2legTrans
    matrices of src/OperatorsMats.cc:26-93, generated by matlab/Ynm.m)."""
    return tuple(_split_blocks(legpoly_ladder(legendre_nodes(p1), p0)[1], p0))


@functools.lru_cache(maxsize=None)
def mat_fourier_d2(p0, p1):
    """Second longitude derivative of MatFourier (src/SHTMats.cc:160-187 up to
    the sqrt(2 pi) factor SphericalHarmonics carries in the Fourier matrix)."""
    M = np.zeros((2 * p0, 2 * p1))
    j = np.arange(2 * p1)
    for k in range(1, p0):
        M[2 * k - 1] = -SQRT2PI * k * k * np.cos(j * k * np.pi / p1)
        M[2 * k] = -SQRT2PI * k * k * np.sin(j * k * np.pi / p1)
    M[2 * p0 - 1] = -SQRT2PI * p0 * p0 * np.cos(j * p0 * np.pi / p1)
    return M


# ---------------------------------------------------------------- matrices
@functools.lru_cache(maxsize=None)
def mat_fourier(p0, p1):
    """MatFourier (src/SphericalHarmonics.cc:606-627): [2p0][2p1]."""
    M = np.zeros((2 * p0, 2 * p1))
    j = np.arange(2 * p1)
    M[0] = SQRT2PI
    for k in range(1, p0):
        M[2 * k - 1] = SQRT2PI * np.cos(j * k * np.pi / p1)
        M[2 * k] = SQRT2PI * np.sin(j * k * np.pi / p1)
    M[2 * p0 - 1] = SQRT2PI * np.cos(j * p0 * np.pi / p1)
    return M


@functools.lru_cache(maxsize=None)
def mat_fourier_inv(p0, p1):
    """MatFourierInv (src/SphericalHarmonics.cc:629-653): [2p0][2p1]."""
    c = 1.0 / np.sqrt(2 * np.pi) / p0
    M = np.zeros((2 * p0, 2 * p1))
    q = min(p1, p0)
    j = np.arange(2 * p0)
    M[:, 0] = c * 0.5
    for k in range(1, q):
        M[:, 2 * k - 1] = c * np.cos(j * k * np.pi / p0)
        M[:, 2 * k] = c * np.sin(j * k * np.pi / p0)
    M[:, 2 * q - 1] = c * np.cos(j * q * np.pi / p0)
    if q == p0:
        M[:, 2 * q - 1] *= 0.5
    return M


@functools.lru_cache(maxsize=None)
def mat_fourier_grad(p0, p1):
    """MatFourierGrad (src/SphericalHarmonics.cc:655-676): [2p0][2p1]."""
    M = np.zeros((2 * p0, 2 * p1))
    j = np.arange(2 * p1)
    for k in range(1, p0):
        M[2 * k - 1] = -SQRT2PI * k * np.sin(j * k * np.pi / p1)
        M[2 * k] = SQRT2PI * k * np.cos(j * k * np.pi / p1)
    M[2 * p0 - 1] = -SQRT2PI * p0 * np.sin(j * p0 * np.pi / p1)
    return M


def _split_blocks(alp, p0):
    out = []
    off = 0
    for i in range(p0 + 1):
        out.append(alp[off:off + p0 + 1 - i])
        off += p0 + 1 - i
    return out


@functools.lru_cache(maxsize=None)
def mat_legendre(p0, p1):
    """MatLegendre (src/SphericalHarmonics.cc:678-701): for m=0..p0 a
    [(p0+1-m)][(p1+1)] block of P_n^m at the p1 nodes."""
    return tuple(_split_blocks(legpoly(legendre_nodes(p1), p0), p0))


@functools.lru_cache(maxsize=None)
def mat_legendre_grad(p0, p1):
    """MatLegendreGrad (src/SphericalHarmonics.cc:735-757)."""
    return tuple(_split_blocks(legpoly_deriv(legendre_nodes(p1), p0), p0))


@functools.lru_cache(maxsize=None)
def mat_legendre_inv(p0, p1):
    """MatLegendreInv (src/SphericalHarmonics.cc:703-733): for m=0..p1 a
    [(p0+1)][(p1+1-m)] block, transposed and times w_k 2 pi."""
    x, w = legendre_nodes_weights(p0)
    blocks = _split_blocks(legpoly(x, p1), p1)
    return tuple((b * (w[None, :] * 2 * np.pi)).T.copy() for b in blocks)


def block_lengths(p):
    return [p + 1 - (j + 1) // 2 for j in range(2 * p)]


def block_offsets(p):
    L = block_lengths(p)
    return np.concatenate([[0], np.cumsum(L)[:-1]]).astype(int), L


# ---------------------------------------------------------------- transforms
def _legendre_stage(B0blocks, Ml, N):
    """Per-order products B1_j = B0_j @ Ml[m] (j = cos / sin blocks of order m)."""
    out = []
    for j, b in enumerate(B0blocks):
        m = (j + 1) // 2
        out.append(b @ Ml[m])
    return out


def shc2grid(S, p0, p1, want_theta=False, want_phi=False):
    """SHC2Grid (src/SphericalHarmonics.cc:3-149).  S: [N][p0(p0+2)] ->
    X: [N][(p1+1)*2p1]; optionally X_theta (Fourier derivative) and X_phi
    (Legendre derivative), named as in the reference."""
    S = np.asarray(S).reshape(-1, p0 * (p0 + 2))
    N = S.shape[0]
    off, L = block_offsets(p0)
    B0 = [S[:, off[j]:off[j] + L[j]] for j in range(2 * p0)]
    Mf = mat_fourier(p0, p1)
    B1 = np.stack(_legendre_stage(B0, mat_legendre(p0, p1), N), axis=0)  # [2p0][N][p1+1]
    B1t = B1.reshape(2 * p0, -1).T                                        # [N(p1+1)][2p0]
    X = (B1t @ Mf).reshape(N, -1)
    res = [X]
    if want_theta:
        res.append((B1t @ mat_fourier_grad(p0, p1)).reshape(N, -1))
    if want_phi:
        B1d = np.stack(_legendre_stage(B0, mat_legendre_grad(p0, p1), N), axis=0)
        res.append((B1d.reshape(2 * p0, -1).T @ Mf).reshape(N, -1))
    return res[0] if len(res) == 1 else tuple(res)


def shc2grid_d(S, p0, p1, dleg=0, dfour=0):
    """SHC2Grid with the Legendre stage differentiated dleg times (latitude
    direction) and the Fourier stage dfour times (longitude), dleg, dfour in
    {0, 1, 2}: the synthesis behind SHTrans::backward, backward_du / dv /
    d2u / d2v / duv (src/SHTrans.cc:123-170)."""
    S = np.asarray(S).reshape(-1, p0 * (p0 + 2))
    N = S.shape[0]
    off, L = block_offsets(p0)
    Ml = (mat_legendre, mat_legendre_grad, mat_legendre_d2)[dleg](p0, p1)
    Mf = (mat_fourier, mat_fourier_grad, mat_fourier_d2)[dfour](p0, p1)
    B0 = [S[:, off[j]:off[j] + L[j]] for j in range(2 * p0)]
    B1 = np.stack(_legendre_stage(B0, Ml, N), axis=0)
    return (B1.reshape(2 * p0, -1).T @ Mf).reshape(N, -1)


def _grid2shc_with(X, p0, p1, Mf, Ml):
    X = np.asarray(X).reshape(-1, 2 * p0 * (p0 + 1))
    N = X.shape[0]
    B2 = X.reshape(N * (p0 + 1), 2 * p0) @ Mf          # [N(p0+1)][2p1]
    B1 = B2.T.reshape(2 * p1, N, p0 + 1)               # [2p1][N][p0+1]
    S = np.zeros((N, p1 * (p1 + 2)))
    off, L = block_offsets(p1)
    for j in range(2 * p1):
        m = (j + 1) // 2
        S[:, off[j]:off[j] + L[j]] = B1[j] @ Ml[m]
    return S


def grid2shc(X, p0, p1):
    """Grid2SHC (src/SphericalHarmonics.cc:151-235).  X: [N][(p0+1)2p0] -> S: [N][p1(p1+2)]."""
    return _grid2shc_with(X, p0, p1, mat_fourier_inv(p0, p1), mat_legendre_inv(p0, p1))


def shc2grid_transpose(X, p0, p1):
    """SHC2GridTranspose (src/SphericalHarmonics.cc:237-322): the adjoint of
    SHC2Grid(p1 -> p0).  X: [N][(p0+1)2p0] -> S: [N][p1(p1+2)]."""
    Mf = mat_fourier(p1, p0).T
    Ml = tuple(b.T for b in mat_legendre(p1, p0))
    return _grid2shc_with(X, p0, p1, Mf, Ml)


def shc2pole(S, p0):
    """SHC2Pole (src/SphericalHarmonics.cc:324-360): [N][p0(p0+2)] -> [N][2]
    values at x = -1 (index 0) and x = +1 (index 1)."""
    S = np.asarray(S).reshape(-1, p0 * (p0 + 2))
    qp = np.stack([legpoly(np.array([-1.0]), p0)[: p0 + 1, 0],
                   legpoly(np.array([1.0]), p0)[: p0 + 1, 0]], axis=1) * SQRT2PI
    return S[:, : p0 + 1] @ qp


# ---------------------------------------------------------------- rotation
@functools.lru_cache(maxsize=None)
def coeff_perm(p0):
    """coeff_perm (src/SphericalHarmonics.cc:761-772): for degree n the list
    of coefficient indices (over j = 0..min(2n+1,2p0)-1) holding degree n."""
    perm = [np.zeros(min(2 * n + 1, 2 * p0), dtype=int) for n in range(p0 + 1)]
    itr = 0
    for i in range(2 * p0):
        m = (i + 1) // 2
        for n in range(m, p0 + 1):
            perm[n][i] = itr
            itr += 1
    return tuple(perm)


@functools.lru_cache(maxsize=None)
def mat_rotate(p0):
    """MatRotate (src/SphericalHarmonics.cc:759-860): per latitude l and
    degree n the block Mr[l*(p0+1)+n] (coefficient -> coefficient) of the
    rotation that brings grid latitude l to the pole."""
    Ncoef = p0 * (p0 + 2)
    Ngrid = 2 * p0 * (p0 + 1)
    x = legendre_nodes(p0)
    lat = np.repeat(x, 2 * p0)
    lon = np.tile(np.pi * np.arange(2 * p0) / p0, p0 + 1)
    s = np.sqrt(1 - lat * lat)
    C0 = np.stack([lat, s * np.sin(lon), s * np.cos(lon)])
    perm = coeff_perm(p0)
    Mr = []
    for l in range(p0 + 1):
        c = -x[l]
        sn = -np.sqrt(1.0 - x[l] * x[l])
        R = np.array([[c, 0, -sn], [0, 1, 0], [sn, 0, c]])
        C1 = R @ C0
        Mleg = legpoly(C1[0], p0)
        theta = np.arctan2(C1[1], C1[2])
        M = np.zeros((Ncoef, Ngrid))
        o0 = 0
        o1 = 0
        for i in range(p0 + 1):
            ln = p0 + 1 - i
            M[o1:o1 + ln] = SQRT2PI * Mleg[o0:o0 + ln] * np.cos(i * theta)[None, :]
            o1 += ln
            if i != 0 and i != p0:
                M[o1:o1 + ln] = SQRT2PI * Mleg[o0:o0 + ln] * np.sin(i * theta)[None, :]
                o1 += ln
            o0 += ln
        C2C = grid2shc(M, p0, p0)   # [Ncoef][Ncoef]
        for n in range(p0 + 1):
            pn = perm[n]
            Mr.append(C2C[np.ix_(pn, pn)].copy())
    return tuple(Mr)


def _phase(S0, p0, j, sign):
    """Multiplies each order-k (cos, sin) pair by exp(sign * i k j pi/p0) as in
    RotateAll / RotateTranspose."""
    out = S0.copy()
    off = 0
    for k in range(p0 + 1):
        ln = p0 + 1 - k
        if k != 0 and k != p0:
            r0 = np.cos(k * j * np.pi / p0)
            r1 = sign * np.sin(k * j * np.pi / p0)
            x0 = S0[..., off:off + ln]
            x1 = S0[..., off + ln:off + 2 * ln]
            out[..., off:off + ln] = x0 * r0 - x1 * r1
            out[..., off + ln:off + 2 * ln] = x0 * r1 + x1 * r0
            off += 2 * ln
        else:
            off += ln
    return out


def rotate_all(S, p0, dof):
    """RotateAll (src/SphericalHarmonics.cc:362-449).  S: [N*dof][Ncoef] ->
    [N][p0+1 (t)][p0 (j)][dof][Ncoef]."""
    Ncoef = p0 * (p0 + 2)
    S = np.asarray(S).reshape(-1, dof, Ncoef)
    N = S.shape[0]
    Mr = mat_rotate(p0)
    perm = coeff_perm(p0)
    B0 = np.stack([_phase(S, p0, j, -1.0) for j in range(p0)], axis=1)   # [N][p0][dof][Ncoef]
    out = np.zeros((N, p0 + 1, p0, dof, Ncoef))
    for t in range(p0 + 1):
        for l in range(p0 + 1):
            pn = perm[l]
            ot = out[:, t]
            ot[..., pn] = B0[..., pn] @ Mr[t * (p0 + 1) + l]
    return out


def rotate_transpose(S_, p0, dof):
    """RotateTranspose (src/SphericalHarmonics.cc:451-541).
    [N][p0+1][p0][dof][Ncoef] -> same shape."""
    Ncoef = p0 * (p0 + 2)
    S_ = np.asarray(S_).reshape(-1, p0 + 1, p0, dof, Ncoef)
    Mr = mat_rotate(p0)
    perm = coeff_perm(p0)
    B = np.zeros_like(S_)
    for t in range(p0 + 1):
        for l in range(p0 + 1):
            pn = perm[l]
            bt = B[:, t]
            bt[..., pn] = S_[:, t][..., pn] @ Mr[t * (p0 + 1) + l].T
    out = np.zeros_like(S_)
    for j in range(p0):
        out[:, :, j] = _phase(B[:, :, j], p0, j, +1.0)
    return out


# ---------------------------------------------------------------- singular
def stokes_singular_integ(X0, p0, p1, want_sl=True, want_dl=True):
    """StokesSingularInteg_ (src/SphericalHarmonics.cc:952-1180) for vesicle
    positions X0 [Nves][3][(p0+1)2p0].  Returns (SL, DL), each
    [Nves][3*Ncoef][3*Ncoef] (rows: density coefficient d*Ncoef+k, columns:
    velocity coefficient d'*Ncoef+k'); None where not requested."""
    Ncoef = p0 * (p0 + 2)
    Ngrid = 2 * p0 * (p0 + 1)
    M1 = 2 * p1 * (p1 + 1)
    X0 = np.asarray(X0).reshape(-1, 3, Ngrid)
    Nv = X0.shape[0]
    S0 = grid2shc(X0.reshape(-1, Ngrid), p0, p0)
    S = rotate_all(S0, p0, 3)                       # [Nv][p0+1][p0][3][Ncoef]
    Sr = S.reshape(-1, Ncoef)
    X, Xt, Xp = shc2grid(Sr, p0, p1, True, True)    # [Nv*(p0+1)*p0*3][M1]
    trg = shc2pole(Sr, p0)                          # [..][2]
    nrot = Nv * (p0 + 1) * p0
    X = X.reshape(nrot, 3, M1)
    Xt = Xt.reshape(nrot, 3, M1)
    Xp = Xp.reshape(nrot, 3, M1)
    trg = trg.reshape(nrot, 3, 2)
    qw = singular_weights(p1)
    n = np.cross(Xt, Xp, axis=1)                    # [nrot][3][M1]
    area = np.sqrt(np.sum(n * n, axis=1))           # [nrot][M1]
    eps = np.finfo(np.float64).eps
    results = []
    for layer, want in (("sl", want_sl), ("dl", want_dl)):
        if not want:
            results.append(None)
            continue
        K0 = np.zeros((nrot, 2, 6, M1))
        for t in range(2):
            w = qw[np.arange(p1 + 1) * t + (p1 - np.arange(p1 + 1)) * (1 - t)]
            wgrid = np.repeat(w, 2 * p1)
            d = trg[:, :, t][:, :, None] - X              # [nrot][3][M1]
            r2 = np.sum(d * d, axis=1)
            rinv = 1.0 / np.sqrt(r2)
            rinv[r2 <= eps] = 0
            rinv2 = rinv * rinv
            dd = [d[:, 0] * d[:, 0], d[:, 0] * d[:, 1], d[:, 0] * d[:, 2],
                  d[:, 1] * d[:, 1], d[:, 1] * d[:, 2], d[:, 2] * d[:, 2]]
            if layer == "dl":
                rn = 3.0 / (4.0 * np.pi) * wgrid[None] * np.sum(n * d, axis=1) * rinv2 * rinv2 * rinv
                for c in range(6):
                    K0[:, t, c] = dd[c] * rn
            else:
                ar = 1.0 / (8.0 * np.pi) * wgrid[None] * area * rinv
                ar2 = ar * rinv2
                for c in range(6):
                    K0[:, t, c] = dd[c] * ar2 + (ar if c in (0, 3, 5) else 0)
        K1 = shc2grid_transpose(K0.reshape(-1, M1), p1, p0)          # [nrot*12][Ncoef]
        K2 = rotate_transpose(K1.reshape(Nv, p0 + 1, p0, 12, Ncoef), p0, 12)  # [Nv][t][j][t2*6+c][Ncoef]
        K2 = K2.reshape(Nv, p0 + 1, p0, 2, 6, Ncoef)
        # Rearrange (src/SphericalHarmonics.cc:1090-1131): B[k + Ncoef*a][grid*3+b]
        sym = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
        M1arr = np.zeros((Nv, 3 * Ncoef, 3, p0 + 1, 2 * p0))
        for a in range(3):
            for b in range(3):
                c = sym[a][b]
                # t=0 pole -> grid (j1, j2); t=1 pole -> grid (p0-j1, p0+j2)
                blk0 = K2[:, :, :, 0, c, :]                  # [Nv][j1][j2][k]
                blk1 = K2[:, :, :, 1, c, :]
                M1arr[:, a * Ncoef:(a + 1) * Ncoef, b, :, :p0] = np.transpose(blk0, (0, 3, 1, 2))
                M1arr[:, a * Ncoef:(a + 1) * Ncoef, b, ::-1, p0:] = np.transpose(blk1, (0, 3, 1, 2))
        G = grid2shc(M1arr.reshape(-1, Ngrid), p0, p0)          # [Nv*3Ncoef*3][Ncoef]
        results.append(G.reshape(Nv, 3 * Ncoef, 3 * Ncoef))
    return results[0], results[1]
