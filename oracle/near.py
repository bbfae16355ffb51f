"""FP64 numpy restatement of NearSingular<Real> (src/NearSingular.cc) without
a periodic box: near-target search, quadratic-patch projection, repulsion
force, SubtractDirect and the near-singular velocity interpolation, for
surface targets (trg_is_surf, the step's own surface points) and for targets
off the surfaces, with single- and double-layer densities.

TEST INFRASTRUCTURE ONLY (the checker of ves3d_amd/csrc/near.hip).

The reference finds near (target, vesicle) pairs with a Morton tree whose
leaves are r_near wide and a 27-neighbour scan (:797-918); the set it builds
is "every target t and every vesicle v other than t's own with a source point
closer than r_near, paired with v's closest point".  This restatement builds
the same set by brute force.  Sources are the upsampled surface with its two
poles first, point-major [nv][M+2][3] with M = 2 p_up (p_up + 1).
"""
import math

import numpy as np

INTERP_DEG = 8                                           # include/NearSingular.h
EPS = 2.0 ** -53                                         # :1592-1599 (halving loop)
EPS_SQRT = math.sqrt(EPS)                                # :1379-1387


def inter_points(i, deg=INTERP_DEG):
    """InterPoints (src/NearSingular.cc:1356-1360): 0, then 1 .. 2 evenly."""
    return 1.0 + (i - 1.0) / (deg - 2) if i else 0.0


def inter_poly(x, j, deg=INTERP_DEG):
    """InterPoly (:1362-1373): Lagrange basis polynomial j at x."""
    y = 1.0
    x0 = inter_points(j, deg)
    for k in range(deg):
        if j != k:
            y *= (x - inter_points(k, deg)) / (x0 - inter_points(k, deg))
    return y


def r_near_of(S, nv, p_up):
    """r_ves = max distance of a source point from its vesicle's centroid
    (centroid = mean of all M+2 points), r_near = r_ves * 2/sqrt(p_up)
    (:83, :118-160)."""
    M = 2 + 2 * p_up * (p_up + 1)
    P = S.reshape(nv, M, 3)
    c = P.sum(axis=1) * (1.0 / M)
    r2 = np.max(np.sum((P - c[:, None, :]) ** 2, axis=2))
    return math.sqrt(r2) * (2.0 / math.sqrt(p_up))


def near_pairs(S, T, nv, p_up, r_near, trg_per_ves):
    """(target, vesicle, closest point) for every target within r_near of a
    point of another vesicle (:797-918).  Ordered by vesicle, then target.
    trg_per_ves None: targets off the surfaces, which belong to no vesicle."""
    M = 2 + 2 * p_up * (p_up + 1)
    P = S.reshape(nv, M, 3)
    r2n = r_near * r_near
    tves = (np.arange(T.shape[0]) // trg_per_ves if trg_per_ves
            else np.full(T.shape[0], -1))
    out = []
    for v in range(nv):
        d2 = np.sum((T[:, None, :] - P[v][None, :, :]) ** 2, axis=2)
        k = np.argmin(d2, axis=1)
        ok = (d2[np.arange(T.shape[0]), k] < r2n) & (tves != v)
        for t in np.nonzero(ok)[0]:
            out.append((t, v, k[t]))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def patch_interp_matrix(p_up):
    """M_patch_interp (:1738-1755): trigonometric interpolation of the 2p_up
    samples of a latitude ring to 8 equally spaced angles."""
    K = 2 * p_up
    k0 = np.arange(8)[:, None, None]
    k1 = np.arange(K)[None, :, None]
    k2 = np.arange(p_up + 1)[None, None, :]
    c = np.where((k2 == 0) | (k2 == p_up), 1.0, 2.0)
    a0, a1 = 2 * np.pi * k2 * (k0 / 8.0), 2 * np.pi * k2 * (k1 / K)
    return np.sum(c * (np.sin(a0) * np.sin(a1) + np.cos(a0) * np.cos(a1)) / K, axis=2)


def pole_fix_matrix():
    """M = M0 pinv(M1) (:1846-1868): re-samples a 3x3 patch whose 8 outer
    values lie on a circle about the pole onto the square stencil."""
    M0, M1 = np.zeros((9, 9)), np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            x0, y0 = i - 1.0, j - 1.0
            x1, y1 = x0, y0
            if x1 and y1:
                x1, y1 = x1 / math.sqrt(2), y1 / math.sqrt(2)
            for m in range(3):
                for n in range(3):
                    M0[i * 3 + j, m * 3 + n] = x0 ** m * y0 ** n
                    M1[i * 3 + j, m * 3 + n] = x1 ** m * y1 ** n
    return M0 @ np.linalg.pinv(M1)


def patch_mesh(p_up, k, sv, Mint=None, Mfix=None):
    """QuadraticPatch::patch_mesh (:1734-1870): the 3x3 stencil [9][3] of a
    per-point field sv [M+2][3] around point k (poles k=0,1)."""
    K, nlat = 2 * p_up, p_up + 1
    out = np.empty((9, 3))
    if k < 2:
        Mint = patch_interp_matrix(p_up) if Mint is None else Mint
        ring = 2 + (0 if k == 0 else (nlat - 1) * K)
        val = Mint @ sv[ring:ring + K]
        order = (0, 7, 6, 1, None, 5, 2, 3, 4) if k == 0 else (0, 1, 2, 7, None, 3, 6, 5, 4)
        for s, o in enumerate(order):
            out[s] = sv[k] if o is None else val[o]
        Mfix = pole_fix_matrix() if Mfix is None else Mfix
        return Mfix @ out
    k0, k1 = (k - 2) // K, (k - 2) % K
    cols = ((k1 + K - 1) % K, k1, (k1 + 1) % K)
    for r, lat in enumerate((k0 - 1, k0, k0 + 1)):
        for c in range(3):
            if lat < 0:
                idx = 0
            elif lat > nlat - 1:
                idx = 1
            else:
                idx = 2 + cols[c] + lat * K
            out[r * 3 + c] = sv[idx]
    return out


class QuadraticPatch:
    """QuadraticPatch (:1560-1573, :1575-1587, :1872-1888): the biquadratic
    through a 3x3 stencil; x runs along a stencil row, y down the rows."""

    def __init__(self, mesh):
        x = np.asarray(mesh).reshape(9, -1)
        dof = x.shape[1]
        self.dof = dof
        tmp = np.empty((3, 3, dof))
        for j in range(3):
            tmp[j, 0] = x[j * 3 + 1]
            tmp[j, 1] = (x[j * 3 + 2] - x[j * 3 + 0]) * 0.5
            tmp[j, 2] = (x[j * 3 + 2] + x[j * 3 + 0]) * 0.5 - x[j * 3 + 1]
        # c[a][b]: a = power of y (row direction), b = power of x (column)
        c = np.empty((3, 3, dof))
        c[0] = tmp[1]
        c[1] = (tmp[2] - tmp[0]) * 0.5
        c[2] = (tmp[2] + tmp[0]) * 0.5 - tmp[1]
        self.c = c

    def eval(self, x, y):
        xs, ys = (1.0, x, x * x), (1.0, y, y * y)
        v = np.zeros(self.dof)
        for i in range(3):
            for j in range(3):
                v += self.c[j, i] * xs[i] * ys[j]
        return v

    def grad(self, x, y):
        xs, ys = (1.0, x, x * x), (1.0, y, y * y)
        dx, dy = (0.0, 1.0, 2 * x), (0.0, 1.0, 2 * y)
        gx, gy = np.zeros(self.dof), np.zeros(self.dof)
        for i in range(3):
            for j in range(3):
                gx += self.c[j, i] * dx[i] * ys[j]
                gy += self.c[j, i] * xs[i] * dy[j]
        return gx, gy

    def project(self, t):
        """Nearest point on the patch by damped Gauss-Newton (:1589-1712).
        Returns (x, y, exterior)."""
        x = y = 0.0
        sc = self.eval(x, y)
        while True:
            dR = t - sc
            dR2 = float(dR @ dR)
            gx, gy = self.grad(x, y)
            n = np.cross(gx, gy)
            n = n / math.sqrt(float(n @ n))
            dRt = dR - float(dR @ n) * n
            if math.sqrt(float(dRt @ dRt)) < 1e-8:
                break
            dxdx, dydy, dxdy = float(gx @ gx), float(gy @ gy), float(gx @ gy)
            dxdR, dydR = float(gx @ dRt), float(gy @ dRt)
            det = dxdx * dydy - dxdy * dxdy
            dx = (dydy * dxdR - dxdy * dydR) / det
            dy = (dxdx * dydR - dxdy * dxdR) / det
            if (x <= -1.2 and dx < 0) or (x >= 1.2 and dx > 0) or \
               (y <= -1.2 and dy < 0) or (y >= 1.2 and dy > 0):
                break
            if x > -1.2 and x + dx < -1.2:
                s = (-1.2 - x) / dx; dx *= s; dy *= s
            if x < 1.2 and x + dx > 1.2:
                s = (1.2 - x) / dx; dx *= s; dy *= s
            if y > -1.2 and y + dy < -1.2:
                s = (-1.2 - y) / dy; dx *= s; dy *= s
            if y < 1.2 and y + dy > 1.2:
                s = (1.2 - y) / dy; dx *= s; dy *= s
            while True:
                r1 = t - self.eval(x + dx * 0.5, y + dy * 0.5)
                r2 = t - self.eval(x + dx, y + dy)
                d1, d2 = float(r1 @ r1), float(r2 @ r2)
                c1 = -d2 + d1 * 4.0 - dR2 * 3.0
                c2 = d2 * 2.0 - d1 * 4.0 + dR2 * 2.0
                if c2 > 0:
                    s = -0.5 * c1 / c2
                    if 0.0 < s < 1.0:
                        dx *= s; dy *= s
                sc = self.eval(x + dx, y + dy)
                r = t - sc
                dR2n = float(r @ r)
                if dR2n < dR2:
                    x += dx; y += dy
                    break
                if dx * dx * dxdx + dy * dy * dydy < 64 * EPS:
                    break
                dx *= 0.5; dy *= 0.5
            if dx * dx * dxdx + dy * dy * dydy < 64 * EPS:
                break
        dR = t - sc
        gx, gy = self.grad(x, y)
        direc = (dR[0] * gx[1] * gy[2] + dR[1] * gx[2] * gy[0] + dR[2] * gx[0] * gy[1]
                 - dR[0] * gx[2] * gy[1] - dR[1] * gx[0] * gy[2] - dR[2] * gx[1] * gy[0])
        return x, y, direc > 0


def stokes_sl(src, q, trg):
    """SL kernel 1/(8 pi) sum q/r + (q.d) d/r^3 (src/PVFMMInterface.cc:161-238)."""
    d = trg[:, None, :] - src[None, :, :]
    r2 = np.sum(d * d, axis=2)
    ri = np.where(r2 > 0, 1.0 / np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0)
    qd = np.sum(q[None, :, :] * d, axis=2)
    v = q[None, :, :] * ri[:, :, None] + (qd * ri ** 3)[:, :, None] * d
    return v.sum(axis=1) / (8 * np.pi)


def stokes_dl(src, qd, trg):
    """DL kernel 3/(4 pi) sum (d.n)(d.g) d / r^5, d = trg - src, qd = [g (3),
    n (3)] per source (src/PVFMMInterface.cc:284-369)."""
    d = trg[:, None, :] - src[None, :, :]
    r2 = np.sum(d * d, axis=2)
    ri = np.where(r2 > 0, 1.0 / np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0)
    dn = np.sum(qd[None, :, 3:] * d, axis=2)
    dg = np.sum(qd[None, :, :3] * d, axis=2)
    v = (dn * dg * ri ** 5)[:, :, None] * d
    return v.sum(axis=1) * (3.0 / (4 * np.pi))


class NearSingular:
    """NearSingular<Real> (src/StokesVelocity.cc:240-242, :146-150):
    coordinate setup (:80-1120), ForceRepul (:1122-1126), SubtractDirect
    (:1143-1212) and operator() (:1375-1551), for surface targets
    (trg_is_surf, trg_per_ves targets per vesicle) or targets off the
    surfaces (trg_is_surf False, trg_per_ves None)."""

    def __init__(self, S, T, nv, p_up, trg_per_ves, repul_dist=5e-2, r_near=None,
                 trg_is_surf=True):
        self.trg_is_surf = trg_is_surf
        self.nv, self.pu, self.ntrg = nv, p_up, T.shape[0]
        self.M = 2 + 2 * p_up * (p_up + 1)
        self.S = np.ascontiguousarray(S, dtype=np.float64).reshape(nv * self.M, 3)
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        self.r_near = r_near_of(self.S, nv, p_up) if r_near is None else r_near
        self.pairs = near_pairs(self.S, self.T, nv, p_up, self.r_near, trg_per_ves)
        self.Mint, self.Mfix = patch_interp_matrix(p_up), pole_fix_matrix()
        npair = len(self.pairs)
        self.xy = np.zeros((npair, 2))
        self.extr = np.zeros(npair, dtype=bool)
        self.proj = np.zeros((npair, 3))
        self.repl = np.zeros((npair, 3))
        self.min_dist = 1e10
        r2inv = (1.0 / repul_dist) ** 2 if repul_dist > 0 else 0.0
        for i, (t, v, k) in enumerate(self.pairs):
            Sv = self.S[v * self.M:(v + 1) * self.M]
            patch = QuadraticPatch(patch_mesh(p_up, k, Sv, self.Mint, self.Mfix))
            x, y, ext = patch.project(self.T[t])
            self.xy[i] = (x, y)
            self.extr[i] = ext
            self.proj[i] = patch.eval(x, y)
            gx, gy = patch.grad(x, y)
            n = np.cross(gx, gy)
            n = n / math.sqrt(float(n @ n))
            r = float((self.T[t] - self.proj[i]) @ n)
            self.min_dist = min(self.min_dist, r)
            f = 0.0
            if r2inv > 0 and r < self.r_near:
                f = math.exp(-r * r * r2inv) / r * (1.0 - r / self.r_near)
            self.repl[i] = n * f
        # interpolation points and Lagrange weights (:1426-1456, :1508-1540)
        self.ipts = np.zeros((npair, INTERP_DEG - 1, 3))
        self.w = np.zeros((npair, INTERP_DEG))
        for i, (t, v, k) in enumerate(self.pairs):
            dR = self.T[t] - self.proj[i]
            dn = math.sqrt(float(dR @ dR))
            if not self.extr[i] and self.trg_is_surf:  # always the exterior normal
                dn = -dn
            oo = 1.0 / dn if dn != 0 else 0.0
            if abs(dn) < EPS_SQRT:
                dn, oo = 0.0, 0.0
            ix = dn / self.r_near
            for l in range(INTERP_DEG - 1):
                self.ipts[i, l] = self.proj[i] + dR * oo * self.r_near * inter_points(l + 1)
            if ix == 0:
                self.w[i, 0] = 1.0
            else:
                self.w[i] = [inter_poly(ix, l) for l in range(INTERP_DEG)]

    def scatter(self, per_pair):
        out = np.zeros((self.ntrg, 3))
        np.add.at(out, self.pairs[:, 0], per_pair)
        return out

    def force_repul(self):
        """Per-target repulsion force [ntrg][3] (point-major)."""
        return self.scatter(self.repl)

    def subtract_direct(self, q, qd=None):
        """Direct SL (q, [M][3]) and DL (qd, [M][6], or None) sums of vesicle
        v at each of its near targets, scattered to the targets (the part
        SubtractDirect removes from the far field)."""
        out = np.zeros((len(self.pairs), 3))
        for v in np.unique(self.pairs[:, 1]) if len(self.pairs) else []:
            sel = np.nonzero(self.pairs[:, 1] == v)[0]
            sl = slice(v * self.M, (v + 1) * self.M)
            T = self.T[self.pairs[sel, 0]]
            if q is not None:
                out[sel] = stokes_sl(self.S[sl], q[sl], T)
            if qd is not None:
                out[sel] += stokes_dl(self.S[sl], qd[sl], T)
        return self.scatter(out)

    def interp(self, q, s_vel_up, qd=None, u_dl=None):
        """operator(): near-singular velocity at each near target from v's
        surface velocity at the projection and the direct sum at 7 points
        along the normal, scattered to the targets.  With a double-layer
        density (qd [M][6] packed as [g w, n], u_dl the density itself on
        the upsampled grid with the poles first, point-major) the direct
        sums carry its DL potential and the surface value its jump,
        +1/2 u_dl (-1/2 at an interior target off the surfaces, :1458-1469)."""
        npair = len(self.pairs)
        out = np.zeros((npair, 3))
        for v in np.unique(self.pairs[:, 1]) if npair else []:
            sel = np.nonzero(self.pairs[:, 1] == v)[0]
            sl = slice(v * self.M, (v + 1) * self.M)
            P = self.ipts[sel].reshape(-1, 3)
            iv = np.zeros((P.shape[0], 3))
            if q is not None:
                iv += stokes_sl(self.S[sl], q[sl], P)
            if qd is not None:
                iv += stokes_dl(self.S[sl], qd[sl], P)
            iv = iv.reshape(len(sel), INTERP_DEG - 1, 3)
            for a, i in enumerate(sel):
                k = self.pairs[i, 2]
                mesh = patch_mesh(self.pu, k, s_vel_up[sl], self.Mint, self.Mfix)
                if u_dl is not None:
                    scal = -0.5 if (not self.extr[i] and not self.trg_is_surf) else 0.5
                    mesh = mesh + scal * patch_mesh(self.pu, k, u_dl[sl], self.Mint, self.Mfix)
                patch = QuadraticPatch(mesh)
                pv = patch.eval(*self.xy[i])
                if self.w[i, 0] == 1.0 and not self.w[i, 1:].any():
                    out[i] = pv
                else:
                    out[i] = self.w[i, 0] * pv + self.w[i, 1:] @ iv[a]
        return self.scatter(out)
