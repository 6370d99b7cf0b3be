"""Exact point classifiers for the surfaces, from their definitions.

Used by the tests to pin the oracle; they share no code or formula with it.

* SES: a point is outside the solvent-excluded surface iff it lies within one
  probe radius p of the probe-free region F = {q : |q - c_i| >= r_i + p for
  every atom i}.  The closest point of F to x lies on a sphere of radius
  r_i + p, on a circle where two such spheres meet, or at a point where three
  meet; the distance is the smallest over the free candidates of those three
  kinds.
* Skin (Edelsbrunner's skin of the weighted points (c_i, w_i = r_i^2 / s)
  with shrink factor s): x is inside iff some convex combination lambda of the
  weighted points gives a ball containing x after shrinking, i.e.
  min over the simplex of G(lambda) = |x - z(lambda)|^2 - s w(lambda) <= 0 with
  z = sum lambda_i c_i and w = sum lambda_i (w_i - |c_i|^2) + |z|^2 (the
  power function of the combination is the combination of the atoms' power
  functions, |x - c_i|^2 - w_i).  G is a convex quadratic in lambda (s < 1);
  for fixed z the rest is linear in lambda,
  so a minimiser exists with at most four non-zero weights (a vertex of
  {lambda >= 0, sum = 1, Z lambda = z}).  Enumerating the supports of size
  1..4 and solving each equality-constrained problem gives the exact minimum.
"""
import itertools

import numpy as np


def _free(q, centers, R, tol=1e-12):
    """q (..., 3): True where q is outside every inflated sphere."""
    d = np.linalg.norm(q[..., None, :] - centers, axis=-1)
    return np.all(d >= R - tol, axis=-1)


def triple_points(c1, r1, c2, r2, c3, r3):
    """The (up to two) points at distance r1, r2, r3 from c1, c2, c3."""
    ex = c2 - c1
    d = np.linalg.norm(ex)
    ex = ex / d
    i = np.dot(ex, c3 - c1)
    ey = c3 - c1 - i * ex
    j = np.linalg.norm(ey)
    if j < 1e-9:
        return []
    ey = ey / j
    ez = np.cross(ex, ey)
    x = (r1 ** 2 - r2 ** 2 + d * d) / (2 * d)
    y = (r1 ** 2 - r3 ** 2 + i * i + j * j) / (2 * j) - i / j * x
    z2 = r1 ** 2 - x * x - y * y
    if z2 < 0:
        return []
    z = np.sqrt(z2)
    base = c1 + x * ex + y * ey
    return [base + z * ez, base - z * ez]


class ExactSES:
    """Distance to the probe-free region for the atoms (centres, radii)."""

    def __init__(self, centers, radii, probe):
        self.c = np.asarray(centers, float)
        self.R = np.asarray(radii, float) + probe
        self.p = probe
        n = len(self.c)
        # circles of intersecting inflated spheres
        circ = []
        for i in range(n):
            for j in range(i + 1, n):
                dv = self.c[j] - self.c[i]
                L = np.linalg.norm(dv)
                if L >= self.R[i] + self.R[j] or L <= abs(self.R[i] - self.R[j]):
                    continue
                w = dv / L
                a = (self.R[i] ** 2 - self.R[j] ** 2 + L * L) / (2 * L)
                circ.append((self.c[i] + a * w, np.sqrt(self.R[i] ** 2 - a * a), w, i, j))
        self.circles = circ
        # free triple points (fixed points, checked once)
        pts = []
        nb = [set() for _ in range(n)]
        for (_, _, _, i, j) in circ:
            nb[i].add(j)
            nb[j].add(i)
        for i in range(n):
            for j in sorted(nb[i]):
                if j <= i:
                    continue
                for k in sorted(nb[i] & nb[j]):
                    if k <= j:
                        continue
                    pts += triple_points(self.c[i], self.R[i], self.c[j], self.R[j], self.c[k], self.R[k])
        pts = np.array(pts).reshape(-1, 3)
        self.triples = pts[_free(pts, self.c, self.R)] if len(pts) else pts

    def dist_to_free(self, points, chunk=4096):
        pts = np.asarray(points, float).reshape(-1, 3)
        out = np.empty(len(pts))
        for a in range(0, len(pts), chunk):
            out[a:a + chunk] = self._dist(pts[a:a + chunk])
        return out

    def _dist(self, P):
        c, R = self.c, self.R
        v = P[:, None, :] - c[None]                      # (m, n, 3)
        dc = np.linalg.norm(v, axis=-1)                  # (m, n)
        best = np.where(np.any(dc < R, axis=1), np.inf, 0.0)
        todo = best > 0
        if not todo.any():
            return best
        Pm, vm, dm = P[todo], v[todo], dc[todo]
        b = np.full(len(Pm), np.inf)
        # projections onto the spheres
        q = c[None] + vm / np.maximum(dm, 1e-30)[..., None] * R[None, :, None]
        ok = _free(q, c, R)
        b = np.minimum(b, np.where(ok, np.abs(dm - R[None]), np.inf).min(axis=1))
        # closest points of the circles
        for (cc, rc, w, _, _) in self.circles:
            u = Pm - cc
            u = u - np.outer(u @ w, w)
            nu = np.linalg.norm(u, axis=1, keepdims=True)
            q = cc + u / np.maximum(nu, 1e-30) * rc
            ok = _free(q, c, R)
            b = np.where(ok, np.minimum(b, np.linalg.norm(Pm - q, axis=1)), b)
        if len(self.triples):
            d = np.linalg.norm(Pm[:, None, :] - self.triples[None], axis=-1).min(axis=1)
            b = np.minimum(b, d)
        best[todo] = b
        return best

    def signed(self, points):
        """p - dist: > 0 outside the SES, < 0 inside."""
        return self.p - self.dist_to_free(points)


def ses_near_ray(centers, radii, probe, axis, u, v):
    """ExactSES restricted to the atoms that can matter on the axis-aligned
    line through (u, v) (the two coordinates other than `axis`): a probe
    within p of a point of the line is blocked only by atoms within
    r + 2p of that line, so the restriction changes no distance <= p."""
    c = np.asarray(centers, float)
    o = [d for d in range(3) if d != axis]
    d = np.hypot(c[:, o[0]] - u, c[:, o[1]] - v)
    keep = d < np.asarray(radii) + 2 * probe + 1e-6
    return ExactSES(c[keep], np.asarray(radii)[keep], probe)


class ExactSkin:
    """min over the simplex of G(lambda) for the atoms and shrink factor s."""

    def __init__(self, centers, radii, s):
        self.c = np.asarray(centers, float)
        self.w = np.asarray(radii, float) ** 2 / s
        self.s = s
        n = len(self.c)
        self.sub = []
        for k in range(1, min(4, n) + 1):
            for S in itertools.combinations(range(n), k):
                S = list(S)
                Z = self.c[S].T                           # (3, k)
                H = 2 * (1 - s) * Z.T @ Z
                K = np.zeros((k + 1, k + 1))
                K[:k, :k] = H
                K[:k, k] = 1
                K[k, :k] = 1
                if np.linalg.matrix_rank(K) < k + 1:
                    continue  # affinely dependent centres: covered by smaller supports
                Ki = np.linalg.inv(K)
                l = s * (self.w[S] - (self.c[S] ** 2).sum(axis=1))
                # lambda(x) = A x + b0 from K [lambda; mu] = [2 Z^T x + l; 1]
                A = Ki[:k, :k] @ (2 * Z.T)
                b0 = Ki[:k, :k] @ l + Ki[:k, k]
                self.sub.append((S, A, b0, l))

    def gmin(self, points):
        P = np.asarray(points, float).reshape(-1, 3)
        best = np.full(len(P), np.inf)
        for S, A, b0, l in self.sub:
            lam = P @ A.T + b0                           # (m, k)
            ok = np.all(lam >= -1e-12, axis=1)
            if not ok.any():
                continue
            z = lam @ self.c[S]                          # (m, 3)
            g = ((P - z) ** 2).sum(axis=1) - self.s * ((z ** 2).sum(axis=1)) - lam @ l
            best = np.where(ok, np.minimum(best, g), best)
        return best
