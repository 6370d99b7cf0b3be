"""Pins the oracle (the CPU restatement the GPU is checked against).

Each check is independent of the formula it tests: closed forms solved
another way, known answers, NumPy polynomial roots, brute force on tiny
inputs and an exact SES point classifier.
"""
import ctypes

import numpy as np
import pytest

import nanoshaper_amd as ns
from oracle import binding


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module")
def orc():
    return binding.load()


# --------------------------------------------------------------- ray/sphere
def test_ray_sphere_matches_geometry(orc):
    rng = np.random.default_rng(0)
    for _ in range(500):
        o = rng.normal(size=3) * 5
        d = np.zeros(3)
        d[rng.integers(3)] = rng.uniform(1, 30)
        c = rng.normal(size=3) * 3
        r = rng.uniform(0.5, 3)
        t = np.zeros(2)
        hit = orc.orc_ray_sphere(_ptr(o), _ptr(d), _ptr(c), r, _ptr(t[0:1]), _ptr(t[1:2]))
        # geometry: distance from centre to the line, then the chord half length
        u = d / np.linalg.norm(d)
        s0 = np.dot(c - o, u)
        h2 = r * r - np.dot(c - o - s0 * u, c - o - s0 * u)
        assert bool(hit) == (h2 >= 0)
        if hit:
            exp = (np.array([s0 - np.sqrt(h2), s0 + np.sqrt(h2)]) / np.linalg.norm(d))
            assert np.allclose(t, exp, atol=1e-12)


# --------------------------------------------------------------- quartic / cubic
def test_quartic_four_real_roots(orc):
    rng = np.random.default_rng(1)
    for _ in range(300):
        rts = np.sort(rng.uniform(-5, 5, 4))
        if np.min(np.diff(rts)) < 0.05:
            continue
        b, c, d, e = np.poly(rts)[1:]
        out = np.zeros(4)
        n = ctypes.c_int()
        orc.orc_quartic(_ptr(out), b, c, d, e, ctypes.byref(n))
        assert n.value == 4
        # the closed form is ill-conditioned near clustered roots (the
        # reference's own accuracy); a wrong term or sign is off by O(1)
        assert np.allclose(np.sort(out), rts, atol=1e-4)


def test_quartic_two_real_roots(orc):
    rng = np.random.default_rng(2)
    for _ in range(300):
        a1, a2 = np.sort(rng.uniform(-5, 5, 2))
        if a2 - a1 < 0.05:
            continue
        p = rng.uniform(-3, 3)
        q = p * p / 4 + rng.uniform(0.5, 4)  # x^2+px+q has no real root
        coef = np.polymul(np.poly([a1, a2]), [1, p, q])
        out = np.zeros(4)
        n = ctypes.c_int()
        orc.orc_quartic(_ptr(out), *coef[1:], ctypes.byref(n))
        assert n.value == 2
        assert np.allclose(np.sort(out[:2]), [a1, a2], atol=1e-4)


def test_quartic_no_real_root(orc):
    rng = np.random.default_rng(3)
    for _ in range(200):
        p1, p2 = rng.uniform(-3, 3, 2)
        coef = np.polymul([1, p1, p1 * p1 / 4 + rng.uniform(0.5, 3)], [1, p2, p2 * p2 / 4 + rng.uniform(0.5, 3)])
        out = np.zeros(4)
        n = ctypes.c_int()
        orc.orc_quartic(_ptr(out), *coef[1:], ctypes.byref(n))
        assert n.value == 0


def test_cubic_returns_a_root(orc):
    rng = np.random.default_rng(4)
    for _ in range(300):
        b, c, d = rng.uniform(-20, 20, 3)
        m = orc.orc_cubic(b, c, d)
        real = np.roots([1, b, c, d])
        real = real[np.abs(real.imag) < 1e-9].real
        assert np.min(np.abs(real - m)) < 1e-6 * max(1.0, abs(m))


# --------------------------------------------------------------- ray/torus
def _random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def test_ray_torus_matches_implicit_quartic(orc):
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(400):
        rot = _random_rotation(rng)  # columns: torus frame axes in world space
        invrot = np.ascontiguousarray(rot.T)
        center = rng.normal(size=3)
        R = rng.uniform(0.5, 3.0)
        r = 1.4
        panel = int(rng.integers(3))
        vc = (0, 2, 1)[panel]
        o = center + rng.normal(size=3) * 2
        o[vc] = center[vc] - 10
        d = np.zeros(3)
        d[vc] = 20.0
        sc = center.copy()
        srad = R + r + 0.5  # clipping sphere containing the whole torus
        t = np.zeros(4)
        t1 = np.zeros(1)
        t2 = np.zeros(1)
        if not orc.orc_ray_sphere(_ptr(o), _ptr(d), _ptr(sc), srad, _ptr(t1), _ptr(t2)):
            continue
        t[0] = t1[0]
        n = ctypes.c_int()
        orc.orc_ray_torus(_ptr(invrot.ravel()), _ptr(center), _ptr(sc), r, R, srad, panel, _ptr(o), _ptr(d),
                          _ptr(t), ctypes.byref(n))
        # independent: (|p|^2 + R^2 - r^2)^2 - 4 R^2 (px^2 + py^2) along the ray in the torus frame
        p0 = invrot @ (o - center)
        u = invrot @ (d / np.linalg.norm(d))
        a2, a1, a0 = np.dot(u, u), 2 * np.dot(p0, u), np.dot(p0, p0) + R * R - r * r
        qa = np.polymul([a2, a1, a0], [a2, a1, a0])
        qb = 4 * R * R * np.polyadd(np.polymul([u[0], p0[0]], [u[0], p0[0]]), np.polymul([u[1], p0[1]], [u[1], p0[1]]))
        roots = np.roots(np.polysub(qa, qb))
        real = np.sort(roots[np.abs(roots.imag) < 1e-7].real) / np.linalg.norm(d)
        got = np.sort(t[:n.value])
        # tangential double roots may be reported or not; compare when well separated
        if len(real) > 1 and np.min(np.diff(real)) * 20 < 1e-4:
            continue
        assert len(got) == len(real)
        assert np.allclose(got, real, atol=1e-9)
        checked += 1
    assert checked > 200


def test_sturm_quartic_roots(orc):
    # getRealRootsSturm: distinct real roots in ascending order to RELERROR
    r = np.zeros(4)
    n = orc.orc_sturm_quartic(_ptr(r), 0.0, -10.0, 0.0, 9.0)  # (x^2 - 1)(x^2 - 9)
    assert n == 4 and np.allclose(r, [-3, -1, 1, 3], atol=1e-6)
    assert orc.orc_sturm_quartic(_ptr(r), 0.0, 2.0, 0.0, 1.0) == 0  # (x^2 + 1)^2
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(300):
        real = np.sort(rng.uniform(-6, 6, size=int(rng.integers(0, 3)) * 2))
        cplx = rng.uniform(-3, 3, size=2 - len(real) // 2) + 1j * rng.uniform(0.5, 2, size=2 - len(real) // 2)
        allr = list(real) + [z for c in cplx for z in (c, np.conj(c))]
        poly = np.real(np.poly(allr))  # monic, highest first
        if len(real) > 1 and np.min(np.diff(real)) < 1e-3:
            continue
        n = orc.orc_sturm_quartic(_ptr(r), poly[1], poly[2], poly[3], poly[4])
        assert n == len(real)
        assert np.all(np.abs(r[:n] - real) <= 1e-6 * np.maximum(1, np.abs(real)))
        checked += 1
    assert checked > 250


def test_root_refinement_lowers_torus_residual(orc):
    # the extended-precision refinement (the mixed-precision yardstick) must
    # never raise the implicit-torus residual of a hit, must lower it on some,
    # and may move a root by at most 1e-3 A
    rng = np.random.default_rng(11)
    ld = np.longdouble
    lowered = compared = 0
    for _ in range(600):
        rot = _random_rotation(rng)
        invrot = np.ascontiguousarray(rot.T)
        center = rng.normal(size=3)
        R = rng.uniform(0.5, 3.0)
        r = 1.4
        panel = int(rng.integers(3))
        vc = (0, 2, 1)[panel]
        o = center + rng.normal(size=3) * 2
        o[vc] = center[vc] - 10
        d = np.zeros(3)
        d[vc] = 1.0
        sc = center.copy()
        srad = R + r + 0.5
        t1, t2 = np.zeros(1), np.zeros(1)
        if not orc.orc_ray_sphere(_ptr(o), _ptr(d), _ptr(sc), srad, _ptr(t1), _ptr(t2)):
            continue
        hits = []
        for refine in (0, 1):
            t = np.zeros(4)
            t[0] = t1[0]
            n = ctypes.c_int()
            orc.orc_set_root_refinement(refine)
            orc.orc_ray_torus(_ptr(invrot.ravel()), _ptr(center), _ptr(sc), r, R, srad, panel, _ptr(o), _ptr(d),
                              _ptr(t), ctypes.byref(n))
            hits.append(np.sort(t[:n.value]))
        orc.orc_set_root_refinement(0)
        if len(hits[0]) != len(hits[1]) or len(hits[0]) == 0:
            assert len(hits[0]) == len(hits[1])
            continue
        assert np.abs(hits[0] - hits[1]).max() <= 1e-3

        def residual(tv):
            p = np.array([ld(x) for x in o]) + np.array([ld(x) for x in d]) * ld(tv)
            q = np.array([[ld(v) for v in row] for row in invrot]) @ (p - np.array([ld(x) for x in center]))
            s2 = q @ q + ld(R) * ld(R) - ld(r) * ld(r)
            return abs(s2 * s2 - 4 * ld(R) * ld(R) * (q[0] * q[0] + q[1] * q[1]))

        for a, b in zip(hits[0], hits[1]):
            fa, fb = residual(a), residual(b)
            assert fb <= fa * (1 + 1e-9) + ld(1e-12)
            lowered += fb < fa
            compared += 1
    assert compared > 400 and lowered > compared // 10


def test_vertex_normals_follow_the_surface():
    # analytic normals of a sphere are radial; on a molecule they are unit
    # vectors that agree with the orientation of the triangles around them
    import nanoshaper_amd as ns
    from oracle.binding import OracleSurface
    a = np.array([[0.3, -0.2, 0.1, 1.7]])
    o = OracleSurface(ns.build_grid(a, 3.0, 80.0), ns.ses_patches(a, 1.4))
    o.get_surf()
    o.triangulate()
    v, _ = o.mesh()
    radial = (v - a[0, :3]) / 1.7
    assert np.abs(o.normals() - radial).max() < 1e-12
    x, _ = ns.generate_chain(40, 6.0, seed=0)
    j, _ = ns.jitter(x)
    o = OracleSurface(ns.build_grid(x, 2.0, 80.0), ns.ses_patches(j, 1.4))
    o.get_surf()
    o.triangulate()
    v, t = o.mesh()
    n = o.normals()
    assert np.abs(np.linalg.norm(n, axis=1) - 1).max() < 1e-12
    acc = np.zeros_like(v)
    w = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    for q in range(3):
        np.add.at(acc, t[:, q], w)
    acc /= np.linalg.norm(acc, axis=1)[:, None]
    assert np.median((n * acc).sum(1)) > 0.95


# --------------------------------------------------------------- generator / grid
def _park_miller_py(seed, n):
    s = seed
    for _ in range(n):
        hi, lo = divmod(s, 127773)
        s = 16807 * lo - 2836 * hi
        if s <= 0:
            s += 2147483647
    return s


def test_park_miller_known_answer_and_jitter():
    # Park & Miller (1988): seed 1 gives 1043618065 after 10,000 draws
    assert _park_miller_py(1, 10000) == 1043618065
    atoms = np.zeros((3333, 4))
    atoms[:, 3] = 1.0
    jit, state = ns.jitter(atoms, amplitude=1.0, seed=1)
    assert state == _park_miller_py(1, 9999)
    # first displacement: u1 - 0.5 with u1 = 16807/m
    assert jit[0, 0] == pytest.approx(16807 / 2147483647 - 0.5, abs=1e-15)


def test_grid_closed_form():
    rng = np.random.default_rng(6)
    atoms = np.column_stack([rng.uniform(-7, 9, (25, 3)), rng.uniform(1.2, 2.0, 25)])
    for scale, perfil in ((2.0, 80.0), (3.0, 70.0), (4.0, 90.0)):
        g = ns.build_grid(atoms, scale, perfil)
        lo = (atoms[:, :3] - atoms[:, 3:]).min(0)
        hi = (atoms[:, :3] + atoms[:, 3:]).max(0)
        mid = (lo + hi) / 2
        rmaxdim = 2 * np.max(np.abs(np.concatenate([hi - mid, lo - mid])))
        n = int(np.floor(scale * 100 / perfil * rmaxdim))
        n += (n % 2 == 0)
        assert (g.nx, g.ny, g.nz) == (n, n, n)
        assert g.xmin == pytest.approx(mid[0] - (n - 1) / (2 * scale), abs=1e-12)
        assert g.zmax == pytest.approx(mid[2] + (n - 1) / (2 * scale), abs=1e-12)
        assert g.hside == pytest.approx(0.5 / scale) and g.A == pytest.approx(1 / scale ** 2)


# --------------------------------------------------------------- marching cubes table
def test_mc_case_table_invariants():
    import re
    src = open(binding.HERE + "/../nanoshaper_amd/csrc/common/mc_cases.h").read()
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]{15})ull", src)]
    assert len(words) == 256
    corners = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    total = 0
    for c, w in enumerate(words):
        active = {e for e, (a, b) in enumerate(corners) if ((c >> a) & 1) != ((c >> b) & 1)}
        tris, used = [], set()
        for q in range(5):
            e = [(w >> (12 * q + 4 * s)) & 15 for s in range(3)]
            if e[0] == 15:
                break
            tris.append(e)
            used |= set(e)
        total += len(tris)
        assert used == active, c  # triangles use exactly the active edges
        # complementary cases have the same number of triangles except the 6
        # ambiguous-face pairs of the published table
        if c in (0, 255):
            assert not tris
    assert total == 820  # triangle count of the published Lorensen-Cline/Bourke table


# --------------------------------------------------------------- cavity numbering
def test_cavity_labels_follow_alternating_scan():
    from scipy import ndimage
    # atoms arranged as a hollow shell with several buried pockets
    xyzr, _ = ns.generate_chain(800, 16.0, seed=5)
    g = ns.build_grid(xyzr, 3.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, 1.4)
    o = binding.OracleSurface(g, ps)
    o.get_surf(fill_cavities=False)
    st0 = o.status()
    o.get_surf()
    st1 = o.status()
    T = st0 == ns.STATUS_TEMP_OUT
    lab, n = ndimage.label(T, structure=ndimage.generate_binary_structure(3, 1))
    assert n >= 3
    NZ, NY, NX = st0.shape
    flat = np.arange(st0.size).reshape(st0.shape)
    brank = (NZ - 1 - np.arange(NZ)[:, None, None]) * NY * NX + flat % (NY * NX)
    fkey = ndimage.minimum(flat, lab, np.arange(1, n + 1))
    bkey = ndimage.minimum(brank, lab, np.arange(1, n + 1))
    remaining = list(range(n))
    want = np.zeros(n, int)
    for t in range(n):
        key = fkey if t % 2 == 0 else bkey
        pick = min(remaining, key=lambda q: key[q])
        remaining.remove(pick)
        want[pick] = 3 + t
    exp = st0.copy()
    exp[T] = want[lab[T] - 1]
    assert np.array_equal(exp, st1)


# --------------------------------------------------------------- whole SES vs exact classifier
def _dist_to_free(points, centers, R):
    """Distance from each point to the probe-free region {q: |q-c_i| >= R_i}."""
    n = len(centers)
    best = np.full(len(points), np.inf)
    inside_any = np.zeros(len(points), bool)
    for i in range(n):
        inside_any |= np.linalg.norm(points - centers[i], axis=1) < R[i]
    best[~inside_any] = 0.0

    def free(q):
        ok = np.ones(len(q), bool)
        for i in range(n):
            ok &= np.linalg.norm(q - centers[i], axis=1) >= R[i] - 1e-12
        return ok

    cands = []
    for i in range(n):  # projections onto single spheres
        v = points - centers[i]
        nv = np.linalg.norm(v, axis=1, keepdims=True)
        cands.append(centers[i] + v / np.maximum(nv, 1e-30) * R[i])
    for i in range(n):  # closest points on circles of intersection
        for j in range(i + 1, n):
            d = centers[j] - centers[i]
            L = np.linalg.norm(d)
            if L >= R[i] + R[j] or L <= abs(R[i] - R[j]):
                continue
            w = d / L
            a = (R[i] ** 2 - R[j] ** 2 + L * L) / (2 * L)
            cc = centers[i] + a * w
            rc = np.sqrt(R[i] ** 2 - a * a)
            v = points - cc
            v = v - np.outer(v @ w, w)
            nv = np.linalg.norm(v, axis=1, keepdims=True)
            cands.append(cc + v / np.maximum(nv, 1e-30) * rc)
    for i in range(n):  # probe positions tangent to three atoms
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                for q in _three_sphere_points(centers[i], R[i], centers[j], R[j], centers[k], R[k]):
                    cands.append(np.repeat(q[None], len(points), 0))
    for q in cands:
        ok = free(q) & inside_any
        dq = np.linalg.norm(points - q, axis=1)
        best = np.where(ok & (dq < best), dq, best)
    return best


def _three_sphere_points(c1, r1, c2, r2, c3, r3):
    ex = c2 - c1
    d = np.linalg.norm(ex)
    ex /= d
    i = np.dot(ex, c3 - c1)
    ey = c3 - c1 - i * ex
    j = np.linalg.norm(ey)
    if j < 1e-9:
        return []
    ey /= j
    ez = np.cross(ex, ey)
    x = (r1 ** 2 - r2 ** 2 + d * d) / (2 * d)
    y = (r1 ** 2 - r3 ** 2 + i * i + j * j) / (2 * j) - i / j * x
    z2 = r1 ** 2 - x * x - y * y
    if z2 < 0:
        return []
    z = np.sqrt(z2)
    base = c1 + x * ex + y * ey
    return [base + z * ez, base - z * ez]


@pytest.mark.parametrize("seed,n", [(11, 3), (12, 5), (13, 7)])
def test_oracle_hits_match_exact_ses(seed, n):
    """Hit depths of the patch path equal the sign changes of the exact SES
    classifier (outside <=> within one probe radius of the probe-free region)."""
    probe = 1.4
    xyzr, _ = ns.generate_chain(n, 5.0, seed=seed)
    g = ns.build_grid(xyzr, 2.0, 60.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, probe)
    o = binding.OracleSurface(g, ps)
    r = o.get_surf(build_epsmap=True, build_status=True, accurate_triangulation=False, fill_cavities=False)
    counts, T = o.ray_hits(0)
    xs, ys, zs = ns.grid_coords(g)
    N = g.nx
    L = xs[-1] - xs[0]
    centers, R = jit[:, :3], jit[:, 3] + probe
    rng = np.random.default_rng(seed)
    rays = rng.choice(N * N, size=min(80, N * N), replace=False)
    s = np.linspace(0, 1, 4001)
    checked = 0
    for ray in rays:
        n_, m_ = divmod(int(ray), N)
        pts = np.column_stack([xs[0] + s * L, np.full_like(s, ys[m_]), np.full_like(s, zs[n_])])
        f = probe - _dist_to_free(pts, centers, R)  # >0 outside the SES
        sign = f > 0
        ch = np.nonzero(sign[1:] != sign[:-1])[0]
        exact = []
        for c in ch:
            a, b = s[c], s[c + 1]
            fa = f[c]
            for _ in range(50):
                mid = 0.5 * (a + b)
                pm = np.array([[xs[0] + mid * L, ys[m_], zs[n_]]])
                fm = probe - _dist_to_free(pm, centers, R)[0]
                if (fm > 0) == (fa > 0):
                    a, fa = mid, fm
                else:
                    b = mid
            exact.append(0.5 * (a + b))
        got = T[ray, :counts[ray]]
        assert len(got) == len(exact), (ray, got, exact)
        if exact:
            assert np.max(np.abs(np.array(got) - np.array(exact))) * L < 1e-6
            checked += 1
    assert r["odd_rays"] == 0
    assert checked > 0


# --------------------------------------------------------------- threaded oracle
@pytest.mark.parametrize("kind", ["ses", "skin"])
def test_threaded_oracle_equals_single_thread(kind):
    """The CPU baseline runs the oracle on all host cores with the reference's
    decomposition; every output equals the single-thread run's."""
    xyzr, _ = ns.generate_chain(300, 9.0, seed=4)
    g = ns.build_grid(xyzr, 3.0, 80.0)
    jit, _ = ns.jitter(xyzr)
    ps = ns.ses_patches(jit, 1.4) if kind == "ses" else ns.skin_patches(jit, 0.45)
    out = []
    for threads in (1, 7):
        binding.set_threads(threads)
        try:
            o = binding.OracleSurface(g, ps)
            r = o.get_surf()
            m = o.triangulate()
            v, t = o.mesh()
            hits = [o.ray_hits(s_) for s_ in range(6)]
            out.append((r, m, o.epsmap(), o.idebmap(), o.status(), o.corner_bits(), v, t, o.normals(), hits,
                        o.cavities()))
        finally:
            binding.set_threads(1)
    a, b = out
    assert a[0] == b[0] and a[1] == b[1]
    for x, y in zip(a[2:8], b[2:8]):
        assert np.array_equal(x, y)
    assert np.abs(a[8] - b[8]).max() < 1e-9  # equal depths of two patches may swap their normals
    for (ca, ta), (cb, tb) in zip(a[9], b[9]):
        assert np.array_equal(ca, cb) and np.array_equal(ta, tb)
    assert all(np.array_equal(p, q) for p, q in zip(a[10], b[10]))
