"""The oracle against what the mathematics fixes, independent of the
reference's records: closed forms, a brute-force SES evaluator (the design
of the reference's sesmesh-testkit oracle: every sphere a blocker, every
pair a circle, every triple a probe position by trilateration), and mesh
invariants (closed, oriented, edge-manifold; vertices on the surface;
thread-count independence)."""
import itertools

import numpy as np
import pytest

import oracle

TOL = 1e-9


def brute_ses(x, centers, radii, probe):
    """SES value r_p - d(x, P) of the probe-accessible region P, by brute
    force over all spheres, pairs and triples (exact in f64 up to round-off)."""
    R = radii + probe
    n = len(centers)
    d0 = np.linalg.norm(centers - x, axis=1)
    if not np.any(d0 < R):
        return probe  # x itself is in P

    def free(q, skip):
        dd = np.linalg.norm(centers - q, axis=1)
        ok = dd >= R - TOL
        ok[list(skip)] = True
        return bool(np.all(ok))

    best = np.inf
    for i in range(n):
        v = x - centers[i]
        ln = np.linalg.norm(v)
        if ln < 1e-12:
            continue
        q = centers[i] + v * (R[i] / ln)
        if free(q, [i]):
            best = min(best, abs(ln - R[i]))
    for i, j in itertools.combinations(range(n), 2):
        d = np.linalg.norm(centers[j] - centers[i])
        if d >= R[i] + R[j] or d <= abs(R[i] - R[j]) or d == 0:
            continue
        axis = (centers[j] - centers[i]) / d
        t = (d * d + R[i] ** 2 - R[j] ** 2) / (2 * d)
        rho2 = R[i] ** 2 - t * t
        if rho2 <= 0:
            continue
        m = centers[i] + axis * t
        y = x - m
        w = y - axis * np.dot(y, axis)
        wl = np.linalg.norm(w)
        if wl < 1e-12:
            continue  # equidistant: a vertex or sphere candidate is as close
        q = m + w / wl * np.sqrt(rho2)
        if free(q, [i, j]):
            best = min(best, np.linalg.norm(x - q))
    for i, j, k in itertools.combinations(range(n), 3):
        for q in trilaterate(centers[i], centers[j], centers[k], R[i], R[j], R[k]):
            if free(q, [i, j, k]):
                best = min(best, np.linalg.norm(x - q))
    return probe - best


def trilaterate(p1, p2, p3, r1, r2, r3):
    ex = p2 - p1
    d = np.linalg.norm(ex)
    if d == 0:
        return []
    ex /= d
    i = np.dot(ex, p3 - p1)
    ey = p3 - p1 - i * ex
    ne = np.linalg.norm(ey)
    if ne < 1e-12:
        return []
    ey /= ne
    ez = np.cross(ex, ey)
    j = np.dot(ey, p3 - p1)
    x = (r1 * r1 - r2 * r2 + d * d) / (2 * d)
    y = (r1 * r1 - r3 * r3 + i * i + j * j) / (2 * j) - i * x / j
    z2 = r1 * r1 - x * x - y * y
    if z2 < 0:
        return []
    z = np.sqrt(z2)
    base = p1 + x * ex + y * ey
    return [base + z * ez, base - z * ez]


def random_chain(rng, n):
    c = [np.zeros(3)]
    while len(c) < n:
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c.append(c[rng.integers(len(c))] + u * rng.uniform(1.3, 1.6))
    return np.array(c), rng.uniform(1.2, 2.0, n)


def test_single_sphere_closed_form():
    c, r, p = np.array([0.3, -0.2, 0.1]), 1.7, 1.4
    s = oracle.Surface(np.array([[*c, r]]), p, 0.25, 10)
    rng = np.random.default_rng(0)
    for _ in range(200):
        x = c + rng.normal(size=3) * 2.0
        v, _, feat = s.sample(x)
        dist = np.linalg.norm(x - c)
        want = p - max(r + p - dist, 0.0) if dist < r + p else p
        assert abs(v - want) < 1e-12, (x, v, want)
    # Lattice samples: exact where |F| < band, else +-band with the right sign.
    keys = s.bricks()
    vals = s.brick_values(keys)
    side = 11
    h = 0.25
    for key, v in zip(keys, vals):
        idx = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"), -1)
        pts = (idx[..., ::-1].reshape(-1, 3) + key * 10) * h
        d = np.linalg.norm(pts - c, axis=1)
        exact = np.where(d < r + p, d - r, p)
        band = s.band
        near = np.abs(exact) < band * 0.99
        assert np.all(np.abs(v[near] - exact[near]) < 1e-4)
        assert np.all(np.sign(v[~near]) == np.sign(exact[~near]))


def test_point_field_matches_brute_force():
    rng = np.random.default_rng(7)
    checked = 0
    for case in range(16):
        n = int(rng.integers(2, 9))
        centers, radii = random_chain(rng, n)
        probe = float(rng.uniform(0.6, 1.8))
        s = oracle.Surface(np.column_stack([centers, radii]), probe, 0.5, 10, stage=1)
        lo, hi = centers.min(0) - 3, centers.max(0) + 3
        for _ in range(25):
            x = rng.uniform(lo, hi)
            v, closest, feat = s.sample(x)
            want = brute_ses(x, centers, radii, probe)
            assert abs(v - want) < 1e-9, (case, x, v, want, feat)
            checked += 1
    assert checked == 400


def mesh_edges_ok(tris):
    """Closed, oriented, edge-manifold: every directed edge once, with its
    reverse present."""
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    key = e[:, 0].astype(np.int64) * (1 << 32) + e[:, 1]
    rev = e[:, 1].astype(np.int64) * (1 << 32) + e[:, 0]
    return len(np.unique(key)) == len(key) and np.array_equal(np.sort(key), np.sort(rev))


@pytest.mark.parametrize("spacing", [0.5, 0.25])
def test_mesh_invariants_on_crambin(structures, spacing):
    s = oracle.Surface(structures("1CRN"), 1.4, spacing, 10, 4)
    m = s.mesh()
    assert mesh_edges_ok(m["triangles"])
    # Vertices on the exact surface (reference tests/surface.rs:46-49):
    # |F(v)| <= 1e-4 h + 2e-6 plus the f32 rounding of the stored position.
    rng = np.random.default_rng(1)
    idx = rng.choice(len(m["positions"]), 400, replace=False)
    for i in idx:
        v, closest, feat = s.sample(m["positions"][i].astype(np.float64))
        assert abs(v) <= 1e-4 * spacing + 2e-6 + 1e-5, (i, v)


def test_thread_count_does_not_change_the_mesh(structures):
    a = oracle.Surface(structures("1UBQ"), 1.4, 0.5, 10, 1).mesh()
    b = oracle.Surface(structures("1UBQ"), 1.4, 0.5, 10, 7).mesh()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_brick_size_changes_sampling_not_the_surface(structures):
    # Different brick sizes give different lattices of bricks but the same
    # samples, hence the same vertex count (reference tests/surface.rs:170-200).
    n = {b: oracle.Surface(structures("1CRN"), 1.4, 0.5, b, 4).n_vertices for b in (6, 8, 10, 12)}
    assert len(set(n.values())) == 1, n
