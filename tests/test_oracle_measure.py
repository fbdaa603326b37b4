"""The oracle's measure stage (validation, components, grouping) against
the reference's own known answers and against independent definitions.

Known answers restated from the reference's tests
(crates/sesmesh/src/mesh/tests.rs): the octahedron's area 4 sqrt(3), volume
4/3 and Euler characteristic 2; the inverted octahedron is a cavity; two
octahedra are two components numbered in vertex order; the defect counts of
`validation_finds_defects` and `duplicates_with_repeated_vertices`.
Independent definitions: on seeded random meshes with defects, edge runs
from sorting every directed edge (the test's own `reference_runs`),
duplicates by sorting vertex triples, components by scipy's connected
components, and grouping by a stable sort. The order-dependent f64 sums are
pinned by the reference's recorded areas and volumes
(tests/test_oracle_golden.py)."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle

OCT_P = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
OCT_T = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                 np.uint32)
NO_DEFECT = dict(boundary_edges=0, non_manifold_edges=0, misoriented_edges=0, degenerate_triangles=0,
                 duplicate_triangles=0, invalid_indices=0, bad_vertices=0)


def test_octahedron_measures():
    m = oracle.measure(OCT_P, OCT_T, OCT_P)
    assert m["validation"] == NO_DEFECT
    assert abs(m["area"] - 4 * 3 ** 0.5) < 1e-12 and abs(m["volume"] - 4 / 3) < 1e-12
    assert m["components"] == 1 and m["cavities"] == 0 and m["euler"] == 2
    assert list(m["vertex_component"]) == [0] * 6
    assert m["comp"]["euler"][0] == 2 and m["comp"]["vertices"][0] == 6 and m["comp"]["triangles"][0] == 8


def test_inverted_octahedron_is_a_cavity():
    m = oracle.measure(OCT_P, OCT_T[:, [0, 2, 1]], OCT_P)
    assert m["cavities"] == 1 and abs(m["comp"]["volume"][0] + 4 / 3) < 1e-12


def test_two_components_in_vertex_order():
    p = np.vstack([OCT_P, OCT_P + [5, 0, 0]])
    t = np.vstack([OCT_T, OCT_T + 6])
    m = oracle.measure(p, t, np.vstack([OCT_P, OCT_P]))
    assert m["components"] == 2 and m["vertex_component"][0] == 0 and m["vertex_component"][6] == 1
    assert m["validation"] == NO_DEFECT
    assert list(m["comp"]["first_triangle"]) == [0, 8]


def test_validation_finds_defects():
    v = lambda t, p=OCT_P, n=OCT_P: oracle.measure(p, t, n)["validation"]
    assert v(OCT_T[:-1])["boundary_edges"] == 3
    t = OCT_T.copy()
    t[0] = t[0, [0, 2, 1]]
    assert v(t)["misoriented_edges"] == 3
    r = v(np.vstack([OCT_T, OCT_T[:1]]))
    assert r["duplicate_triangles"] == 1 and r["non_manifold_edges"] == 3
    assert v(np.vstack([OCT_T, [[0, 0, 1]]]))["degenerate_triangles"] >= 1
    assert v(np.vstack([OCT_T, [[0, 1, 99]]]))["invalid_indices"] == 1
    p, n = OCT_P.copy(), OCT_P.copy()
    n[0] = 0
    p[1, 0] = np.nan
    assert v(OCT_T, p, n)["bad_vertices"] == 2


def test_duplicates_with_repeated_vertices():
    extra = [[0, 1, 1], [1, 0, 1], [1, 1, 0], [2, 2, 2], [2, 2, 2], [3, 4, 3], [3, 3, 4], [4, 3, 4]]
    assert oracle.measure(OCT_P, np.vstack([OCT_T, extra]), OCT_P)["validation"]["duplicate_triangles"] == 4


def defect_meshes(seed=0x5E5, count=120):
    """Octahedra with random extra vertices, random triangles and re-used
    (possibly flipped) triangles, as the reference's defect_meshes."""
    rng = np.random.default_rng(seed)
    out = []
    for case in range(count):
        n = 6 + int(rng.integers(0, 1 + case % 40))
        extra = rng.normal(size=(n - 6, 3))
        extra /= np.linalg.norm(extra, axis=1, keepdims=True) + 1e-30
        p = np.vstack([OCT_P, extra.astype(np.float32)])
        t = [list(x) for x in OCT_T]
        for _ in range(int(rng.integers(0, 1 + case))):
            t.append(list(rng.integers(0, n, 3)))
        for _ in range(int(rng.integers(0, 1 + case % 12))):
            a = t[int(rng.integers(len(t)))]
            r = int(rng.integers(3))
            c = [a[r], a[(r + 1) % 3], a[(r + 2) % 3]]
            if rng.integers(2):
                c[1], c[2] = c[2], c[1]
            t.append(c)
        out.append((p, np.array(t, np.uint32)))
    return out


def runs_by_sorting(t):
    a, b = t.reshape(-1), np.roll(t, -1, axis=1).reshape(-1)
    keys = sorted(zip(np.minimum(a, b).tolist(), np.maximum(a, b).tolist(), (a < b).tolist()))
    runs = {}
    for lo, hi, fwd in keys:
        r = runs.setdefault((lo, hi), [0, 0])
        r[0 if fwd else 1] += 1
    return runs


@pytest.mark.parametrize("seed", [0x5E5, 7])
def test_random_defect_meshes_match_independent_definitions(seed):
    dups = 0
    for p, t in defect_meshes(seed):
        m = oracle.measure(p, t, p)
        runs = runs_by_sorting(t)
        want = dict(boundary_edges=0, non_manifold_edges=0, misoriented_edges=0)
        for (lo, hi), (f, b) in runs.items():
            if lo == hi or (f, b) == (1, 1):
                continue
            want[{1: "boundary_edges", 2: "misoriented_edges"}.get(f + b, "non_manifold_edges")] += 1
        for k, x in want.items():
            assert m["validation"][k] == x, k
        s = np.sort(t, axis=1)
        assert m["validation"]["duplicate_triangles"] == len(s) - len(np.unique(s, axis=0))
        dups += m["validation"]["duplicate_triangles"] > 0
        # components: scipy's partition, numbered by smallest vertex
        n = len(p)
        g = coo_matrix((np.ones(2 * len(t)), (np.r_[t[:, 0], t[:, 0]], np.r_[t[:, 1], t[:, 2]])), shape=(n, n))
        _, lab = connected_components(g, directed=False)
        first = {}
        for v, l in enumerate(lab):
            first.setdefault(l, len(first))
        assert np.array_equal(m["vertex_component"], [first[l] for l in lab])
        vc = m["vertex_component"]
        comp_t = vc[t[:, 0]]
        assert np.array_equal(m["grouped"], t[np.argsort(comp_t, kind="stable")])
        edges = np.zeros(m["components"], np.int64)
        for (lo, hi) in runs:
            edges[vc[lo]] += 1
        c = m["comp"]
        assert np.array_equal(c["vertices"], np.bincount(vc, minlength=m["components"]))
        assert np.array_equal(c["triangles"], np.bincount(comp_t, minlength=m["components"]))
        assert np.array_equal(c["euler"], c["vertices"].astype(np.int64) - edges + c["triangles"].astype(np.int64))
        assert np.array_equal(c["first_triangle"], np.r_[0, np.cumsum(c["triangles"])[:-1]])
        # per-component sums equal a plain f64 sum to rounding
        o = p[0].astype(np.float64)
        q = p.astype(np.float64) - o
        a_, b_, d_ = q[t[:, 0]], q[t[:, 1]], q[t[:, 2]]
        area = 0.5 * np.linalg.norm(np.cross(b_ - a_, d_ - a_), axis=1)
        vol = np.einsum("ij,ij->i", a_, np.cross(b_, d_)) / 6
        for k in range(m["components"]):
            sel = comp_t == k
            assert abs(c["area"][k] - area[sel].sum()) <= 1e-12 * max(1, area[sel].sum())
            assert abs(c["volume"][k] - vol[sel].sum()) <= 1e-12 * max(1, abs(vol).sum())
    assert dups > 20


def test_surface_mesh_is_closed_and_grouped(structures):
    s = oracle.Surface(structures("1CRN"), 1.4, 0.5, 10, threads=4)
    m = s.measure_full()
    assert m["validation"] == NO_DEFECT
    assert m["components"] == len(m["comp"]["euler"]) and all(e % 2 == 0 and e <= 2 for e in m["comp"]["euler"])
    t = s.mesh()["triangles"]
    comp_t = m["vertex_component"][t[:, 0]]
    assert np.array_equal(m["grouped"], t[np.argsort(comp_t, kind="stable")])
    assert m["area"] == s.measure()["area"]
