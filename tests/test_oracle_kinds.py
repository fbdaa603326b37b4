"""The oracle's SAS and vdW kinds (unions of spheres) pinned by closed forms.

The reference tests its sphere kinds on one sphere against the analytic
area and volume, the surface distance |F(v)| of every vertex and the angle
of every normal (crates/sesmesh/tests/surface.rs:56-84). The same checks
run here, plus closed forms the reference does not state:

* the union of two spheres: area 4 pi (R1^2 + R2^2) less the two caps each
  sphere hides of the other, volume less the lens;
* F of a union of spheres is min_k(|x - c_k| - R_k): every lattice sample
  the sphere kernel evaluates is that minimum (clamped at band) to f32
  rounding, every sample a uniform sub-block assigns has its sign, and
  every vertex's atom is a sphere its point lies on;
* SAS(radii r, probe p) is vdW(radii r + p), bit for bit: the expanded
  radii are the same f64 values, and nothing else of the arrangement
  enters these kinds' fields (a slip to the van der Waals radius in the SAS
  path breaks it);
* an SES of probe 0 is the vdW surface (reference field/point.rs:224-228).
"""
import math

import numpy as np
import pytest

import oracle
from conftest import cpu_threads


def rel(a, b):
    return abs(a - b) / abs(b)


def sphere_F(points, xyzr):
    """min_k(|x - c_k| - R_k) in f64 and its argmin."""
    d = np.linalg.norm(points[:, None, :] - xyzr[None, :, :3], axis=2) - xyzr[None, :, 3]
    return d.min(axis=1), d


def expanded(xyzr, kind, probe):
    e = xyzr.copy()
    if kind == "sas":
        e[:, 3] += probe
    return e


def check_surface_accuracy(o, xyzr_expanded, h):
    m = o.mesh()
    pos = m["positions"].astype(np.float64)
    F, d = sphere_F(pos, xyzr_expanded)
    # tests/surface.rs:46-49: the edge clamp, the refinement tolerance and
    # f32 storage bound |F(v)|.
    assert np.abs(F).max() < 1e-4 * h + 2e-6
    # Normals: the unit radial direction from the nearest sphere's centre.
    k = d.argmin(axis=1)
    radial = pos - xyzr_expanded[k, :3]
    radial /= np.linalg.norm(radial, axis=1)[:, None]
    n = m["normals"].astype(np.float64)
    dot = np.clip((radial * n).sum(axis=1), -1.0, 1.0)
    cross = np.linalg.norm(np.cross(radial, n), axis=1)
    angle = np.degrees(np.arctan2(cross, dot))
    return m, F, d, angle


@pytest.mark.parametrize("kind,R", [("ses", 1.8), ("sas", 3.2), ("vdw", 1.8)])
def test_single_sphere_all_kinds(kind, R):
    """crates/sesmesh/tests/surface.rs:56-84 on the oracle."""
    s = np.array([[0.3, -0.2, 0.1, 1.8]])
    h = 0.25
    o = oracle.Surface(s, 1.4, h, 10, cpu_threads(), kind=kind)
    st = o.measure_full()
    assert st["components"] == 1 and st["euler"] == 2
    area, volume = 4.0 * math.pi * R * R, 4.0 / 3.0 * math.pi * R ** 3
    assert rel(st["area"], area) < 5e-3
    assert rel(st["volume"], volume) < 1.2e-2
    # An inscribed mesh lies inside the sphere: both fall short.
    assert st["area"] < area and st["volume"] < volume
    e = s.copy()
    e[:, 3] = R
    m, F, d, angle = check_surface_accuracy(o, e, h)
    assert angle.max() < 0.01
    assert set(np.unique(m["patches"])) == {0}
    assert set(np.unique(m["atoms"])) == {0}


def union_of_two(r1, r2, dist):
    """Area and volume of the union of two overlapping balls."""
    a = (dist * dist + r1 * r1 - r2 * r2) / (2.0 * dist)  # plane of the circle, from centre 1
    h1, h2 = r1 - a, r2 - (dist - a)  # heights of the hidden caps
    area = 4.0 * math.pi * (r1 * r1 + r2 * r2) - 2.0 * math.pi * (r1 * h1 + r2 * h2)
    lens = math.pi * h1 * h1 * (3.0 * r1 - h1) / 3.0 + math.pi * h2 * h2 * (3.0 * r2 - h2) / 3.0
    volume = 4.0 / 3.0 * math.pi * (r1 ** 3 + r2 ** 3) - lens
    return area, volume


@pytest.mark.parametrize("kind", ["sas", "vdw"])
@pytest.mark.parametrize("r1,r2,dist", [(1.7, 1.5, 2.5), (1.6, 1.6, 3.0), (1.8, 1.2, 2.2)])
def test_two_spheres_union(kind, r1, r2, dist):
    s = np.array([[0.0, 0.0, 0.0, r1], [dist, 0.0, 0.0, r2]])
    probe = 1.4
    e = expanded(s, kind, probe)
    area, volume = union_of_two(e[0, 3], e[1, 3], dist)
    err = {}
    for h in (0.25, 0.125):
        o = oracle.Surface(s, probe, h, 10, cpu_threads(), kind=kind)
        st = o.measure_full()
        assert st["components"] == 1 and st["euler"] == 2
        err[h] = ((area - st["area"]) / area, (volume - st["volume"]) / volume)
    # Inscribed: short of both. The crease where the spheres meet is cut by
    # the cells next to it, so the area converges more slowly than on a
    # smooth surface; halving the spacing at least halves both errors.
    ea, ev = err[0.25]
    assert 0 < ea < 2e-2 and 0 < ev < 1.5e-2
    assert 0 < err[0.125][0] < 0.5 * ea and 0 < err[0.125][1] < 0.5 * ev
    h = 0.125
    m, F, d, angle = check_surface_accuracy(o, e, h)
    # Away from the crease the normal is the sphere's radial direction.
    sorted_d = np.sort(d, axis=1)
    away = sorted_d[:, 1] - sorted_d[:, 0] > 0.5
    assert angle[away].max() < 0.01
    # The atom is a sphere the vertex lies on (the crease: either).
    on = np.abs(d - F[:, None]) < 1e-3
    assert on[np.arange(len(F)), m["atoms"]].all()


def test_separated_spheres_are_two_components():
    s = np.array([[0.0, 0.0, 0.0, 1.8], [5.0, 0.0, 0.0, 1.2]])
    o = oracle.Surface(s, 1.4, 0.25, 10, cpu_threads(), kind="vdw")
    st = o.measure_full()
    assert st["components"] == 2 and st["euler"] == 4 and st["cavities"] == 0
    o = oracle.Surface(s, 1.4, 0.25, 10, cpu_threads(), kind="sas")  # 3.2 + 2.6 > 5: one
    assert o.measure_full()["components"] == 1


@pytest.mark.parametrize("kind", ["sas", "vdw"])
def test_samples_are_the_union_distance(structures, kind):
    """Every brick sample: the sphere kernel's min(band, min(|x-c| - R)) to
    f32 rounding where a sub-block is sampled, +-band of the right sign
    where uniform sub-blocks assign it (kernel.rs:489-546, 953-981)."""
    xyzr = structures("1CRN")
    probe, h, b = 1.4, 0.5, 10
    o = oracle.Surface(xyzr, probe, h, b, cpu_threads(), stage=2, kind=kind)
    keys = o.bricks()[::7]
    vals = o.brick_values(keys).astype(np.float64)
    e = expanded(xyzr, kind, probe)
    side = b + 1
    ii = np.arange(side ** 3)
    local = np.stack([ii % side, (ii // side) % side, ii // (side * side)], axis=1)
    fband = float(np.float32(o.band))
    n_exact = 0
    for key, v in zip(keys, vals):
        pts = (key.astype(np.int64) * b + local) * h
        F, _ = sphere_F(pts, e)
        at_band = np.abs(v) == fband
        # Evaluated: the clamped union distance, to f32 rounding.
        assert np.all(np.abs(v[~at_band] - np.minimum(F[~at_band], o.band)) < 2e-5)
        # +-band (clamped, or assigned by a uniform sub-block): the sign of F.
        assert np.all((v[at_band] > 0) == (F[at_band] > 0))
        n_exact += int((~at_band).sum())
    assert n_exact > 1000


def test_sas_is_vdw_of_expanded_radii(structures):
    xyzr = structures("1CRN")
    probe, h = 1.4, 0.5
    a = oracle.Surface(xyzr, probe, h, 10, cpu_threads(), kind="sas")
    b = oracle.Surface(expanded(xyzr, "sas", probe), probe, h, 10, cpu_threads(), kind="vdw")
    assert a.edge_clamp == b.edge_clamp and a.touched == b.touched
    assert np.array_equal(a.bricks(), b.bricks())
    ma, mb = a.mesh(), b.mesh()
    for k in ("positions", "normals", "t"):
        assert np.array_equal(ma[k].view(np.uint8), mb[k].view(np.uint8)), k
    for k in ("atoms", "patches", "triangles", "evals"):
        assert np.array_equal(ma[k], mb[k]), k
    assert a.n_vertices > 1000


def test_ses_with_probe_zero_is_vdw():
    s = np.array([[0.0, 0.0, 0.0, 1.6], [2.2, 0.4, 0.0, 1.5], [1.0, 2.0, 0.3, 1.7]])
    a = oracle.Surface(s, 0.0, 0.3, 10, cpu_threads(), kind="ses")
    b = oracle.Surface(s, 1.4, 0.3, 10, cpu_threads(), kind="vdw")
    assert a.n_circles == 0 and b.n_circles == 0
    ma, mb = a.mesh(), b.mesh()
    for k in ma:
        assert np.array_equal(ma[k], mb[k]), k


def test_sphere_kind_field_samples():
    """SesField::sample of the sphere kinds (point.rs:340-368): value
    |x - c| - R of the nearest sphere, closest point on it, feature that
    sphere."""
    s = np.array([[0.0, 0.0, 0.0, 1.6], [3.0, 0.0, 0.0, 1.2]])
    o = oracle.Surface(s, 1.4, 0.5, 10, 1, stage=1, kind="sas")
    arr = o.arrangement()
    for x in ([0.0, 0.0, 5.0], [3.0, 0.5, 0.0], [1.5, 0.0, 0.0], [-4.0, 1.0, 1.0]):
        x = np.array(x)
        v, closest, (fk, fid) = o.sample(x)
        F, d = sphere_F(x[None, :], expanded(s, "sas", 1.4))
        assert abs(v - F[0]) < 1e-12
        k = int(np.argmin(d[0]))
        assert fk == 2 and arr["orig"][fid] == k
        c = s[k, :3]
        want = c + (x - c) / np.linalg.norm(x - c) * (s[k, 3] + 1.4)
        assert np.allclose(closest, want, atol=1e-12)
