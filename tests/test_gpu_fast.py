"""Fast (f32 compensated) refinement and the exact point evaluator it is
measured with.

* smg_field_clamped equals the CPU back end's exact field (SesField::value /
  clamped_value, reference crates/sesmesh/src/field/point.rs:287-330,
  759-765) bit for bit at points near the surface.
* Fast mode keeps the exact mode's topology (the same samples, so the same
  crossing edges and triangles) and moves each vertex by about f32
  rounding: its vertices lie on the surface to within the exact mode's own
  distance plus 1e-5 A, and area and volume agree to 1e-5 relative. The
  reference has no fast mode; its surface tolerance |F| <= 1e-4 h + 2e-6
  (tests/surface.rs:46-49, 202-241) bounds both.
"""
import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu


def test_field_clamped_equals_cpu_near_the_surface(device, structures):
    xyzr = structures("1CRN")
    probe, h = 1.4, 0.5
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, probe))
    g = res.mesh(h, 10).arrays()
    rng = np.random.default_rng(5)
    pts = g["positions"][rng.choice(len(g["positions"]), 600, replace=False)].astype(np.float64)
    pts = np.concatenate([pts, pts + rng.uniform(-0.8, 0.8, pts.shape)])
    got = res.field_clamped(pts, 4.0)
    o = oracle.Surface(xyzr, probe, h, 10, cpu_threads(), stage=1)
    want = np.array([o.sample(p)[0] for p in pts])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("case", ["1CRN-0.25", "globule-0.5", "1UBQ-0.35"])
def test_fast_mode_tracks_the_exact_surface(device, structures, case):
    name, h = case.split("-")
    xyzr = sesmesh_amd.synth_globule(9, 1200) if name == "globule" else structures(name)
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4))
    rep = sesmesh_amd.precision_report(res, float(h))
    e, f = rep["exact"], rep["fast"]
    assert rep["same_topology"]
    assert f["max_abs_F"] <= e["max_abs_F"] + 1e-5
    assert f["max_abs_F"] <= 1e-4 * float(h) + 2e-6 + 3e-5
    assert f["mean_abs_F"] <= 1e-5
    # Both vertices are roots within the 1e-5 tolerance on F; where F is
    # nearly flat along the edge that allows a larger shift along it.
    assert rep["max_vertex_shift"] <= 0.01 * float(h)
    assert rep["mean_vertex_shift"] <= 1e-5
    assert rep["rel_area_error"] <= 1e-5 and rep["rel_volume_error"] <= 1e-5


def test_fast_mode_needs_one_mode_per_run(device):
    xyzr = sesmesh_amd.synth_globule(5, 100)
    r = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4))
    params = [sesmesh_amd._mesh_params(0.5, 10, "exact", 0), sesmesh_amd._mesh_params(0.5, 10, "fast", 0)]
    with pytest.raises(sesmesh_amd.BackendError):
        sesmesh_amd.Mesh(None, None, 10, "exact", 0, _batch=([r, r], params))


def test_unconverged_fast_edges_are_redone_exactly(device, structures):
    """Edges whose f32 Newton / Illinois search does not converge are
    redone by the exact f64 search (the reference redoes such device edges
    on the CPU, crates/sesmesh/src/mesh/extract.rs:336-358): those vertices
    (evaluation word bit 30) equal exact mode's bit for bit."""
    redone = 0
    for xyzr, h in ((structures("1UBQ"), 0.25), (sesmesh_amd.synth_globule(9, 3000), 0.5), (structures("2PTN"), 0.35)):
        res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4))
        e = res.mesh(h, 10, refine="exact").arrays()
        f = res.mesh(h, 10, refine="fast").arrays()
        assert np.array_equal(e["triangles"], f["triangles"])
        sel = ((f["evals"] >> 30) & 1).astype(bool)
        redone += int(sel.sum())
        for k in ("positions", "normals"):
            assert np.array_equal(f[k][sel].view(np.uint32), e[k][sel].view(np.uint32)), k
        for k in ("atoms", "patches"):
            assert np.array_equal(f[k][sel], e[k][sel]), k
        assert np.array_equal(f["t"][sel].view(np.uint64), e["t"][sel].view(np.uint64))
        assert np.array_equal(f["evals"][sel] & ~np.uint32(1 << 30), e["evals"][sel])
        # converged fast edges carry bit 31, redone ones never
        assert not np.any((f["evals"][sel] >> 31) & 1 & ((e["evals"][sel] >> 31) ^ 1))
    assert redone > 0
