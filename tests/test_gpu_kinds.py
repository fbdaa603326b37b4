"""The SAS and vdW kinds on the device against the oracle, bit for bit.

The reference's device path meshes all three kinds and checks each against
its CPU back end (crates/sesmesh/src/field/gpu.rs:846-1010 loop over
SurfaceKind::{Ses, Sas, Vdw}); so do these tests: samples, vertices,
normals, atoms, patches, triangles, edge parameters and evaluation counts.
The analytic single-sphere check of crates/sesmesh/tests/surface.rs:56-84
runs on the device's measure stage, batched runs mix the kinds, and the
fast refinement mode (which takes the exact search for these kinds) must
give the exact mode's mesh.
"""
import math

import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads
from test_gpu_batch import INFO, KEYS, same_bits

pytestmark = pytest.mark.gpu


def assert_same_as_oracle(o, m, values=True):
    g = m.arrays(values=values)
    assert m.info.touched_bricks == o.touched
    keys = o.bricks()
    assert np.array_equal(g["bricks"], keys)
    if values:
        assert np.array_equal(g["values"].view(np.uint32), o.brick_values(keys).view(np.uint32))
    om = o.mesh()
    assert m.n_vertices == o.n_vertices and m.n_triangles == o.n_triangles
    for k in ("positions", "normals"):
        assert np.array_equal(g[k].view(np.uint32), om[k].view(np.uint32)), k
    for k in ("atoms", "patches", "triangles", "evals"):
        assert np.array_equal(g[k], om[k]), k
    assert np.array_equal(g["t"].view(np.uint64), om["t"].view(np.uint64))
    assert m.info.refine_evaluations == o.evals
    return g


def both(device, xyzr, kind, spacing, brick=10, probe=1.4):
    print(f"kind {kind} spacing {spacing} brick {brick} atoms {len(xyzr)}", flush=True)
    o = oracle.Surface(xyzr, probe, spacing, brick, cpu_threads(), kind=kind)
    arr = sesmesh_amd.Arrangement(xyzr, probe, kind=kind)
    assert arr.kind == kind and arr.ref.n_circles == 0 and arr.ref.n_vertices == 0
    res = sesmesh_amd.Resident(device, arr)
    return o, res, res.mesh(spacing, brick)


@pytest.mark.parametrize("kind", ["sas", "vdw"])
@pytest.mark.parametrize("spacing", [0.5, 0.25, 1.0])
def test_crambin_equals_cpu(device, structures, kind, spacing):
    o, _, m = both(device, structures("1CRN"), kind, spacing)
    assert_same_as_oracle(o, m)
    assert m.n_vertices > 1000


@pytest.mark.parametrize("kind", ["sas", "vdw"])
def test_globule_and_large_bricks_equal_cpu(device, kind):
    xyzr = sesmesh_amd.synth_globule(3, 900)
    for brick, spacing in ((10, 0.35), (24, 0.5)):
        o, _, m = both(device, xyzr, kind, spacing, brick)
        assert_same_as_oracle(o, m, values=brick == 10)


@pytest.mark.parametrize("kind,R", [("sas", 3.2), ("vdw", 1.8)])
def test_single_sphere_matches_the_closed_form(device, kind, R):
    """crates/sesmesh/tests/surface.rs:56-84 on the device's mesh and
    measure stage."""
    s = np.array([[0.3, -0.2, 0.1, 1.8]])
    o, _, m = both(device, s, kind, 0.25)
    assert_same_as_oracle(o, m)
    m.measure()
    st = m.measure_info()
    assert st["components"] == 1 and st["euler"] == 2
    assert abs(st["area"] - 4 * math.pi * R * R) / (4 * math.pi * R * R) < 5e-3
    assert abs(st["volume"] - 4 / 3 * math.pi * R ** 3) / (4 / 3 * math.pi * R ** 3) < 1.2e-2
    g = m.arrays()
    pos = g["positions"].astype(np.float64)
    F = np.linalg.norm(pos - s[0, :3], axis=1) - R
    assert np.abs(F).max() < 1e-4 * 0.25 + 2e-6
    assert set(np.unique(g["patches"])) == {0}


def test_sas_is_vdw_of_expanded_radii_on_the_device(device, structures):
    xyzr = structures("1CRN")
    e = xyzr.copy()
    e[:, 3] += 1.4
    _, _, a = both(device, xyzr, "sas", 0.5)
    _, _, b = both(device, e, "vdw", 0.5)
    ga, gb = a.arrays(values=True), b.arrays(values=True)
    for k in KEYS:
        assert same_bits(ga[k], gb[k]), k


def test_batched_run_mixes_kinds(device, structures):
    """One smg_mesh_many run over SES, SAS and vdW molecules: part i equals
    the mesh of molecule i alone."""
    mols = [(structures("1CRN"), "ses", 0.5), (structures("1CRN"), "sas", 0.5), (structures("1CRN"), "vdw", 0.35),
            (sesmesh_amd.synth_globule(4, 400), "sas", 0.4), (sesmesh_amd.synth_globule(5, 300), "vdw", 0.5)]
    res = [sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(x, 1.4, kind=k)) for x, k, _ in mols]
    hs = [h for _, _, h in mols]
    batch = sesmesh_amd.mesh_many(res, hs)
    for i, (r, h) in enumerate(zip(res, hs)):
        alone = r.mesh(h, 10)
        a, b = batch.arrays(values=True, part=i), alone.arrays(values=True)
        for k in KEYS:
            assert same_bits(a[k], b[k]), (i, k)
        for f in INFO:
            assert getattr(batch.parts[i], f) == getattr(alone.parts[0], f), (i, f)


@pytest.mark.parametrize("kind", ["sas", "vdw"])
def test_fast_mode_takes_the_exact_search(device, structures, kind):
    xyzr = structures("1CRN")
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4, kind=kind))
    ex = res.mesh(0.5, 10).arrays(values=True)
    fa = res.mesh(0.5, 10, refine="fast").arrays(values=True)
    for k in KEYS:
        assert same_bits(ex[k], fa[k]), k
    assert not np.any(fa["evals"] & (1 << 30))


@pytest.mark.parametrize("kind", ["sas", "vdw"])
def test_field_clamped_equals_cpu(device, structures, kind):
    xyzr = structures("1CRN")
    h = 0.5
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4, kind=kind))
    g = res.mesh(h, 10).arrays()
    rng = np.random.default_rng(7)
    pts = g["positions"][rng.choice(len(g["positions"]), 500, replace=False)].astype(np.float64)
    pts = np.concatenate([pts, pts + rng.uniform(-0.8, 0.8, pts.shape)])
    got = res.field_clamped(pts, 4.0)
    o = oracle.Surface(xyzr, 1.4, h, 10, cpu_threads(), stage=1, kind=kind)
    want = np.array([o.sample(p)[0] for p in pts])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_plan_replays_sphere_kinds(device, structures):
    mols = [(structures("1CRN"), "sas", 0.5), (sesmesh_amd.synth_globule(6, 500), "vdw", 0.5),
            (structures("1CRN"), "ses", 0.5)]
    res = [sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(x, 1.4, kind=k)) for x, k, _ in mols]
    hs = [h for _, _, h in mols]
    want = sesmesh_amd.mesh_many(res, hs)
    plan = sesmesh_amd.Plan(device, sesmesh_amd.Plan.caps_for(res, hs, mesh=want, margin=0.25))
    for rep in range(3):
        got = plan.run(res, hs)
        for i in range(len(mols)):
            a, b = got.arrays(values=True, part=i), want.arrays(values=True, part=i)
            for k in KEYS:
                assert same_bits(a[k], b[k]), (rep, i, k)
        del got
        device.selftest(1 << 12, rep + 1)


def test_kinds_are_checked(device):
    xyzr = sesmesh_amd.synth_globule(5, 50)
    with pytest.raises(ValueError):
        sesmesh_amd.Arrangement(xyzr, 1.4, kind="sphere")
    a = sesmesh_amd.Arrangement(xyzr, 0.0, kind="ses")
    assert a.kind == "vdw" and a.probe == 0.0
    b = sesmesh_amd.Arrangement(xyzr, 1.4, kind="vdw")
    assert b.kind == "vdw" and b.probe == 0.0
