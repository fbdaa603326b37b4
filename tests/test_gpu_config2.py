"""The headline workload against the oracle, bit for bit, at full size.

The 58k-atom two-ring assembly (14 copies of a seeded 4.1k-atom globule in
two stacked 7-fold rings), SES probe 1.4 A, spacing 0.25 A, brick 10, exact
refinement: the configuration the speed is quoted on. Meshed on the device
through the plain path (smg_mesh_new) and through a capture-ready plan
replayed as a graph, and compared with the oracle's CPU mesh stage: surface
brick keys, every sample, vertex positions, normals, atoms, patches, edge
parameters and evaluation counts, and triangles. These are the reference's
acceptance checks for its device path (crates/sesmesh/src/field/gpu.rs:
846-903 samples, 906-962 meshes), applied to every brick rather than a
sample. The hidden-candidate cover margin and the lower-bound pruning of the
refinement rest on margin arguments; this is their test at the size the
bench runs.
"""
import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu

PROBE, SPACING, BRICK = 1.4, 0.25, 10


@pytest.fixture(scope="module")
def assembly():
    xyzr = sesmesh_amd.synth_rings(seed=2)
    assert len(xyzr) == 57400
    print(f"[config2] oracle on {cpu_threads()} threads", flush=True)
    o = oracle.Surface(xyzr, PROBE, SPACING, BRICK, cpu_threads())
    keys = o.bricks()
    ref = {"bricks": keys, "values": o.brick_values(keys), **o.mesh()}
    print(f"[config2] oracle: {o.n_vertices} vertices, {o.n_triangles} triangles", flush=True)
    return xyzr, o, ref


def assert_equal_to_oracle(o, ref, m, g):
    assert m.info.touched_bricks == o.touched
    assert m.n_bricks == o.n_bricks and m.n_vertices == o.n_vertices and m.n_triangles == o.n_triangles
    assert np.array_equal(g["bricks"], ref["bricks"])
    assert np.array_equal(g["values"].view(np.uint32), ref["values"].view(np.uint32))
    for k in ("positions", "normals"):
        assert np.array_equal(g[k].view(np.uint32), ref[k].view(np.uint32)), k
    for k in ("atoms", "patches", "triangles", "evals"):
        assert np.array_equal(g[k], ref[k]), k
    assert np.array_equal(g["t"].view(np.uint64), ref["t"].view(np.uint64))
    assert m.info.refine_evaluations == o.evals


def test_assembly_plain_and_replayed_plan_equal_cpu(device, assembly):
    xyzr, o, ref = assembly
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, PROBE))
    m = res.mesh(SPACING, BRICK)
    assert m.n_vertices == 6588892 and m.n_triangles == 13179908
    assert_equal_to_oracle(o, ref, m, m.arrays(values=True))
    plan = sesmesh_amd.Plan(device, sesmesh_amd.Plan.caps_for([res], [SPACING], brick=BRICK, mesh=m, margin=0.05))
    del m
    for run in range(3):  # the first run records the graph, the later ones replay it
        v = plan.run([res], [SPACING])
        if run:
            assert_equal_to_oracle(o, ref, v, v.arrays(values=True))
        del v
    st = plan.stats()
    assert st["fallbacks"] == 0 and st["replays"] == 2 and st["captures"] == 1


def test_assembly_measure_equals_cpu(device, assembly):
    """The device measure stage of the 58k assembly's mesh (validation,
    components, chunk-ordered f64 sums, grouping) equals the oracle's."""
    from test_gpu_measure import assert_measure_equal

    xyzr, o, ref = assembly
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, PROBE))
    m = res.mesh(SPACING, BRICK)
    ms = m.measure()
    print(f"[config2] measure stage on the device: {ms:.2f} ms", flush=True)
    om = o.measure_full()
    assert_measure_equal(m, om)
    assert om["validation"] == {k: 0 for k in om["validation"]}


def test_assembly_files_equal_cpu(device, assembly, tmp_path):
    """OBJ, PLY and MSMS files of the measured 58k-assembly mesh, written on
    the device as the reference's command line writes them, equal the
    oracle's files byte for byte."""
    from test_gpu_export import cli_files

    xyzr, o, ref = assembly
    om = o.measure_full()
    omesh = {k: ref[k] for k in ("positions", "normals", "atoms", "patches")}
    omesh["triangles"], omesh["vertex_component"] = om["grouped"], om["vertex_component"]
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, PROBE))
    m = res.mesh(SPACING, BRICK)
    m.measure()
    groups, hd = cli_files(omesh, om, len(xyzr), PROBE, "rings.pdb")
    m.write(tmp_path / "a.obj", "obj", groups=groups)
    assert (tmp_path / "a.obj").read_bytes() == oracle.write_obj(omesh, 4, True, (), "ses", None, groups)
    m.write(tmp_path / "a.ply", "ply")
    assert (tmp_path / "a.ply").read_bytes() == oracle.write_ply(omesh)
    v, f = oracle.write_msms(omesh, hd["description"], hd["spheres"], hd["density"], hd["probe_radius"])
    m.write(tmp_path / "a.vert", "msms_vert", **hd)
    m.write(tmp_path / "a.face", "msms_face", **hd)
    assert (tmp_path / "a.vert").read_bytes() == v and (tmp_path / "a.face").read_bytes() == f
