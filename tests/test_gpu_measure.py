"""The measure stage on the device against the oracle, bit for bit.

Validation counts, components (vertices, triangles, f64 area and volume
summed in the reference's chunk order, Euler characteristic, first
triangle), vertex components, triangles grouped by component and the
stats (crates/sesmesh/src/mesh/mod.rs:91-705, post.rs:14-34,
surface.rs:1081-1168): on the reference's known-answer meshes and random
defect meshes uploaded to the device, on device meshes of real structures,
where the counts, components, cavities and every printed digit of the area
are the reference's recorded `acc` runs, and on the parts of a batched run.
"""
import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads
from test_oracle_measure import OCT_P, OCT_T, defect_meshes

pytestmark = pytest.mark.gpu


def assert_measure_equal(dm, om, part=0, grouped=None):
    """Device mesh `dm` (measured) against oracle result `om`."""
    info = dm.measure_info(part)
    assert info["validation"] == om["validation"]
    assert info["measured"] == om["indices_valid"]
    if not om["indices_valid"]:
        return
    assert info["components"] == om["components"] and info["cavities"] == om["cavities"]
    assert info["euler"] == om["euler"]
    assert np.float64(info["area"]).view(np.uint64) == np.float64(om["area"]).view(np.uint64)
    assert np.float64(info["volume"]).view(np.uint64) == np.float64(om["volume"]).view(np.uint64)
    comp, vc = dm.components(part)
    assert np.array_equal(vc, om["vertex_component"])
    for k in ("vertices", "triangles", "euler", "first_triangle"):
        assert np.array_equal(comp[k].astype(np.int64), om["comp"][k].astype(np.int64)), k
    for k in ("area", "volume"):
        assert np.array_equal(comp[k].astype(np.float64).view(np.uint64), om["comp"][k].view(np.uint64)), k
    tris = dm.arrays(part=part)["triangles"] if grouped is None else grouped
    assert np.array_equal(tris, om["grouped"])


def test_known_answer_meshes_equal_oracle(device):
    cases = [(OCT_P, OCT_T, OCT_P), (OCT_P, OCT_T[:, [0, 2, 1]], OCT_P),
             (np.vstack([OCT_P, OCT_P + [5, 0, 0]]), np.vstack([OCT_T, OCT_T + 6]), np.vstack([OCT_P, OCT_P])),
             (OCT_P, OCT_T[:-1], OCT_P), (OCT_P, np.vstack([OCT_T, OCT_T[:1]]), OCT_P),
             (OCT_P, np.vstack([OCT_T, [[0, 0, 1]]]), OCT_P), (OCT_P, np.vstack([OCT_T, [[0, 1, 99]]]), OCT_P),
             (OCT_P, np.vstack([OCT_T, [[0, 1, 1], [1, 0, 1], [1, 1, 0], [2, 2, 2], [2, 2, 2], [3, 4, 3],
                                        [3, 3, 4], [4, 3, 4]]]), OCT_P),
             (OCT_P, OCT_T, None), (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), None)]
    bad_p, bad_n = OCT_P.copy(), OCT_P.copy()
    bad_p[1, 0] = np.nan
    bad_n[0] = 0
    cases.append((bad_p, OCT_T, bad_n))
    for p, t, n in cases:
        om = oracle.measure(p, t, n)
        dm = sesmesh_amd.upload_mesh(device, p, t, normals=n)
        dm.measure()
        assert_measure_equal(dm, om)
    om = oracle.measure(OCT_P, OCT_T, OCT_P)
    assert abs(om["area"] - 4 * 3 ** 0.5) < 1e-12 and om["euler"] == 2


def test_random_defect_meshes_equal_oracle(device):
    for seed in (0x5E5, 7):
        for p, t in defect_meshes(seed):
            om = oracle.measure(p, t, p)
            dm = sesmesh_amd.upload_mesh(device, p, t, normals=p)
            dm.measure()
            assert_measure_equal(dm, om)


@pytest.mark.parametrize("entry,spacing", [("1CRN", 0.5), ("1CRN", 0.25), ("1UBQ", 0.5), ("1UBQ", 0.25),
                                           ("2PTN", 0.5), ("2PTN", 0.25)])
def test_device_meshes_measure_as_the_reference_recorded(device, structures, golden, entry, spacing):
    xyzr = structures(entry)
    o = oracle.Surface(xyzr, 1.4, spacing, 10, cpu_threads())
    om = o.measure_full()
    m = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4)).mesh(spacing, 10)
    ms = m.measure()
    assert ms > 0
    assert_measure_equal(m, om)
    info = m.measure_info()
    assert info["validation"] == {k: 0 for k in info["validation"]}
    r = [x for x in golden["runs"] if x["campaign"] == "acc" and x["entry"] == entry and x["spacing"] == spacing][0]
    assert (m.n_vertices, m.n_triangles, info["components"], info["cavities"]) == (
        r["vertices"], r["triangles"], r["components"], r["cavities"])
    assert info["area"] == r["area"]
    m.measure()  # measuring again leaves the grouped mesh and its results as they are
    assert_measure_equal(m, om)


def test_batched_parts_measure_as_single_runs(device):
    xs = [sesmesh_amd.synth_globule(40 + i, n) for i, n in enumerate((300, 1200, 25, 800, 4000))]
    rs = [sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(x, 1.4)) for x in xs]
    hs = [0.5, 0.35, 0.25, 0.5, 0.4]
    mb = sesmesh_amd.mesh_many(rs, hs, 10)
    mb.measure()
    for i, (r, h) in enumerate(zip(rs, hs)):
        m = r.mesh(h, 10)
        m.measure()
        a, b = mb.measure_info(i), m.measure_info()
        assert a == b
        ca, va = mb.components(i)
        cb, vb = m.components()
        assert np.array_equal(va, vb) and all(np.array_equal(ca[k], cb[k]) for k in ca)
        assert np.array_equal(mb.arrays(part=i)["triangles"], m.arrays()["triangles"])
        if i == 3:
            o = oracle.Surface(xs[i], 1.4, h, 10, cpu_threads())
            assert_measure_equal(mb, o.measure_full(), part=i)
