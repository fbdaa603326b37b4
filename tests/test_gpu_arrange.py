"""The SAS arrangement built on the device (smg_resident_build) against the
host builder, array for array and bit for bit.

The host builder is pinned to the oracle (tests/test_library_host.py), and
through it to the reference's Arrangement::build_with
(crates/sesmesh/src/geom/arrangement.rs:437-645) and the field's grids
(field/point.rs:213-274). Compared here: the sphere order and every sphere
array, the full set of circles with their blockers and triple points in
order, the field's neighbour rows, and the three cell grids (origin, cell,
dimensions, starts, items). Meshes from a device-built resident equal the
oracle's. A sphere set whose exposed-arc test falls within the device's
atan2 margin is built on the host (smg_resident_info says which ran); the
arrays are equal either way.
"""
import time

import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def compare(dev_arr, host_arr):
    d, h = dev_arr.arrays(), host_arr.arrays()
    for k in ("orig", "centers", "expanded", "radii", "status", "circles_ij", "circles", "vertices", "vertex_atoms"):
        assert same(d[k], h[k]), k
    for k in ("neighbors", "circle_blockers"):
        assert same(d[k][0], h[k][0]) and same(d[k][1], h[k][1]), k
    assert d["max_expanded"] == h["max_expanded"] and d["rho_max"] == h["rho_max"]
    for g in ("sphere_grid", "circle_grid", "vertex_grid"):
        dg, hg = d["grids"][g], h["grids"][g]
        assert dg["dims"] == hg["dims"] and dg["cell"] == hg["cell"], g
        assert same(dg["origin"], hg["origin"]) and same(dg["starts"], hg["starts"]) and same(dg["items"], hg["items"]), g
    assert dev_arr.kind == host_arr.kind and dev_arr.probe == host_arr.probe


def build_both(device, xyzr, probe=1.4, kind="ses"):
    print(f"arrangement: {len(xyzr)} spheres, probe {probe}, {kind}", flush=True)
    t = time.perf_counter()
    r = sesmesh_amd.Resident.build(device, xyzr, probe, kind)
    t_dev = time.perf_counter() - t
    t = time.perf_counter()
    h = sesmesh_amd.Arrangement(xyzr, probe, kind=kind)
    t_host = time.perf_counter() - t
    compare(r.arrangement(), h)
    return r, h, t_dev, t_host


@pytest.mark.parametrize("entry", ["1CRN", "1UBQ", "2PTN", "1EJG"])
def test_structures(device, structures, entry):
    r, h, _, _ = build_both(device, structures(entry))
    info = r.info()
    assert info["n_circles"] == h.ref.n_circles and info["n_vertices"] == h.ref.n_vertices
    assert info["n_circles"] > 100


@pytest.mark.parametrize("kind", ["sas", "vdw"])
def test_sphere_kinds(device, structures, kind):
    r, h, _, _ = build_both(device, structures("1CRN"), kind=kind)
    assert r.info()["built_on_device"] and h.ref.n_circles == 0


def test_degenerate_and_edge_sets(device):
    rng = np.random.default_rng(5)
    g = np.arange(3) * 3.0
    lattice = np.array([[x, y, z, 1.7] for x in g for y in g for z in g])
    chain = np.array([[2.0 * i, 0, 0, 1.5] for i in range(8)])
    box = np.column_stack([rng.uniform(0, 10, (30, 3)), rng.uniform(1.2, 2.0, 30)])
    box = np.vstack([box, box[::4], [box[0] + [0.1, 0, 0, -0.5]], [[1, 1, 1, 0.0]]])
    sets = [(lattice, 1.4), (chain, 1.0), (np.abs(box), 1.4), (np.array([[0.3, -0.2, 0.1, 1.8]]), 1.4),
            (np.zeros((0, 4)), 1.4), (np.array([[0, 0, 0, 0.0], [1, 1, 1, 0.0]]), 1.4)]
    for seed in range(6):
        n = int(rng.integers(2, 60))
        c = rng.normal(size=(n, 3)) * rng.uniform(1, 5)
        sets.append((np.column_stack([c, rng.uniform(1.0, 2.2, n)]), float(rng.uniform(0.5, 2.0))))
    for xyzr, probe in sets:
        build_both(device, xyzr, probe)


def test_meshes_from_a_device_built_resident_equal_cpu(device, structures):
    xyzr = structures("1CRN")
    r = sesmesh_amd.Resident.build(device, xyzr, 1.4)
    assert r.info()["built_on_device"]
    m = r.mesh(0.5, 10)
    g = m.arrays(values=True)
    o = oracle.Surface(xyzr, 1.4, 0.5, 10, cpu_threads())
    om = o.mesh()
    assert np.array_equal(g["values"].view(np.uint32), o.brick_values(o.bricks()).view(np.uint32))
    for k in ("positions", "normals"):
        assert np.array_equal(g[k].view(np.uint32), om[k].view(np.uint32)), k
    for k in ("atoms", "patches", "triangles", "evals"):
        assert np.array_equal(g[k], om[k]), k


def test_assembly_of_58k_atoms(device):
    """The two-ring assembly meshed at 0.25 A (57,400 atoms)."""
    xyzr = sesmesh_amd.synth_rings(seed=2)
    r, h, t_dev, t_host = build_both(device, xyzr)
    print(f"[arrangement] 58k atoms: device {r.info()['build_ms']:.1f} ms ({t_dev * 1e3:.1f} ms wall), "
          f"host {t_host * 1e3:.1f} ms, on device: {r.info()['built_on_device']}", flush=True)
    assert r.info()["n_circles"] == h.ref.n_circles


def test_shell_of_2_4m_atoms(device):
    """The 2.4M-atom shell of 60 globule copies."""
    xyzr = sesmesh_amd.synth_shell()
    t = time.perf_counter()
    r = sesmesh_amd.Resident.build(device, xyzr, 1.4)
    t_dev = time.perf_counter() - t
    info = r.info()
    print(f"[arrangement] 2.4M atoms: device {info['build_ms']:.1f} ms ({t_dev * 1e3:.1f} ms wall), "
          f"{info['n_circles']} circles, {info['n_vertices']} vertices, on device: {info['built_on_device']}",
          flush=True)
    t = time.perf_counter()
    h = sesmesh_amd.Arrangement(xyzr, 1.4, threads=16)
    print(f"[arrangement] 2.4M atoms: host {1e3 * (time.perf_counter() - t):.1f} ms on 16 threads", flush=True)
    compare(r.arrangement(), h)
