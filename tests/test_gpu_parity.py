"""Device mesh stage against the oracle (CPU back end), bit for bit.

The reference's acceptance checks for its device path are restated here
(crates/sesmesh/src/field/gpu.rs:846-962): the device's samples equal the
CPU's, and the device's meshes equal the CPU's (positions, normals, atoms,
patches, triangles). This build goes further than the reference: the
classification and the f64 vertex refinement also run on the device, and
every sample value (not only those the CPU evaluates) is compared.
"""
import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu


def run_both(device, xyzr, probe, spacing, brick=10, values=True):
    o = oracle.Surface(xyzr, probe, spacing, brick, cpu_threads())
    arr = sesmesh_amd.Arrangement(xyzr, probe)
    res = sesmesh_amd.Resident(device, arr)
    m = res.mesh(spacing, brick)
    g = m.arrays(values=values)
    return o, m, g


def assert_same_mesh(o, m, g, values=True):
    assert m.info.touched_bricks == o.touched
    keys = o.bricks()
    assert np.array_equal(g["bricks"], keys)
    if values:
        ov = o.brick_values(keys)
        assert np.array_equal(g["values"].view(np.uint32), ov.view(np.uint32))
    om = o.mesh()
    assert m.n_vertices == o.n_vertices and m.n_triangles == o.n_triangles
    for k in ("positions", "normals"):
        assert np.array_equal(g[k].view(np.uint32), om[k].view(np.uint32)), k
    for k in ("atoms", "patches", "triangles", "evals"):
        assert np.array_equal(g[k], om[k]), k
    assert np.array_equal(g["t"].view(np.uint64), om["t"].view(np.uint64))
    assert m.info.refine_evaluations == o.evals


def test_device_opens_with_selftest(device):
    assert device.arch.startswith("gfx950")
    for seed in (1, 2):
        assert device.selftest(1 << 20, seed) == (0, 0)


@pytest.mark.parametrize("spacing", [0.5, 0.25, 0.75, 1.0])
def test_crambin_equals_cpu(device, structures, spacing):
    # 0.75 and 1.0 A: larger sub-block balls, long near lists and overflow
    # tiers in the gathers.
    o, m, g = run_both(device, structures("1CRN"), 1.4, spacing)
    assert_same_mesh(o, m, g)
    assert m.n_vertices == {0.5: 13812, 0.25: 55542, 0.75: 6064, 1.0: 3422}[spacing]


@pytest.mark.parametrize("brick,spacing", [(18, 0.25), (24, 0.5), (26, 0.5), (32, 0.5)])
def test_large_bricks_equal_cpu(device, structures, brick, spacing):
    """The edge and cell passes keep a brick's sample signs and (cell pass)
    the edge masks of it and its seven upper neighbours in LDS; at brick 32
    the cell pass asks for 156,056 dynamic bytes, within a workgroup's 160
    KiB with its static tables. At brick 18 and 0.25 A, 51 of crambin's 233
    surface bricks list more triangles (up to 1,694) than the cell pass's
    1,024 LDS slots and take the per-thread path. The mesh must equal the
    CPU back end's at the same brick size
    (crates/sesmesh/src/field/gpu.rs:906-962)."""
    o, m, g = run_both(device, structures("1CRN"), 1.4, spacing, brick)
    assert_same_mesh(o, m, g)


def test_synthetic_330_atom_globule_equals_cpu(device):
    xyzr = sesmesh_amd.synth_globule(1, 330)
    o, m, g = run_both(device, xyzr, 1.4, 0.5)
    assert_same_mesh(o, m, g)


def _chain(rng, n):
    c = [np.zeros(3)]
    while len(c) < n:
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c.append(c[rng.integers(len(c))] + u * rng.uniform(1.3, 1.6))
    return np.column_stack([np.array(c), rng.uniform(1.2, 2.0, n)])


def test_random_clusters_equal_cpu(device):
    rng = np.random.default_rng(2026)
    for case in range(12):
        n = int(rng.integers(1, 40))
        if case % 3 == 2:
            xyzr = np.column_stack([rng.uniform(0, 2 + n ** (1 / 3) * 3, (n, 3)), rng.uniform(1.0, 2.2, n)])
        else:
            xyzr = _chain(rng, n)
        probe = float(rng.uniform(0.6, 1.8))
        spacing = float(rng.choice([0.5, 0.35, 0.25]))
        brick = int(rng.choice([4, 6, 8, 10, 12]))
        o, m, g = run_both(device, xyzr, probe, spacing, brick)
        assert_same_mesh(o, m, g)


def test_symmetric_and_degenerate_sets_equal_cpu(device):
    g3 = np.arange(3) * 3.0
    lattice = np.array([[x, y, z, 1.7] for x in g3 for y in g3 for z in g3])
    phi = (1 + 5 ** 0.5) / 2
    ico = []
    for a, b in ((1, phi), (-1, phi), (1, -phi), (-1, -phi)):
        ico += [(0, a, b), (a, b, 0), (b, 0, a)]
    cage = np.array([np.array(p) / np.linalg.norm(p) * 3.5 for p in ico])
    cage = np.column_stack([cage, np.full(len(cage), 1.6)])
    collinear = np.array([[2.0 * i, 0, 0, 1.5] for i in range(8)])
    single = np.array([[0.3, -0.2, 0.1, 1.8]])
    rng = np.random.default_rng(3)
    box = np.column_stack([rng.uniform(0, 10, (30, 3)), rng.uniform(1.2, 2.0, 30)])
    dup = np.vstack([box, box[::4], [box[0] + [0.1, 0, 0, -0.5]]])
    for xyzr, probe, h in ((lattice, 1.4, 0.5), (cage, 1.4, 0.25), (collinear, 1.0, 0.5), (single, 1.4, 0.25),
                           (dup, 1.4, 0.5), (lattice, 0.4, 0.35)):
        o, m, gd = run_both(device, xyzr, probe, h)
        assert_same_mesh(o, m, gd)


@pytest.mark.parametrize("entry", ["1UBQ", "2PTN"])
def test_structures_at_quarter_angstrom_equal_cpu(device, structures, golden, entry):
    o, m, g = run_both(device, structures(entry), 1.4, 0.25)
    assert_same_mesh(o, m, g)
    r = [x for x in golden["runs"] if x["campaign"] == "acc" and x["entry"] == entry and x["spacing"] == 0.25][0]
    assert m.n_vertices == r["vertices"] and m.n_triangles == r["triangles"]


def test_device_reads_static_lds_of_its_kernels(device, tmp_path):
    """The LDS plan of every run uses the static LDS the runtime reports for
    the loaded kernels: it equals the code object's kernel descriptors."""
    import ctypes

    from test_lds_budget import KINDS, kernel_statics

    want = kernel_statics(tmp_path)
    got = (ctypes.c_uint32 * len(KINDS))()
    assert sesmesh_amd.lib().smg_device_static_lds(device._h, got) == 0
    assert list(got) == want


def test_objects_outlive_a_closed_device_handle(structures):
    """The device lives until the last object made on it is freed: closing
    the handle first, then copying a mesh and freeing in any order works."""
    xyzr = structures("1CRN")
    dev = sesmesh_amd.Device(0)
    res = sesmesh_amd.Resident(dev, sesmesh_amd.Arrangement(xyzr, 1.4))
    m = res.mesh(0.5, 10)
    dev.close()
    g = m.arrays()
    assert len(g["positions"]) == m.n_vertices == 13812
    m2 = res.mesh(0.5, 10)
    del res
    assert np.array_equal(m2.arrays()["triangles"], g["triangles"])
    del m, m2


def test_interpreter_exit_with_live_objects():
    """Objects left alive at interpreter exit (device closed first, or not
    at all) end the process cleanly."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r); import sesmesh_amd as s, numpy as np\n"
            "d = s.Device(0); x = s.synth_globule(1, 330)\n"
            "r = s.Resident(d, s.Arrangement(x, 1.4)); m = r.mesh(0.5, 10)\n"
            "p = s.Plan(d, s.Plan.caps_for([r], [0.5], mesh=m)); v = p.run([r], [0.5])\n"
            "d.close(); print('done', m.n_vertices, v.n_vertices)\n") % str(sesmesh_amd.HERE.parent)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.startswith("done 11900 11900")
