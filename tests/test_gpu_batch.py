"""Batched mesh stage (smg_mesh_many): each part of a batched run equals, byte
for byte, the single-molecule run of the same molecule, and through it the
CPU back end. The reference meshes one molecule per `fold_resident` run
(crates/sesmesh/src/field/gpu.rs:532-712); its device acceptance check
(gpu.rs:906-962, device meshes equal CPU meshes) is restated per part.
"""
import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu

KEYS = ("positions", "normals", "atoms", "patches", "triangles", "bricks", "t", "evals", "values")
INFO = ("touched_bricks", "surface_bricks", "vertices", "triangles", "refine_evaluations", "sampled_subblocks",
        "evaluated_samples", "band", "edge_clamp")


def same_bits(a, b):
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32 if a.itemsize == 4 else np.uint64),
                              b.view(np.uint32 if b.itemsize == 4 else np.uint64))
    return np.array_equal(a, b)


def batch_molecules(seed, count, hi):
    atoms, probe, spacing, seeds = sesmesh_amd.synth_batch_spec(seed, count, lo=20, hi=hi)
    mols = [(sesmesh_amd.synth_globule(int(sd), int(n)), float(p), float(h))
            for n, p, h, sd in zip(atoms, probe, spacing, seeds)]
    return mols


def check_batch(device, mols, brick=10, oracle_parts=()):
    residents = [sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(x, p)) for x, p, _ in mols]
    batch = sesmesh_amd.mesh_many(residents, [h for _, _, h in mols], brick)
    assert len(batch) == len(mols)
    for i, (res, (xyzr, probe, h)) in enumerate(zip(residents, mols)):
        single = res.mesh(h, brick)
        a, b = batch.arrays(values=True, part=i), single.arrays(values=True)
        for k in KEYS:
            assert same_bits(a[k], b[k]), (i, k)
        for f in INFO:
            assert getattr(batch.parts[i], f) == getattr(single.info, f), (i, f)
        assert list(batch.parts[i].frame) == list(single.info.frame)
        if i in oracle_parts:
            o = oracle.Surface(xyzr, probe, h, brick, cpu_threads())
            om = o.mesh()
            assert batch.parts[i].vertices == o.n_vertices and batch.parts[i].triangles == o.n_triangles
            for k in ("positions", "normals"):
                assert same_bits(a[k], om[k]), (i, k)
            for k in ("atoms", "patches", "triangles"):
                assert np.array_equal(a[k], om[k]), (i, k)
    return batch


def test_batch_of_uneven_molecules_equals_single_runs(device):
    mols = batch_molecules(7, 24, 1500)
    check_batch(device, mols, oracle_parts=(0, 5, 11))


def test_batch_with_tiny_repeated_and_degenerate_molecules(device):
    one = np.array([[0.0, 0.0, 0.0, 1.7]])
    two = np.array([[0.0, 0.0, 0.0, 1.7], [2.5, 0.3, -0.2, 1.52]])
    nested = np.array([[0.0, 0.0, 0.0, 3.0], [0.2, 0.1, 0.0, 1.0]])  # the small atom is contained
    glob = sesmesh_amd.synth_globule(3, 200)
    mols = [(one, 1.4, 0.5), (two, 1.2, 0.25), (glob, 1.4, 0.5), (nested, 1.3, 0.5), (glob, 1.4, 0.5),
            (one, 1.4, 0.2)]
    check_batch(device, mols, oracle_parts=(0, 1, 3))


def test_batch_of_one_is_the_single_run(device):
    xyzr = sesmesh_amd.synth_globule(5, 400)
    check_batch(device, [(xyzr, 1.4, 0.25)])


def test_batch_needs_one_brick_size(device):
    xyzr = sesmesh_amd.synth_globule(5, 100)
    r = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4))
    params = [sesmesh_amd._mesh_params(0.5, 10, "exact", 0), sesmesh_amd._mesh_params(0.5, 8, "exact", 0)]
    with pytest.raises(sesmesh_amd.BackendError):
        sesmesh_amd.Mesh(None, None, 10, "exact", 0, _batch=([r, r], params))
