"""Batched serving at its stated size: 4096 seeded molecules, atom counts
log-uniform in 20-5000 (3.7 M atoms), probe 1.2-1.4 A, spacing 0.2 / 0.25 /
0.5 A, meshed in one batched pass (smg_mesh_many) and by a replayed plan.
Every part's counts equal its single-molecule run; a spread of parts equals
its single run byte for byte (all arrays, samples included); the smallest
parts equal the oracle's CPU mesh stage."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu

COUNT, SEED = 4096, 4


def test_config4_batch_parts_equal_single_runs(device):
    atoms, probe, spacing, seeds = sesmesh_amd.synth_batch_spec(SEED, COUNT)
    assert atoms.min() >= 20 and atoms.max() <= 5000 and set(np.unique(spacing)) == {0.2, 0.25, 0.5}

    def build(i):
        xyzr = sesmesh_amd.synth_globule(int(seeds[i]), int(atoms[i]))
        return xyzr, sesmesh_amd.Arrangement(xyzr, float(probe[i]), threads=1)

    with ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
        built = list(ex.map(build, range(COUNT)))
    print(f"[config4] {COUNT} molecules, {int(atoms.sum())} atoms resident", flush=True)
    residents = [sesmesh_amd.Resident(device, a) for _, a in built]
    mb = sesmesh_amd.mesh_many(residents, spacing, 10)
    print(f"[config4] batched pass: {mb.times.total_ms:.1f} ms device time, "
          f"{sum(p.vertices for p in mb.parts)} vertices", flush=True)
    check = set(np.linspace(0, COUNT - 1, 16).astype(int).tolist())
    small = [int(i) for i in np.argsort(atoms)[:3]]
    for i in range(COUNT):
        m = residents[i].mesh(float(spacing[i]), 10)
        p = mb.parts[i]
        assert (p.touched_bricks, p.surface_bricks, p.vertices, p.triangles, p.refine_evaluations) == (
            m.info.touched_bricks, m.info.surface_bricks, m.n_vertices, m.n_triangles, m.info.refine_evaluations), i
        if i in check or i in small:
            a, b = mb.arrays(values=True, part=i), m.arrays(values=True)
            for k in a:
                assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (i, k)
        if i in small:
            o = oracle.Surface(built[i][0], float(probe[i]), float(spacing[i]), 10, cpu_threads())
            om = o.mesh()
            a = mb.arrays(part=i)
            assert np.array_equal(a["positions"].view(np.uint32), om["positions"].view(np.uint32))
            assert np.array_equal(a["triangles"], om["triangles"])
        del m
    print("[config4] every part equals its single run", flush=True)
    caps = sesmesh_amd.Plan.caps_for(residents, spacing, mesh=mb, margin=0.05)
    del mb
    plan = sesmesh_amd.Plan(device, caps)
    for run in range(2):
        v = plan.run(residents, spacing)
        if run:
            for i in sorted(check):
                assert v.parts[i].vertices == residents[i].mesh(float(spacing[i]), 10).n_vertices
        del v
    assert plan.stats()["fallbacks"] == 0
