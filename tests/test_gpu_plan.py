"""Capture-ready plans (smg_plan_*): the whole mesh stage replayed as a
hipGraph over buffers sized once by capacities, with persistent workgroups
taking work from device-side queues. Every run's meshes equal, byte for
byte, smg_mesh_many's for the same molecules (and through them the CPU back
end); runs beyond the capacities fall back to the exact path and grow the
plan. The reference has no such loop: it is the serving loop around
`fold_resident` (crates/sesmesh/src/field/gpu.rs:532-712).

Each graph test replays several times with eager library work and
synchronisations between the replays, and checks every replay.
"""
import numpy as np
import pytest

import sesmesh_amd
from test_gpu_batch import INFO, KEYS, batch_molecules, same_bits

pytestmark = pytest.mark.gpu


def residents_of(device, mols):
    return [sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(x, p)) for x, p, _ in mols]


def assert_same_meshes(got, want, n):
    assert len(got) == len(want) == n
    for i in range(n):
        a, b = got.arrays(values=True, part=i), want.arrays(values=True, part=i)
        for k in KEYS:
            assert same_bits(a[k], b[k]), (i, k)
        for f in INFO:
            assert getattr(got.parts[i], f) == getattr(want.parts[i], f), (i, f)


def test_plan_replays_equal_the_exact_path(device):
    mols = batch_molecules(11, 12, 1200)
    res = residents_of(device, mols)
    hs = [h for _, _, h in mols]
    want = sesmesh_amd.mesh_many(res, hs)
    plan = sesmesh_amd.Plan(device, sesmesh_amd.Plan.caps_for(res, hs, mesh=want, margin=0.25))
    for run in range(4):
        got = plan.run(res, hs)
        assert_same_meshes(got, want, len(mols))
        del got
        # Eager work and a synchronisation between replays.
        res[0].mesh(hs[0], 10)
    st = plan.stats()
    assert st["runs"] == 4 and st["captures"] == 1 and st["replays"] == 3 and st["fallbacks"] == 0


def test_plan_graph_serves_other_molecules(device):
    a = batch_molecules(21, 10, 900)
    b = batch_molecules(22, 8, 900)
    ra, rb = residents_of(device, a), residents_of(device, b)
    ha, hb = [h for *_, h in a], [h for *_, h in b]
    want_a = sesmesh_amd.mesh_many(ra, ha)
    caps = sesmesh_amd.Plan.caps_for(ra, ha, mesh=want_a, margin=3.0)
    caps.locals, caps.spheres, caps.box_bricks = 4 * caps.locals, 4 * caps.spheres, 4 * caps.box_bricks
    plan = sesmesh_amd.Plan(device, caps)
    assert_same_meshes(plan.run(ra, ha), want_a, len(a))
    want_b = sesmesh_amd.mesh_many(rb, hb)
    for _ in range(2):
        got = plan.run(rb, hb)  # replays of the graph captured for the first batch
        assert_same_meshes(got, want_b, len(b))
        del got
    got = plan.run(ra, ha)
    assert_same_meshes(got, want_a, len(a))
    del got
    st = plan.stats()
    assert st["fallbacks"] == 0 and st["captures"] == 1 and st["replays"] == 3


def test_plan_overflow_falls_back_and_grows(device):
    small = batch_molecules(31, 4, 200)
    big = batch_molecules(32, 6, 1500)
    rs, rb = residents_of(device, small), residents_of(device, big)
    hs, hb = [h for *_, h in small], [h for *_, h in big]
    caps = sesmesh_amd.Plan.caps_for(rs, hs, mesh=sesmesh_amd.mesh_many(rs, hs), margin=0.0)
    caps.molecules, caps.locals, caps.spheres, caps.box_bricks = 8, 10**6, 10**6, 10**6
    plan = sesmesh_amd.Plan(device, caps)
    want = sesmesh_amd.mesh_many(rb, hb)
    got = plan.run(rb, hb)  # surface bricks beyond the capacity: found on the device
    assert_same_meshes(got, want, len(big))
    st = plan.stats()
    assert st["fallbacks"] == 1 and st["last_overflow"] & 1
    del got
    for _ in range(2):  # the plan grew: the next runs fit and replay
        got = plan.run(rb, hb)
        assert_same_meshes(got, want, len(big))
        del got
    st = plan.stats()
    assert st["fallbacks"] == 1 and st["replays"] == 1 and st["captures"] == 1


def test_plan_table_overflow_falls_back(device):
    mols = batch_molecules(41, 5, 300)
    res = residents_of(device, mols)
    hs = [h for *_, h in mols]
    want = sesmesh_amd.mesh_many(res, hs)
    caps = sesmesh_amd.Plan.caps_for(res[:2], hs[:2], mesh=sesmesh_amd.mesh_many(res[:2], hs[:2]))
    plan = sesmesh_amd.Plan(device, caps)
    got = plan.run(res, hs)  # more molecules than table rows: found on the host
    assert_same_meshes(got, want, len(mols))
    assert plan.stats()["last_overflow"] & 8


def test_plan_view_must_be_released_before_the_next_run(device):
    mols = batch_molecules(51, 3, 300)
    res = residents_of(device, mols)
    hs = [h for *_, h in mols]
    plan = sesmesh_amd.Plan(device, sesmesh_amd.Plan.caps_for(res, hs, mesh=sesmesh_amd.mesh_many(res, hs)))
    view = plan.run(res, hs)
    with pytest.raises(sesmesh_amd.BackendError):
        plan.run(res, hs)
    del view
    plan.run(res, hs)


def test_plan_fast_mode_equals_fast_batch(device):
    mols = batch_molecules(61, 5, 800)
    res = residents_of(device, mols)
    hs = [h for *_, h in mols]
    want = sesmesh_amd.mesh_many(res, hs, refine="fast")
    plan = sesmesh_amd.Plan(device, sesmesh_amd.Plan.caps_for(res, hs, refine="fast", mesh=want))
    for _ in range(2):
        got = plan.run(res, hs)
        assert_same_meshes(got, want, len(mols))
        del got
