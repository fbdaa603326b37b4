"""Cancellation and progress of device runs (smg_device_set_control), as
the reference's Control (crates/sesmesh/src/control.rs): a set flag stops a
run with Error::Cancelled promptly, on the device too (reference tests
crates/sesmesh/src/field/gpu.rs:964-992, crates/sesmesh/src/surface.rs:
1341-1381, within 100 ms), runs after a cancellation are correct, and the
progress callback covers every stage with (0, total) ... (total, total)."""
import ctypes
import gc
import time
from contextlib import contextmanager
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import sesmesh_amd
from conftest import cpu_threads

pytestmark = pytest.mark.gpu


@contextmanager
def no_gc_pauses():
    """The latencies measured are the library's: earlier tests' device objects
    left in reference cycles are freed first, and Python's cyclic collector
    (whose collections free such objects, each freeing device memory) is
    paused while a latency is measured."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def same(a, b):
    return all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)) for k in a)


@pytest.fixture(scope="module")
def batch(device):
    atoms, probe, spacing, seeds = sesmesh_amd.synth_batch_spec(4, 1024)

    def build(i):
        return sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(int(seeds[i]), int(atoms[i])), float(probe[i]),
                                       threads=1)

    with ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
        arrs = list(ex.map(build, range(len(atoms))))
    return [sesmesh_amd.Resident(device, a) for a in arrs], spacing


def test_cancelled_before_the_run(device):
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(1, 330), 1.4))
    want = res.mesh(0.5, 10).arrays()
    flag = ctypes.c_int32(1)
    device.set_control(cancel=flag)
    try:
        with pytest.raises(sesmesh_amd.BackendError) as e:
            res.mesh(0.5, 10)
        assert e.value.kind == "Cancelled"
        flag.value = 0
        assert same(res.mesh(0.5, 10).arrays(), want)
    finally:
        device.set_control()


@pytest.mark.parametrize("stage,after", [("classify", 0), ("mesh", 0), ("mesh", 1)])
def test_cancellation_of_a_batch_is_prompt(device, batch, stage, after):
    residents, spacing = batch
    t = time.perf_counter()
    full = sesmesh_amd.mesh_many(residents, spacing, 10)
    full_s = time.perf_counter() - t
    parts = [full.arrays(part=i) for i in (0, 500, 1023)]
    del full
    flag = ctypes.c_int32(0)
    at = []

    def progress(st, done, total):
        if st == stage and done >= after and not at:  # once
            at.append(time.perf_counter())
            flag.value = 1

    device.set_control(cancel=flag, progress=progress)
    try:
        with no_gc_pauses(), pytest.raises(sesmesh_amd.BackendError) as e:
            sesmesh_amd.mesh_many(residents, spacing, 10)
        latency = time.perf_counter() - at[0]
        assert e.value.kind == "Cancelled"
        print(f"[control] batch {full_s:.3f} s, cancelled at {stage} {after} after {latency * 1e3:.1f} ms", flush=True)
        assert latency < 0.1 and latency < full_s / 2
        flag.value = 0
        again = sesmesh_amd.mesh_many(residents, spacing, 10)
        assert all(same(again.arrays(part=i), p) for i, p in zip((0, 500, 1023), parts))
    finally:
        device.set_control()


def test_cancellation_of_a_plan_replay_is_prompt(device, batch):
    residents, spacing = batch
    m = sesmesh_amd.mesh_many(residents, spacing, 10)
    want = m.arrays(part=700)
    plan = sesmesh_amd.Plan(device, sesmesh_amd.Plan.caps_for(residents, spacing, mesh=m, margin=0.05))
    del m
    for _ in range(2):  # record the graph, replay it once
        v = plan.run(residents, spacing)
        del v
    t = time.perf_counter()
    v = plan.run(residents, spacing)
    replay_s = time.perf_counter() - t
    del v
    flag = ctypes.c_int32(0)
    at = []

    def cancel_later():  # from another thread, while the replay runs
        time.sleep(replay_s / 4)
        at.append(time.perf_counter())
        flag.value = 1

    device.set_control(cancel=flag)
    try:
        with no_gc_pauses(), ThreadPoolExecutor(max_workers=1) as ex:
            ex.submit(cancel_later)
            with pytest.raises(sesmesh_amd.BackendError) as e:
                plan.run(residents, spacing)
        latency = time.perf_counter() - at[0]
        assert e.value.kind == "Cancelled"
        print(f"[control] replay {replay_s:.3f} s, cancelled after {latency * 1e3:.1f} ms", flush=True)
        assert latency < 0.1 and latency < replay_s / 2
        flag.value = 0
        for _ in range(2):  # the replays after a cancelled one are whole
            v = plan.run(residents, spacing)
            assert same(v.arrays(part=700), want)
            del v
        assert plan.stats()["fallbacks"] == 0
    finally:
        device.set_control()


def test_progress_covers_every_stage(device):
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(3, 800), 1.4))
    seen = []
    device.set_control(progress=lambda st, done, total: seen.append((st, done, total)))
    try:
        m = res.mesh(0.35, 10)
        m.measure()
    finally:
        device.set_control()
    nb = m.n_bricks
    stages = [s for s, _, _ in seen]
    assert stages == sorted(stages, key=["classify", "mesh", "assemble", "measure"].index)
    for st, total in (("classify", 1), ("mesh", nb), ("assemble", 1), ("measure", 1)):
        reps = [(d, t) for s, d, t in seen if s == st]
        assert reps[0] == (0, total) and reps[-1] == (total, total), (st, reps)
        assert all(t == total and 0 <= d <= total for d, t in reps)
        assert [d for d, _ in reps] == sorted(d for d, _ in reps)
