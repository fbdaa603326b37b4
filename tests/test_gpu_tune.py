"""Launch-shape tuning (smg_device_set_tuning, smg_tune): every tuning gives
the same mesh, byte for byte, as the default one; the autotuner returns a
valid tuning, leaves it on the device and keeps the mesh. The reference's
launch shapes are fixed (crates/sesmesh-gpu/src/hip/resident.rs:675-683),
so its acceptance check (device mesh equal to the CPU mesh,
crates/sesmesh/src/field/gpu.rs:906-962) is what each tuning must keep.
"""
import pytest

import sesmesh_amd
from test_gpu_batch import INFO, KEYS, same_bits

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(sample_first_tier=1),
    dict(sample_first_tier=2, refine_first_tier=2),
    dict(refine_first_tier=1, refine_group_max=16),
    dict(refine_group_max=1),
    dict(refine_group_max=8, uniform_shared=0),
    dict(sample_first_tier=1, refine_first_tier=1, refine_group_max=32, uniform_shared=0),
    dict(scratch_workgroups=16),
    dict(sample_first_tier=2, refine_first_tier=2, scratch_workgroups=4096),
]


@pytest.fixture
def tuned_device(device):
    yield device
    device.tuning = None


def assert_same(a_mesh, b_mesh):
    a, b = a_mesh.arrays(values=True), b_mesh.arrays(values=True)
    for k in KEYS:
        assert same_bits(a[k], b[k]), k
    for f in INFO:
        assert getattr(a_mesh.info, f) == getattr(b_mesh.info, f), f


@pytest.mark.parametrize("case", ["1CRN-0.25", "globule-0.5", "globule-1.0"])
def test_every_tuning_gives_the_same_mesh(tuned_device, structures, case):
    name, h = case.split("-")
    xyzr = sesmesh_amd.synth_globule(4, 2500) if name == "globule" else structures(name)
    res = sesmesh_amd.Resident(tuned_device, sesmesh_amd.Arrangement(xyzr, 1.4))
    tuned_device.tuning = None
    want = res.mesh(float(h), 10)
    for t in TUNINGS:
        tuned_device.tuning = t
        assert tuned_device.tuning == {**sesmesh_amd.DEFAULT_TUNING, **t}
        assert_same(res.mesh(float(h), 10), want)


def test_fast_mode_is_tuning_independent(tuned_device):
    res = sesmesh_amd.Resident(tuned_device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(6, 900), 1.4))
    tuned_device.tuning = None
    want = res.mesh(0.35, 10, refine="fast")
    for t in TUNINGS[2:4]:
        tuned_device.tuning = t
        assert_same(res.mesh(0.35, 10, refine="fast"), want)


def test_autotuner_keeps_the_mesh(tuned_device):
    xyzr = sesmesh_amd.synth_globule(8, 3000)
    res = sesmesh_amd.Resident(tuned_device, sesmesh_amd.Arrangement(xyzr, 1.4))
    tuned_device.tuning = None
    want = res.mesh(0.5, 10)
    out = sesmesh_amd.tune([res], [0.5], repeats=1)
    assert out["brick"] == 10 and out["ms"] > 0
    assert tuned_device.tuning == out["tuning"]
    assert_same(res.mesh(0.5, 10), want)


def test_autotuner_over_brick_sizes(tuned_device):
    xyzr = sesmesh_amd.synth_globule(9, 800)
    res = sesmesh_amd.Resident(tuned_device, sesmesh_amd.Arrangement(xyzr, 1.4))
    out = sesmesh_amd.tune([res], [0.5], repeats=1, bricks=(6, 10, 16))
    assert out["brick"] in (6, 10, 16) and len(out["trials"]) == 3
    assert out["ms"] == min(t["ms"] for t in out["trials"])


def test_invalid_tunings_are_rejected(tuned_device):
    tuned_device.tuning = None
    for bad in (dict(sample_first_tier=4), dict(refine_group_max=12), dict(refine_group_max=128),
                dict(uniform_shared=2), dict(refine_first_tier=-1), dict(scratch_workgroups=8),
                dict(scratch_workgroups=9000)):
        with pytest.raises(sesmesh_amd.BackendError):
            tuned_device.tuning = bad
    assert tuned_device.tuning == sesmesh_amd.DEFAULT_TUNING
