"""smg_mesh_into, the mesh stage that allocates nothing, against
smg_mesh_new: two calls in a row on one device (the second on a smaller
molecule, into the same output buffers) write meshes byte for byte equal
to smg_mesh_new's; capacities and scratch too small are rejected."""
import numpy as np
import pytest

import sesmesh_amd

pytestmark = pytest.mark.gpu


def device_buffers(torch, nv, nt):
    return {"positions": torch.empty(3 * nv, dtype=torch.float32, device="cuda"),
            "normals": torch.empty(3 * nv, dtype=torch.float32, device="cuda"),
            "atoms": torch.empty(nv, dtype=torch.int32, device="cuda"),
            "patches": torch.empty(nv, dtype=torch.uint8, device="cuda"),
            "triangles": torch.empty(3 * nt, dtype=torch.int32, device="cuda")}


def run_into(res, spacing, bufs, nv_cap, nt_cap, brick=10):
    return sesmesh_amd.mesh_into(res, spacing, bufs["positions"].data_ptr(), bufs["normals"].data_ptr(),
                                 bufs["atoms"].data_ptr(), bufs["patches"].data_ptr(), bufs["triangles"].data_ptr(),
                                 nv_cap, nt_cap, brick=brick)


def test_mesh_into_equals_mesh_new_twice():
    torch = pytest.importorskip("torch")
    dev = sesmesh_amd.Device(0)
    cases = [(sesmesh_amd.synth_globule(5, 1200), 0.35), (sesmesh_amd.synth_globule(1, 330), 0.5)]
    bufs = None
    for k, (xyzr, spacing) in enumerate(cases):
        print(f"[no-alloc] case {k}: {len(xyzr)} atoms at {spacing} A", flush=True)
        res = sesmesh_amd.Resident(dev, sesmesh_amd.Arrangement(xyzr, 1.4))
        m = res.mesh(spacing, 10)
        ref = m.arrays()
        nv, nt = m.n_vertices, m.n_triangles
        if bufs is None:  # sized for the first (larger) molecule, reused for the second
            cap_v, cap_t = nv + 1000, nt + 2000
            bufs = device_buffers(torch, cap_v, cap_t)
        info = run_into(res, spacing, bufs, cap_v, cap_t)
        torch.cuda.synchronize()
        assert (info.vertices, info.triangles, info.surface_bricks) == (nv, nt, m.n_bricks)
        assert info.refine_evaluations == m.info.refine_evaluations
        got = {k2: v.cpu().numpy() for k2, v in bufs.items()}
        assert np.array_equal(got["positions"][:3 * nv].view(np.uint32), ref["positions"].reshape(-1).view(np.uint32))
        assert np.array_equal(got["normals"][:3 * nv].view(np.uint32), ref["normals"].reshape(-1).view(np.uint32))
        assert np.array_equal(got["atoms"][:nv].view(np.uint32), ref["atoms"])
        assert np.array_equal(got["patches"][:nv], ref["patches"])
        assert np.array_equal(got["triangles"][:3 * nt].view(np.uint32), ref["triangles"].reshape(-1))
        del m


def test_mesh_into_rejects_small_capacities_and_large_molecules():
    torch = pytest.importorskip("torch")
    dev = sesmesh_amd.Device(0)
    res = sesmesh_amd.Resident(dev, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(1, 330), 1.4))
    m = res.mesh(0.5, 10)
    nv, nt = m.n_vertices, m.n_triangles
    bufs = device_buffers(torch, nv, nt)
    with pytest.raises(sesmesh_amd.BackendError) as e:
        run_into(res, 0.5, bufs, nv - 1, nt)
    assert e.value.kind == "Length" and "vertex capacity" in str(e.value)
    with pytest.raises(sesmesh_amd.BackendError) as e:
        run_into(res, 0.5, bufs, nv, nt - 1)
    assert e.value.kind == "Length" and "triangle capacity" in str(e.value)
    # The call after a rejected one still meshes correctly.
    info = run_into(res, 0.5, bufs, nv, nt)
    assert (info.vertices, info.triangles) == (nv, nt)
    got = bufs["triangles"].cpu().numpy().view(np.uint32)
    assert np.array_equal(got, m.arrays()["triangles"].reshape(-1))
    big = sesmesh_amd.Resident(dev, sesmesh_amd.Arrangement(sesmesh_amd.synth_rings(seed=2), 1.4))
    with pytest.raises(sesmesh_amd.BackendError) as e:
        run_into(big, 0.25, bufs, nv, nt)
    assert e.value.kind == "Length" and "scratch" in str(e.value)
