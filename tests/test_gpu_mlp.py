"""Learned per-brick sizing (smg_model_*) on the GPU, checked against the
CPU oracle oracle/mlp.py: the BF16 matrix-core forward pass and the
hand-written backward pass, the features and labels collected from the
exact pipeline, training, and what the trained model gains. The model may
only change where a sub-block is tried first, so every mesh made with it
must equal the mesh made without it, byte for byte, including under
injected mispredictions. The reference has no such model; its lists are
sized by a count pass.
"""
import numpy as np
import pytest

import sesmesh_amd
from oracle import mlp
from test_gpu_batch import KEYS, same_bits

pytestmark = pytest.mark.gpu


def inputs(rng, n):
    return np.sqrt(rng.integers(0, 40, (n, 64)).astype(np.float32))


def test_forward_matches_oracle(device):
    model = sesmesh_amd.Model(device, seed=3)
    rng = np.random.default_rng(1)
    P = model.params
    P[mlp.P_B1:mlp.P_W2] = rng.normal(0, 0.1, 128)
    P[mlp.P_B2:mlp.P_W3] = rng.normal(0, 0.1, 128)
    P[mlp.P_B3:] = [0.3, -0.2]
    model.params = P
    X = inputs(rng, 333)  # not a multiple of the 64-row tile
    y = model.forward(X)
    ref, _ = mlp.forward(P, X, bf16_ops=True)
    assert np.all(np.abs(y - ref) <= 1e-2 * (1.0 + np.abs(ref))), np.abs(y - ref).max()
    # Without the BF16 roundings the oracle is measurably further away: the
    # device rounds where the oracle says it does.
    exact, _ = mlp.forward(P, X, bf16_ops=False)
    assert np.abs(y - ref).mean() < np.abs(y - exact).mean()


def test_gradients_match_oracle(device):
    model = sesmesh_amd.Model(device, seed=5)
    rng = np.random.default_rng(2)
    P = model.params
    n = 300
    X = inputs(rng, n)
    occ = rng.integers(0, 2, n).astype(np.uint8)
    tgt = rng.uniform(0, 6, n).astype(np.float32)
    g, loss = model.gradients(X, occ, tgt)
    (bce, mse), ref = mlp.loss_and_grads(P, X, occ, tgt, bf16_ops=True)
    assert abs(loss[0] - bce) <= 1e-3 * bce and abs(loss[1] - mse) <= 1e-3 * mse + 1e-6
    for name, lo, hi in [("W1", mlp.P_W1, mlp.P_B1), ("b1", mlp.P_B1, mlp.P_W2), ("W2", mlp.P_W2, mlp.P_B2),
                         ("b2", mlp.P_B2, mlp.P_W3), ("W3", mlp.P_W3, mlp.P_B3), ("b3", mlp.P_B3, mlp.N_PARAMS)]:
        d = np.linalg.norm(g[lo:hi] - ref[lo:hi]) / np.linalg.norm(ref[lo:hi])
        assert d < 2e-2, (name, d)


def test_features_and_labels_follow_the_exact_pipeline(device, structures):
    xyzr = structures("1CRN")
    arr = sesmesh_amd.Arrangement(xyzr, 1.4)
    res = sesmesh_amd.Resident(device, arr)
    model = sesmesh_amd.Model(device)
    model.collect([res], [0.5])
    x, occ, t = model.dataset()
    mesh = res.mesh(0.5, 10)
    keys = mesh.arrays()["bricks"]
    assert len(x) == mesh.info.touched_bricks and occ.sum() == len(keys)
    a = arr.arrays()
    want = mlp.features(a["centers"], 1.4, a["max_expanded"], 10, 0.5, keys)
    assert np.array_equal(x[occ == 1].view(np.uint32), want.view(np.uint32))
    assert np.all(t[occ == 0] == 0.0) and np.all(t[occ == 1] >= 0.0) and (t[occ == 1] > 0).mean() > 0.9


@pytest.fixture(scope="module")
def trained(device):
    model = sesmesh_amd.Model(device, seed=7)
    mols = [(sesmesh_amd.synth_globule(100 + i, n), h)
            for i, (n, h) in enumerate([(2000, 0.35), (6000, 0.4), (15000, 0.5), (30000, 0.6), (4000, 0.75),
                                        (20000, 0.8), (12000, 1.0), (40000, 1.0), (8000, 0.9), (25000, 0.7),
                                        (700, 0.5), (50000, 0.45)])]
    for xyzr, h in mols:
        model.collect([sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4))], [h])
    before = model.train(epochs=0, batch=4096, lr=1e-3)
    after = model.train(epochs=40, batch=4096, lr=2e-3)
    return model, before, after


def test_training_lowers_the_loss(trained):
    model, before, after = trained
    assert after["samples"] > 30_000
    assert after["occupancy_loss"] < 0.6 * before["occupancy_loss"]
    assert after["occupancy_accuracy"] > 0.85
    assert after["load_rmse"] < 0.5 * before["load_rmse"]


def test_sizing_keeps_the_mesh_and_cuts_retries(device, trained):
    # What the trained network gains: placed by its predictions (every tier
    # allowed), most sub-blocks that do not fit the first lists skip them,
    # so the wasted gathers (retries) drop; the default, occupancy-aware
    # placement only skips for loads well past the first lists and never
    # retries more. The mesh is the same in every case.
    model, _, _ = trained
    model.use()
    try:
        res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(999, 30000), 1.4))
        for h in (0.75, 1.0):
            plain = res.mesh(h, 10)
            a = plain.arrays(values=True)
            r0 = plain.info.sample_retries + plain.info.refine_retries
            for policy, bound in (((1.0, 2), 0.5), ((0.3, 1), 1.0)):
                model.policy(*policy)
                sized = res.mesh(h, 10, sizing=1)
                b = sized.arrays(values=True)
                for k in KEYS:
                    assert same_bits(a[k], b[k]), (h, policy, k)
                r1 = sized.info.sample_retries + sized.info.refine_retries
                assert r0 > 200 and r1 <= bound * r0, (h, policy, r0, r1)
        model.policy()
    finally:
        sesmesh_amd.lib().smg_device_set_model(device._h, None)


@pytest.mark.parametrize("bias", [30.0, -30.0])
def test_mispredicted_sizing_falls_back_to_the_exact_lists(device, bias):
    # W3 = 0, b3[1] = +-30: every sub-block predicted for the largest LDS
    # lists or for the smallest; the exact gathers take over.
    model = sesmesh_amd.Model(device, seed=9)
    model.policy(1.0, 2)
    P = model.params
    P[mlp.P_W3:mlp.P_B3] = 0.0
    P[mlp.P_B3 + 1] = bias
    model.params = P
    model.use()
    try:
        res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(31, 3000), 1.4))
        for h in (0.5, 1.0):
            plain, sized = res.mesh(h, 10), res.mesh(h, 10, sizing=1)
            a, b = plain.arrays(values=True), sized.arrays(values=True)
            for k in KEYS:
                assert same_bits(a[k], b[k]), (h, k)
            if bias < 0:
                assert sized.info.sample_retries == plain.info.sample_retries
                assert sized.info.refine_retries == plain.info.refine_retries
            else:  # only sub-blocks beyond the largest LDS lists retry
                assert sized.info.sample_retries <= plain.info.sample_retries
                assert sized.info.refine_retries <= plain.info.refine_retries
    finally:
        sesmesh_amd.lib().smg_device_set_model(device._h, None)


def test_sizing_needs_a_model(device):
    res = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(3, 100), 1.4))
    sesmesh_amd.lib().smg_device_set_model(device._h, None)
    with pytest.raises(sesmesh_amd.BackendError):
        res.mesh(0.5, 10, sizing=1)


def test_predicted_capacities_size_a_plan(device, trained):
    """A plan sized by the model (smg_plan_caps_predict) for molecules no
    run has seen meshes them as smg_mesh_new does; with capacities far too
    low, the plan falls back to the exact path and grows, with the same
    mesh."""
    model, _, _ = trained
    res = [sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(sesmesh_amd.synth_globule(2000 + i, n), 1.4))
           for i, n in enumerate((9000, 2500))]
    hs = [0.35, 0.5]
    caps, est = sesmesh_amd.Plan.caps_predict(res, hs, model, margin=0.15)
    want = [r.mesh(h, 10) for r, h in zip(res, hs)]
    nb = sum(m.n_bricks for m in want)
    nv = sum(m.n_vertices for m in want)
    assert est["touched_bricks"] == sum(m.info.touched_bricks for m in want)
    assert abs(est["surface_bricks"] - nb) < 0.15 * nb and abs(est["vertices"] - nv) < 0.25 * nv, (est, nb, nv)
    for expect_fallback in (False, True):
        if expect_fallback:
            low, _ = sesmesh_amd.Plan.caps_predict(res, hs, model, margin=0.15)
            low.vertices = 10  # far too few
            plan = sesmesh_amd.Plan(device, low)
        else:
            plan = sesmesh_amd.Plan(device, caps)
        for run in range(2):
            v = plan.run(res, hs)
            for i, m in enumerate(want):
                a, b = v.arrays(values=True, part=i), m.arrays(values=True)
                for k in KEYS:
                    assert same_bits(a[k], b[k]), (run, i, k)
            del v
        st = plan.stats()
        if expect_fallback:
            assert st["fallbacks"] == 1 and st["last_overflow"] & 2
        del plan
