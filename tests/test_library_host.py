"""The C-ABI library on the CPU: it loads, exports every function the
header declares, and its host-side parts (arrangement, Marching Cubes table)
equal the oracle's bit for bit. No GPU call is made here."""
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import ROOT


def header_functions():
    text = (ROOT / "include" / "sesmesh_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = sesmesh_amd.lib()
    declared = header_functions()
    assert len(declared) >= 14
    assert set(declared) == set(sesmesh_amd.ABI_FUNCTIONS)
    for name in declared:
        assert hasattr(lib, name), name


def test_mc_table_matches_reference(golden):
    off, tris = sesmesh_amd.mc_table()
    data = off.astype("<u2").tobytes() + tris.astype(np.uint8).tobytes()
    assert hashlib.sha256(data).hexdigest() == golden["mc_table"]["sha256"]


def _compare_arrangements(xyzr, probe):
    o = oracle.Surface(xyzr, probe, 0.5, 10, stage=1)
    oa = o.arrangement()
    pa = sesmesh_amd.Arrangement(xyzr, probe, threads=4).arrays()
    assert np.array_equal(pa["orig"], oa["orig"])
    assert np.array_equal(pa["centers"], oa["centers"])
    assert np.array_equal(pa["expanded"], oa["expanded"])
    assert np.array_equal(pa["radii"], oa["radii"])
    assert np.array_equal(pa["status"], oa["status"])
    for k in ("neighbors", "circle_blockers"):
        assert np.array_equal(pa[k][0], oa[k][0]), k
        assert np.array_equal(pa[k][1], oa[k][1]), k
    assert np.array_equal(pa["circles_ij"], oa["circles_ij"])
    assert np.array_equal(pa["circles"], oa["circles"])  # bitwise f64
    assert np.array_equal(pa["vertices"], oa["vertices"])
    assert np.array_equal(pa["vertex_atoms"], oa["vertex_atoms"])
    assert pa["max_expanded"] == o.max_expanded and pa["rho_max"] == o.rho_max
    return oa


@pytest.mark.parametrize("entry", ["1CRN", "1EJG", "2PTN"])
def test_arrangement_equals_oracle_on_structures(structures, entry):
    oa = _compare_arrangements(structures(entry), 1.4)
    assert len(oa["circles"]) > 100


def test_arrangement_equals_oracle_on_degenerate_sets():
    rng = np.random.default_rng(5)
    sets = []
    # cubic lattice (many exactly equal distances) and a collinear chain
    g = np.arange(3) * 3.0
    lat = np.array([[x, y, z, 1.7] for x in g for y in g for z in g])
    sets.append((lat, 1.4))
    sets.append((np.array([[2.0 * i, 0, 0, 1.5] for i in range(8)]), 1.0))
    # duplicates, a nested sphere and zero radii
    box = np.column_stack([rng.uniform(0, 10, (30, 3)), rng.uniform(1.2, 2.0, 30)])
    box = np.vstack([box, box[::4], [box[0] + [0.1, 0, 0, -0.5]], [[1, 1, 1, 0.0]]])
    sets.append((box, 1.4))
    for xyzr, probe in sets:
        _compare_arrangements(xyzr, probe)


def test_synthetic_globule_is_deterministic():
    a = sesmesh_amd.synth_globule(1, 330)
    b = sesmesh_amd.synth_globule(1, 330)
    assert np.array_equal(a, b) and a.shape == (330, 4)
    assert np.all(np.linalg.norm(a[:, :3], axis=1) <= 14.0 + 1e-9)
    d = np.linalg.norm(a[:, None, :3] - a[None, :, :3], axis=2) + np.eye(330) * 9
    assert d.min() >= 1.25 - 1e-12
    assert set(np.unique(a[:, 3])) <= {1.70, 1.55, 1.52, 1.80}


def test_invalid_input_raises():
    with pytest.raises(sesmesh_amd.BackendError) as e:
        sesmesh_amd.Arrangement(np.array([[0, 0, np.nan, 1.7]]), 1.4)
    assert e.value.kind == "InvalidInput"
    with pytest.raises(sesmesh_amd.BackendError):
        sesmesh_amd.Arrangement(np.array([[0, 0, 0, -1.0]]), 1.4)
