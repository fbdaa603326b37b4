"""Mesh files written on the device against the oracle's writers, byte for
byte: OBJ (precisions, normals, colours, comments, groups), binary PLY
(normals, colours, component ids) and MSMS .vert/.face, of uploaded meshes
and of measured device meshes of real structures, written as the reference's
command line writes them (crates/sesmesh-cli/src/main.rs:935-1005: one
group per component named component_i / cavity_i, MSMS density = vertices /
area) (crates/sesmesh/src/export/obj.rs, ply.rs, msms.rs)."""
import numpy as np
import pytest

import oracle
import sesmesh_amd
from conftest import cpu_threads
from test_oracle_export import octahedron

pytestmark = pytest.mark.gpu


def test_uploaded_mesh_files_equal_oracle(device, tmp_path):
    m = octahedron()
    rgb = np.linspace(-0.2, 1.2, 18).reshape(6, 3).astype(np.float32)
    for comps in (False, True):
        dm = sesmesh_amd.upload_mesh(device, m["positions"], m["triangles"], normals=m["normals"], atoms=m["atoms"],
                                     patches=m["patches"])
        om = dict(m)
        if comps:
            dm.measure()
            _, vc = dm.components()
            om["vertex_component"] = vc
            om["triangles"] = dm.arrays()["triangles"]
        else:
            om.pop("vertex_component")
        for prec, normals, colors, groups in ((4, True, None, ()), (0, False, rgb, ((0, "a"), (5, "b\nc"))),
                                               (9, True, rgb, ((3, "x"), (3, "y"), (99, "z")))):
            kw = dict(precision=prec, normals=normals, comments=["one", "two\rlines"], vertex_colors=colors,
                      groups=groups)
            assert dm.export("obj", **kw) == oracle.write_obj(om, prec, normals, kw["comments"], "ses", colors, groups)
        for normals, colors in ((True, None), (False, rgb)):
            assert dm.export("ply", normals=normals, comments=["c"], vertex_colors=colors) == \
                oracle.write_ply(om, normals, ["c"], colors)
        v, f = oracle.write_msms(om, "oct.pdb", 6, 1.234, 1.4)
        hd = dict(description="oct.pdb", spheres=6, density=1.234, probe_radius=1.4)
        assert dm.export("msms_vert", **hd) == v and dm.export("msms_face", **hd) == f
        dm.write(tmp_path / "o.obj", "obj")
        assert (tmp_path / "o.obj").read_bytes() == oracle.write_obj(om)


def cli_files(dev_or_oracle_mesh, measured, atoms, probe, name):
    comps = measured["comp"]
    groups = [(int(comps["first_triangle"][c]), f"{'cavity' if comps['volume'][c] < 0 else 'component'}_{c}")
              for c in range(len(comps["first_triangle"]))]
    density = len(dev_or_oracle_mesh["positions"]) / max(measured["area"], 1e-12)
    return groups, dict(description=name, spheres=atoms, density=density, probe_radius=probe)


@pytest.mark.parametrize("entry,spacing", [("1CRN", 0.5), ("1UBQ", 0.25)])
def test_structure_files_equal_oracle(device, structures, entry, spacing):
    xyzr = structures(entry)
    o = oracle.Surface(xyzr, 1.4, spacing, 10, cpu_threads())
    om = o.measure_full()
    omesh = o.mesh()
    omesh["triangles"], omesh["vertex_component"] = om["grouped"], om["vertex_component"]
    m = sesmesh_amd.Resident(device, sesmesh_amd.Arrangement(xyzr, 1.4)).mesh(spacing, 10)
    m.measure()
    groups, hd = cli_files(omesh, om, len(xyzr), 1.4, f"{entry}.pdb")
    comments = ["input: " + entry, "probe 1.4 A"]
    assert m.export("obj", comments=comments, groups=groups) == oracle.write_obj(omesh, 4, True, comments, "ses",
                                                                                 None, groups)
    assert m.export("ply", comments=comments) == oracle.write_ply(omesh, True, comments)
    v, f = oracle.write_msms(omesh, **{"description": hd["description"], "spheres": hd["spheres"],
                                       "density": hd["density"], "probe": hd["probe_radius"]})
    assert m.export("msms_vert", **hd) == v and m.export("msms_face", **hd) == f
