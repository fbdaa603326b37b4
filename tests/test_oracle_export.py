"""The oracle's mesh writers (OBJ, binary PLY, MSMS) pinned on the CPU.

* The reference's literal OBJ snapshot (crates/sesmesh/tests/surface.rs:
  341-355).
* push_fixed (export/fmt.rs:30-54) against Python's exact decimal
  formatting, which like Rust's `{:.p}` rounds the exact binary value half
  to even and keeps the sign of negative values that round to zero: the
  property the reference tests against Rust's formatter (fmt.rs:81-134),
  on its edge-case table and on random f32 bit patterns and coordinates.
* Whole files against independent restatements from the format specs:
  OBJ and MSMS lines built with Python format specs (`{:9.3f}`, `{:7d}`:
  the same fill, alignment and rounding as the reference's Rust specs), PLY
  records built as numpy structured arrays.
"""
import numpy as np

import oracle
from test_oracle_measure import OCT_P, OCT_T

VERSION = "1.3.2"


def test_obj_snapshot():
    m = {"positions": [[0.0, 0.0, 0.0], [1.0, -0.5, 0.0], [0.0, 1.25, -2.71849]], "normals": [[0.0, 0.0, 1.0]] * 3,
         "triangles": [[0, 1, 2]]}
    want = (f"# sesmesh {VERSION}\no ses\nv 0.0000 0.0000 0.0000\nv 1.0000 -0.5000 0.0000\n"
            "v 0.0000 1.2500 -2.7185\nvn 0.0000 0.0000 1.0000\nvn 0.0000 0.0000 1.0000\n"
            "vn 0.0000 0.0000 1.0000\nf 1//1 2//2 3//3\n")
    assert oracle.write_obj(m).decode() == want


def py_fixed(x, p):
    return format(float(np.float32(x)), f".{p}f")


def test_fixed_matches_exact_decimal_formatting():
    edge = np.array([0.0, -0.0, 0.5, 1.5, 2.5, -0.5, 0.25, 0.125, 0.375, 0.03125, -0.00001, 1e-9, 123456.79,
                     -98765.43, np.finfo(np.float32).tiny, 1e15, -3e30, np.finfo(np.float32).max], np.float32)
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    bits = bits[np.isfinite(bits)]
    coords = rng.uniform(-2000, 2000, 20000).astype(np.float32)
    for values in (edge, bits, coords):
        for p in range(10):
            got = [s.decode() for s in oracle.fixed(values, p)]
            assert got == [py_fixed(x, p) for x in values], p
    assert [s.decode() for s in oracle.fixed(np.array([0.5, 1.5, 2.5, -0.0, -0.00001], np.float32), 0)] == \
        ["0", "2", "2", "-0", "-0"]


def octahedron():
    rng = np.random.default_rng(4)
    p = (OCT_P * 1.7 + rng.normal(scale=0.3, size=OCT_P.shape)).astype(np.float32)
    n = OCT_P.astype(np.float32)
    return {"positions": p, "normals": n, "triangles": OCT_T, "atoms": np.arange(6, dtype=np.uint32) * 7,
            "patches": np.array([0, 1, 2, 0, 1, 2], np.uint8), "vertex_component": np.array([0, 0, 1, 1, 0, 1],
                                                                                              np.uint32)}


def test_obj_file_matches_format_specs():
    m = octahedron()
    rgb = np.linspace(-0.2, 1.2, 18).reshape(6, 3).astype(np.float32)
    for prec, normals, colors, groups in ((4, True, None, ()), (2, False, rgb, ((0, "component_0"), (5, "cav\nity_1"))),
                                           (0, True, rgb, ((3, "a"), (3, "b"), (9, "c")))):
        got = oracle.write_obj(m, prec, normals, ["test header", "two\rlines"], "ses", colors, groups).decode()
        lines = [f"# sesmesh {VERSION}", "# test header", "# two lines", "o ses"]
        for i, v in enumerate(m["positions"]):
            s = "v " + " ".join(py_fixed(c, prec) for c in v)
            if colors is not None:
                s += "".join(" " + py_fixed(min(max(c, 0.0), 1.0), 4) for c in colors[i])
            lines.append(s)
        if normals:
            lines += ["vn " + " ".join(py_fixed(c, prec) for c in v) for v in m["normals"]]
        face = lambda t: "f " + " ".join(f"{i + 1}//{i + 1}" if normals else f"{i + 1}" for i in t)
        starts = {}
        for first, name in groups:
            starts[min(first, len(OCT_T))] = name.replace("\n", " ")
        for k, t in enumerate(OCT_T):
            if k in starts:
                lines.append("g " + starts[k])
            lines.append(face(t))
        if len(OCT_T) in starts:
            lines.append("g " + starts[len(OCT_T)])
        assert got == "\n".join(lines) + "\n"


def test_msms_files_match_format_specs():
    m = octahedron()
    vert, face = oracle.write_msms(m, "in\nput.pdb", 6, 2.3456, 1.4)
    head = lambda what, n: (f"# MSMS solvent excluded surface {what} for in put.pdb\n#{what} #sphere density probe_r\n"
                            f"{n:7d} {6:7d} {2.3456:5.2f} {1.4:5.2f}\n")
    ty = {0: 3, 1: 1, 2: 2}
    want = head("vertices", 6) + "".join(
        " ".join(format(float(c), "9.3f") for c in list(m["positions"][v]) + list(m["normals"][v])) +
        f" {m['vertex_component'][v] + 1:7d} {m['atoms'][v] + 1:7d} {ty[int(m['patches'][v])]:2d}\n" for v in range(6))
    assert vert.decode() == want
    rows = []
    for t in OCT_T:
        types = {ty[int(m["patches"][i])] for i in t}
        rows.append(f"{t[0] + 1:6d} {t[1] + 1:6d} {t[2] + 1:6d} {types.pop() if len(types) == 1 else 1:2d} "
                    f"{m['vertex_component'][t[0]] + 1:6d}\n")
    assert face.decode() == head("faces", 8) + "".join(rows)
    # without component ids every vertex and face is in component 1
    m.pop("vertex_component")
    vert, face = oracle.write_msms(m, "x", 6, 1.0, 1.4)
    assert all(line.split()[6] == "1" for line in vert.decode().splitlines()[3:])


def test_ply_file_matches_numpy_records():
    m = octahedron()
    rgb = np.linspace(-0.2, 1.2, 18).reshape(6, 3).astype(np.float32)
    for normals, colors, comps in ((True, None, True), (False, rgb, False), (True, rgb, True)):
        mm = dict(m)
        if not comps:
            mm.pop("vertex_component")
        got = oracle.write_ply(mm, normals, ["c1", "c\n2"], colors)
        fields = [("x", "<f4"), ("y", "<f4"), ("z", "<f4")]
        head = f"ply\nformat binary_little_endian 1.0\ncomment sesmesh {VERSION}\ncomment c1\ncomment c 2\n" \
               f"element vertex 6\nproperty float x\nproperty float y\nproperty float z\n"
        if normals:
            fields += [("nx", "<f4"), ("ny", "<f4"), ("nz", "<f4")]
            head += "property float nx\nproperty float ny\nproperty float nz\n"
        if colors is not None:
            fields += [("r", "u1"), ("g", "u1"), ("b", "u1")]
            head += "property uchar red\nproperty uchar green\nproperty uchar blue\n"
        fields += [("atom", "<u4"), ("patch", "u1")]
        head += "property uint atom\nproperty uchar patch\n"
        if comps:
            fields += [("comp", "<u4")]
            head += "property uint component\n"
        head += "element face 8\nproperty list uchar uint vertex_indices\nend_header\n"
        v = np.zeros(6, np.dtype(fields))
        v["x"], v["y"], v["z"] = m["positions"].T
        if normals:
            v["nx"], v["ny"], v["nz"] = m["normals"].T
        if colors is not None:
            c = np.floor(np.clip(colors, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
            v["r"], v["g"], v["b"] = c.T
        v["atom"], v["patch"] = m["atoms"], m["patches"]
        if comps:
            v["comp"] = m["vertex_component"]
        f = np.zeros(8, np.dtype([("n", "u1"), ("i", "<u4", 3)]))
        f["n"], f["i"] = 3, OCT_T
        assert got == head.encode() + v.tobytes() + f.tobytes()
