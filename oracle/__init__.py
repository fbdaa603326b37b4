"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a scalar C++ restatement of the reference's CPU mesh stage
(see oracle.hpp). Only the tests, ``__graft_entry__.smoke`` and the CPU
baseline leg of ``bench.py`` use it, as the checker; the product package
``sesmesh_amd`` never imports it.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_c_p = ctypes.c_void_p
_lib = None


def build(force=False):
    """Compiles liboracle.so with the Makefile beside this file."""
    if force or not LIB.exists():
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.orc_surface_new.restype = _c_p
        L.orc_surface_new.argtypes = [_c_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, ctypes.c_size_t,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.orc_surface_free.argtypes = [_c_p]
        L.orc_field_sample.restype = ctypes.c_double
        L.orc_time_mesh_sample.restype = ctypes.c_double
        L.orc_time_mesh_sample.argtypes = [_c_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                           _c_p]
        L.orc_count_ops.restype = None
        L.orc_count_ops.argtypes = [_c_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _c_p]
        L.orc_csr.restype = ctypes.c_size_t
        for name in ("orc_surface_counts", "orc_surface_params", "orc_surface_times", "orc_surface_bricks",
                     "orc_surface_mesh", "orc_surface_measure", "orc_brick_values", "orc_sub_states",
                     "orc_uniform_sign", "orc_arrangement", "orc_circles", "orc_sas_vertices"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_c_p) if a is not None else None


KINDS = {"ses": 0, "sas": 1, "vdw": 2}


class Surface:
    """The oracle's pipeline on one sphere set (`stage` 1: field and engine,
    2: plus brick classification, 3: plus meshing and assembly) for surface
    kind `kind` ("ses", "sas" or "vdw"; the vdW kind takes probe 0)."""

    def __init__(self, xyzr, probe=1.4, spacing=0.5, brick=10, threads=None, stage=3, kind="ses"):
        L = lib()
        a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
        self.xyzr = a
        if threads is None:
            threads = min(8, os.cpu_count() or 1)
        err = ctypes.create_string_buffer(512)
        h = L.orc_surface_new(_ptr(a), len(a), probe, spacing, brick, threads, stage, KINDS[kind], err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        self.probe, self.spacing, self.brick, self.kind = probe, spacing, brick, kind
        c = (ctypes.c_int64 * 14)()
        L.orc_surface_counts(_c_p(h), c)
        (self.touched, self.n_bricks, self.n_vertices, self.n_triangles, self.evals, self.n_spheres,
         self.n_circles, self.n_sas_vertices, self.n_contained, self.n_exposed, f32_ok, fx, fy, fz) = list(c)
        self.f32_ok = bool(f32_ok)
        self.frame = (fx, fy, fz)
        p = (ctypes.c_double * 6)()
        L.orc_surface_params(_c_p(h), p)
        self.band, self.coord_max, self.edge_clamp, self.rho_max, self.max_expanded, self.extent = list(p)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_surface_free(_c_p(h))
            self._h = None

    @property
    def handle(self):
        return _c_p(self._h)

    def bricks(self):
        out = np.zeros((self.n_bricks, 3), dtype=np.int32)
        lib().orc_surface_bricks(self.handle, _ptr(out))
        return out

    def mesh(self):
        nv, nt = self.n_vertices, self.n_triangles
        pos = np.zeros((nv, 3), np.float32)
        nrm = np.zeros((nv, 3), np.float32)
        atoms = np.zeros(nv, np.uint32)
        patches = np.zeros(nv, np.uint8)
        tris = np.zeros((nt, 3), np.uint32)
        t = np.zeros(nv, np.float64)
        ev = np.zeros(nv, np.uint32)
        lib().orc_surface_mesh(self.handle, _ptr(pos), _ptr(nrm), _ptr(atoms), _ptr(patches), _ptr(tris), _ptr(t),
                               _ptr(ev))
        return {"positions": pos, "normals": nrm, "atoms": atoms, "patches": patches, "triangles": tris, "t": t,
                "evals": ev}

    def measure(self):
        m = (ctypes.c_double * 4)()
        lib().orc_surface_measure(self.handle, m)
        return {"components": int(m[0]), "cavities": int(m[1]), "area": m[2], "volume": m[3]}

    def measure_full(self):
        """The measure stage of this surface's mesh: validation, components
        (with Euler characteristics and first triangles), vertex components,
        triangles grouped by component, and the stats."""
        L = lib()
        L.orc_surface_measure_new.restype = _c_p
        L.orc_measure_free.argtypes = [_c_p]
        h = L.orc_surface_measure_new(self.handle)
        try:
            out = _read_measure(h)
            vc = np.zeros(self.n_vertices, np.uint32)
            g = np.zeros((self.n_triangles, 3), np.uint32)
            L.orc_measure_arrays(_c_p(h), _ptr(vc), _ptr(g))
            out["vertex_component"], out["grouped"] = vc, g
        finally:
            L.orc_measure_free(h)
        return out

    def times(self):
        m = (ctypes.c_double * 4)()
        lib().orc_surface_times(self.handle, m)
        return dict(zip(("geometry", "classify", "mesh", "assemble"), list(m)))

    def brick_values(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 3)
        side = self.brick + 1
        out = np.zeros((len(keys), side ** 3), np.float32)
        lib().orc_brick_values(self.handle, _ptr(keys), ctypes.c_size_t(len(keys)), _ptr(out))
        return out

    def sub_states(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 3)
        out = np.zeros((len(keys), 8), np.uint8)
        lib().orc_sub_states(self.handle, _ptr(keys), ctypes.c_size_t(len(keys)), _ptr(out))
        return out

    def uniform_sign(self, centers, rho):
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(c), np.uint8)
        lib().orc_uniform_sign(self.handle, _ptr(c), ctypes.c_size_t(len(c)), ctypes.c_double(rho), _ptr(out))
        return out

    def sample(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        closest = np.zeros(3, np.float64)
        feat = np.zeros(2, np.uint32)
        v = lib().orc_field_sample(self.handle, _ptr(x), _ptr(closest), _ptr(feat))
        return v, closest, (int(feat[0]), int(feat[1]))

    def arrangement(self):
        n = self.n_spheres
        c = np.zeros((n, 3)); e = np.zeros(n); r = np.zeros(n)
        st = np.zeros(n, np.uint8); orig = np.zeros(n, np.uint32)
        lib().orc_arrangement(self.handle, _ptr(c), _ptr(e), _ptr(r), _ptr(st), _ptr(orig))
        out = {"centers": c, "expanded": e, "radii": r, "status": st, "orig": orig}
        for which, name in ((0, "neighbors"), (1, "circle_blockers")):
            rows = n if which == 0 else self.n_circles
            m = lib().orc_csr(self.handle, which, None, None)
            off = np.zeros(rows + 1, np.uint32); items = np.zeros(m, np.uint32)
            lib().orc_csr(self.handle, which, _ptr(off), _ptr(items))
            out[name] = (off, items)
        ij = np.zeros((self.n_circles, 2), np.uint32); geo = np.zeros((self.n_circles, 10))
        lib().orc_circles(self.handle, _ptr(ij), _ptr(geo))
        out["circles_ij"], out["circles"] = ij, geo
        vp = np.zeros((self.n_sas_vertices, 3)); va = np.zeros((self.n_sas_vertices, 3), np.uint32)
        lib().orc_sas_vertices(self.handle, _ptr(vp), _ptr(va))
        out["vertices"], out["vertex_atoms"] = vp, va
        return out

    def time_mesh_sample(self, first, count, stride, threads):
        """Wall seconds to sample and mesh bricks first + k*stride (k < count)
        with `threads` threads, and the vertices made."""
        v = ctypes.c_uint64()
        s = lib().orc_time_mesh_sample(self.handle, first, count, stride, threads, ctypes.byref(v))
        return s, v.value

    def count_point(self, point, rho=0.01):
        """Algorithmic op counts at one point: the f32 lane kernel over the
        sampling lists of the ball (point, rho) and the f64 eval_local over
        its refinement lists: (f32 ops, f64 ops)."""
        out = (ctypes.c_uint64 * 2)()
        lib().orc_count_point(self.handle, (ctypes.c_double * 3)(*[float(v) for v in point]), ctypes.c_double(rho),
                              out)
        return int(out[0]), int(out[1])

    def count_ops(self, first, count, stride):
        """Algorithmic op counts over bricks first + k*stride: f32 ops and
        evaluated samples of the lane kernel, f64 ops and evaluations of the
        refinement, vertices."""
        out = (ctypes.c_uint64 * 5)()
        lib().orc_count_ops(self.handle, first, count, stride, out)
        return dict(zip(("f32_ops", "samples", "f64_ops", "evals", "vertices"), list(out)))


def _read_measure(h):
    L = lib()
    val = (ctypes.c_uint64 * 7)()
    st = (ctypes.c_double * 4)()
    cnt = (ctypes.c_int64 * 2)()
    L.orc_measure_get(_c_p(h), val, st, cnt)
    names = ("boundary_edges", "non_manifold_edges", "misoriented_edges", "degenerate_triangles",
             "duplicate_triangles", "invalid_indices", "bad_vertices")
    out = {"validation": dict(zip(names, [int(x) for x in val])), "area": st[0], "volume": st[1],
           "components": int(st[2]), "cavities": int(st[3]), "euler": int(cnt[0]), "indices_valid": bool(cnt[1])}
    nc = out["components"]
    vt = np.zeros((nc, 2), np.uint64)
    av = np.zeros((nc, 2), np.float64)
    eu = np.zeros(nc, np.int64)
    first = np.zeros(nc, np.uint64)
    L.orc_measure_components(_c_p(h), _ptr(vt), _ptr(av), _ptr(eu), _ptr(first))
    out["comp"] = {"vertices": vt[:, 0], "triangles": vt[:, 1], "area": av[:, 0], "volume": av[:, 1], "euler": eu,
                   "first_triangle": first}
    return out


def measure(positions, triangles, normals=None):
    """The measure stage (validation, components, grouping; reference
    mesh/mod.rs, post.rs) of a mesh given as arrays."""
    L = lib()
    L.orc_measure_new.restype = _c_p
    L.orc_measure_new.argtypes = [_c_p, ctypes.c_size_t, _c_p, ctypes.c_size_t, _c_p, ctypes.c_size_t]
    L.orc_measure_free.argtypes = [_c_p]
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    nr = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    h = L.orc_measure_new(_ptr(p), len(p), _ptr(nr), 0 if nr is None else len(nr), _ptr(t), len(t))
    try:
        out = _read_measure(h)
        if out["indices_valid"]:
            vc = np.zeros(len(p), np.uint32)
            g = np.zeros((len(t), 3), np.uint32)
            L.orc_measure_arrays(_c_p(h), _ptr(vc), _ptr(g))
            out["vertex_component"], out["grouped"] = vc, g
    finally:
        L.orc_measure_free(h)
    return out


def _bytes(h):
    L = lib()
    L.orc_bytes.restype = ctypes.c_size_t
    L.orc_bytes.argtypes = [_c_p, _c_p]
    L.orc_bytes_free.argtypes = [_c_p]
    try:
        n = L.orc_bytes(h, None)
        buf = ctypes.create_string_buffer(max(n, 1))
        L.orc_bytes(h, buf)
        return buf.raw[:n]
    finally:
        L.orc_bytes_free(h)


def _strings(items):
    return b"".join(str(x).encode() + b"\0" for x in items) or b"\0"


def _mesh_args(mesh):
    p = np.ascontiguousarray(mesh["positions"], np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(mesh["triangles"], np.uint32).reshape(-1, 3)
    get = lambda k, dt: None if mesh.get(k) is None else np.ascontiguousarray(mesh[k], dt)
    return p, t, get("normals", np.float32), get("atoms", np.uint32), get("patches", np.uint8), \
        get("vertex_component", np.uint32)


def fixed(values, prec):
    """push_fixed (export/fmt.rs) of each f32 value."""
    L = lib()
    L.orc_fixed.restype = _c_p
    L.orc_fixed.argtypes = [_c_p, ctypes.c_size_t, ctypes.c_int]
    x = np.ascontiguousarray(values, np.float32)
    return _bytes(L.orc_fixed(_ptr(x), len(x), prec)).split(b"\0")[:-1]


def write_obj(mesh, precision=4, normals=True, comments=(), object_name="ses", vertex_colors=None, groups=()):
    """OBJ bytes of a mesh dict (export/obj.rs write_obj)."""
    L = lib()
    L.orc_write_obj.restype = _c_p
    L.orc_write_obj.argtypes = [_c_p, _c_p, ctypes.c_size_t, _c_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, _c_p, _c_p, ctypes.c_char_p,
                                ctypes.c_size_t]
    p, t, nr, _, _, _ = _mesh_args(mesh)
    rgb = None if vertex_colors is None else np.ascontiguousarray(vertex_colors, np.float32)
    gf = np.array([g[0] for g in groups], np.uint64)
    h = L.orc_write_obj(_ptr(p), _ptr(nr), len(p), _ptr(t), len(t), precision, int(normals), _strings(comments),
                        len(comments), object_name.encode(), _ptr(rgb), _ptr(gf), _strings([g[1] for g in groups]),
                        len(groups))
    return _bytes(h)


def write_ply(mesh, normals=True, comments=(), vertex_colors=None):
    """Binary PLY bytes of a mesh dict (export/ply.rs write_ply)."""
    L = lib()
    L.orc_write_ply.restype = _c_p
    L.orc_write_ply.argtypes = [_c_p] * 5 + [ctypes.c_size_t, _c_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_char_p,
                                             ctypes.c_size_t, _c_p]
    p, t, nr, at, pa, vc = _mesh_args(mesh)
    rgb = None if vertex_colors is None else np.ascontiguousarray(vertex_colors, np.float32)
    h = L.orc_write_ply(_ptr(p), _ptr(nr), _ptr(at), _ptr(pa), _ptr(vc), len(p), _ptr(t), len(t), int(normals),
                        _strings(comments), len(comments), _ptr(rgb))
    return _bytes(h)


def write_msms(mesh, description, spheres, density, probe):
    """MSMS (.vert bytes, .face bytes) of a mesh dict (export/msms.rs)."""
    L = lib()
    L.orc_write_msms.restype = _c_p
    L.orc_write_msms.argtypes = [_c_p] * 5 + [ctypes.c_size_t, _c_p, ctypes.c_size_t, ctypes.c_char_p,
                                              ctypes.c_size_t, ctypes.c_double, ctypes.c_double, _c_p]
    p, t, nr, at, pa, vc = _mesh_args(mesh)
    vl = ctypes.c_size_t()
    h = L.orc_write_msms(_ptr(p), _ptr(nr), _ptr(at), _ptr(pa), _ptr(vc), len(p), _ptr(t), len(t),
                         description.encode(), spheres, density, probe, ctypes.byref(vl))
    b = _bytes(h)
    return b[:vl.value], b[vl.value:]


def mc_table():
    off = np.zeros(257, np.uint16)
    tris = np.zeros(3 * 1024, np.uint8)
    lib().orc_mc_table(_ptr(off), _ptr(tris), ctypes.c_size_t(1024))
    return off, tris[:3 * int(off[256])].reshape(-1, 3)
