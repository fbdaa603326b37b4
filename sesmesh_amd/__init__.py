"""sesmesh_amd: sesmesh's SES mesh stage on AMD MI355X (gfx950).

The product is ``libsesmesh_hip.so``: hand-written gfx950 HIP kernels behind
the C ABI of ``include/sesmesh_gpu.h``. This module is the thin Python host
over that ABI (ctypes, no torch types), mirroring the reference's library
API for the path: ``compute_surface(spheres, params)`` with
``SurfaceParams`` (reference crates/sesmesh/src/surface.rs:195-300, 599-685)
and the device/resident objects of ``sesmesh_gpu::hip`` (reference
crates/sesmesh-gpu/src/hip/mod.rs, resident.rs).

There is no CPU fallback: without the built library, or without a gfx950
device, every call raises.
"""
import ctypes
import os
import sys
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libsesmesh_hip.so"

SMG_ERRORS = {
    -1: "NoRuntime", -2: "NoDevice", -3: "Unsupported", -4: "Call", -5: "SelfTest", -6: "Length",
    -7: "InvalidInput", -8: "Internal", -9: "Cancelled",
}

STAGES = {0: "geometry", 1: "classify", 2: "mesh", 3: "assemble", 4: "measure"}
_PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64)


class BackendError(RuntimeError):
    """A failed C-ABI call (reference sesmesh_gpu::Error wrapped as
    sesmesh::Error::Backend, crates/sesmesh-gpu/src/error.rs:10-66)."""

    def __init__(self, code, message):
        super().__init__(f"hip: {SMG_ERRORS.get(code, code)}: {message}")
        self.code = code
        self.kind = SMG_ERRORS.get(code, str(code))


_vp = ctypes.c_void_p


class _GridRef(ctypes.Structure):
    _fields_ = [("origin", _vp), ("cell", ctypes.c_double), ("dims", ctypes.c_uint32 * 3), ("starts", _vp),
                ("items", _vp), ("n_items", ctypes.c_size_t)]


class ArrangementRef(ctypes.Structure):
    """smg_arrangement_ref of include/sesmesh_gpu.h."""
    _fields_ = [("n_spheres", ctypes.c_size_t), ("center", _vp), ("radius", _vp), ("expanded", _vp),
                ("status", _vp), ("orig", _vp), ("nb_off", _vp), ("nb_items", _vp), ("n_circles", ctypes.c_size_t),
                ("circle_ij", _vp), ("circle", _vp), ("cb_off", _vp), ("cb_items", _vp),
                ("n_vertices", ctypes.c_size_t), ("vertex", _vp), ("vertex_atoms", _vp), ("sphere_grid", _GridRef),
                ("circle_grid", _GridRef), ("vertex_grid", _GridRef), ("probe", ctypes.c_double),
                ("max_expanded", ctypes.c_double), ("rho_max", ctypes.c_double), ("kind", ctypes.c_int32)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64), ("compute_units", ctypes.c_int),
                ("clock_khz", ctypes.c_int), ("total_memory", ctypes.c_size_t), ("lds_per_block", ctypes.c_size_t),
                ("wavefront", ctypes.c_int)]


class MeshParams(ctypes.Structure):
    _fields_ = [("spacing", ctypes.c_double), ("brick", ctypes.c_uint32), ("refine", ctypes.c_int32),
                ("sizing", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class MeshTimes(ctypes.Structure):
    _fields_ = [("classify_ms", ctypes.c_double), ("sample_ms", ctypes.c_double), ("edges_ms", ctypes.c_double),
                ("refine_ms", ctypes.c_double), ("cells_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("k_classify_ms", ctypes.c_double), ("k_substates_ms", ctypes.c_double),
                ("k_sample_ms", ctypes.c_double), ("k_refine_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshInfo(ctypes.Structure):
    _fields_ = [("touched_bricks", ctypes.c_size_t), ("surface_bricks", ctypes.c_size_t),
                ("vertices", ctypes.c_size_t), ("triangles", ctypes.c_size_t),
                ("refine_evaluations", ctypes.c_uint64), ("sampled_subblocks", ctypes.c_size_t),
                ("overflow_subblocks", ctypes.c_size_t), ("evaluated_samples", ctypes.c_uint64),
                ("frame", ctypes.c_int64 * 3), ("band", ctypes.c_double),
                ("edge_clamp", ctypes.c_double), ("sample_retries", ctypes.c_uint64),
                ("refine_retries", ctypes.c_uint64)]


class PlanCaps(ctypes.Structure):
    """smg_plan_caps of include/sesmesh_gpu.h."""
    _fields_ = [("molecules", ctypes.c_uint32), ("brick", ctypes.c_uint32), ("refine", ctypes.c_int32),
                ("pad", ctypes.c_uint32), ("locals", ctypes.c_uint64), ("spheres", ctypes.c_uint64),
                ("box_bricks", ctypes.c_uint64), ("surface_bricks", ctypes.c_uint64),
                ("vertices", ctypes.c_uint64), ("triangles", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class Validation(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("boundary_edges", "non_manifold_edges", "misoriented_edges",
                                               "degenerate_triangles", "duplicate_triangles", "invalid_indices",
                                               "bad_vertices")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Component(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_uint64), ("triangles", ctypes.c_uint64), ("area", ctypes.c_double),
                ("volume", ctypes.c_double), ("euler", ctypes.c_int64), ("first_triangle", ctypes.c_uint64)]


class MeasureInfo(ctypes.Structure):
    _fields_ = [("validation", Validation), ("components", ctypes.c_uint64), ("cavities", ctypes.c_uint64),
                ("area", ctypes.c_double), ("volume", ctypes.c_double), ("euler", ctypes.c_int64),
                ("measured", ctypes.c_int32), ("pad", ctypes.c_int32)]


class ExportOptions(ctypes.Structure):
    """smg_export_options of include/sesmesh_gpu.h."""
    _fields_ = [("precision", ctypes.c_int32), ("normals", ctypes.c_int32), ("object_name", ctypes.c_char_p),
                ("comments", ctypes.POINTER(ctypes.c_char_p)), ("n_comments", ctypes.c_size_t),
                ("vertex_colors", _vp), ("group_first", _vp), ("group_names", ctypes.POINTER(ctypes.c_char_p)),
                ("n_groups", ctypes.c_size_t), ("description", ctypes.c_char_p), ("spheres", ctypes.c_uint64),
                ("density", ctypes.c_double), ("probe_radius", ctypes.c_double)]


def _export_options(precision=4, normals=True, object_name="ses", comments=(), vertex_colors=None, groups=(),
                    description="", spheres=0, density=0.0, probe_radius=0.0):
    """The options struct and the objects it points into (keep both alive)."""
    keep = []
    com = (ctypes.c_char_p * max(len(comments), 1))(*[str(c).encode() for c in comments])
    names = (ctypes.c_char_p * max(len(groups), 1))(*[str(g[1]).encode() for g in groups])
    first = np.array([g[0] for g in groups] or [0], np.uint64)
    rgb = None if vertex_colors is None else np.ascontiguousarray(vertex_colors, np.float32)
    keep += [com, names, first, rgb]
    o = ExportOptions(int(precision), int(bool(normals)), object_name.encode(), com, len(comments), _ptr(rgb),
                      _ptr(first), names, len(groups), description.encode(), int(spheres), float(density),
                      float(probe_radius))
    return o, keep


class MeshOut(ctypes.Structure):
    """smg_mesh_out of include/sesmesh_gpu.h: caller-owned device buffers."""
    _fields_ = [("positions", _vp), ("normals", _vp), ("atoms", _vp), ("patches", _vp), ("triangles", _vp),
                ("vertex_capacity", ctypes.c_uint64), ("triangle_capacity", ctypes.c_uint64)]


NOALLOC_SCRATCH_BYTES = 256 << 20


def mesh_into(resident, spacing, positions, normals, atoms, patches, triangles, vertex_capacity, triangle_capacity,
              brick=10, refine="exact", sizing=0):
    """The mesh stage of one molecule with no allocation (smg_mesh_into):
    the mesh is written to caller-owned device buffers given as addresses
    (positions and normals 3 f32 per vertex, atoms u32, patches u8,
    triangles 3 u32), intermediates use the device's fixed scratch array.
    Returns the counts (MeshInfo)."""
    out = MeshOut(positions, normals, atoms, patches, triangles, int(vertex_capacity), int(triangle_capacity))
    p = _mesh_params(spacing, brick, refine, sizing)
    info = MeshInfo()
    _check(lib().smg_mesh_into(resident._h, ctypes.byref(p), ctypes.byref(out), ctypes.byref(info)))
    return info


class Tuning(ctypes.Structure):
    """smg_tuning of include/sesmesh_gpu.h: launch shapes that leave the
    meshes unchanged."""
    _fields_ = [("sample_first_tier", ctypes.c_int32), ("refine_first_tier", ctypes.c_int32),
                ("refine_group_max", ctypes.c_int32), ("uniform_shared", ctypes.c_int32),
                ("scratch_workgroups", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


DEFAULT_TUNING = dict(sample_first_tier=0, refine_first_tier=0, refine_group_max=64, uniform_shared=1,
                      scratch_workgroups=0)


class TrainStats(ctypes.Structure):
    _fields_ = [("samples", ctypes.c_uint64), ("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32),
                ("occupancy_loss", ctypes.c_double), ("occupancy_accuracy", ctypes.c_double),
                ("load_rmse", ctypes.c_double), ("train_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlanStats(ctypes.Structure):
    _fields_ = [("runs", ctypes.c_uint64), ("replays", ctypes.c_uint64), ("fallbacks", ctypes.c_uint64),
                ("captures", ctypes.c_uint64), ("last_overflow", ctypes.c_uint32), ("pad", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


# Every function include/sesmesh_gpu.h declares, plus the helpers the
# library exports for tests and synthetic inputs.
ABI_FUNCTIONS = [
    "smg_last_error", "smg_device_open", "smg_device_close", "smg_selftest", "smg_arrangement_build",
    "smg_arrangement_build_kind", "smg_resident_build", "smg_resident_info_get", "smg_resident_arrangement",
    "smg_arrangement_free", "smg_arrangement_view", "smg_resident_new", "smg_resident_free", "smg_resident_bytes",
    "smg_mesh_new", "smg_mesh_free", "smg_mesh_info_get", "smg_mesh_copy", "smg_mesh_many", "smg_mesh_count",
    "smg_mesh_part_info", "smg_mesh_part_copy", "smg_field_clamped", "smg_plan_new", "smg_plan_caps_for",
    "smg_plan_run", "smg_plan_stats_get", "smg_plan_free", "smg_device_set_tuning", "smg_device_get_tuning",
    "smg_tune", "smg_model_new", "smg_model_free", "smg_model_param_count", "smg_model_get_params",
    "smg_model_set_params", "smg_model_collect", "smg_model_samples", "smg_model_dataset", "smg_model_forward",
    "smg_model_gradients", "smg_model_train", "smg_device_set_model", "smg_model_set_policy", "smg_mesh_into",
    "smg_mesh_measure", "smg_mesh_part_measure", "smg_mesh_part_components", "smg_mesh_upload",
    "smg_export_defaults", "smg_export", "smg_export_file", "smg_plan_caps_predict", "smg_device_set_control",
]

FORMATS = {"obj": 0, "ply": 1, "msms_vert": 2, "msms_face": 3}

_lib = None


def build(force=False):
    """Builds libsesmesh_hip.so (hipcc --offload-arch=gfx950) in place."""
    import subprocess
    args = ["make", "-C", str(HERE / "csrc")]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    """The loaded C-ABI library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with sesmesh_amd.build() (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.smg_last_error.restype = ctypes.c_char_p
        L.smg_resident_bytes.restype = ctypes.c_size_t
        L.smg_mesh_count.restype = ctypes.c_size_t
        L.smg_mesh_part_info.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.smg_mesh_part_copy.argtypes = [ctypes.c_void_p, ctypes.c_size_t] + [ctypes.c_void_p] * 9
        L.smg_field_clamped.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double,
                                        ctypes.c_void_p]
        L.smg_mesh_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                    ctypes.c_void_p]
        L.smg_plan_caps_for.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_double, ctypes.c_void_p]
        L.smg_plan_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_void_p, ctypes.c_void_p]
        L.smg_tune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p]
        L.smg_model_new.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        L.smg_model_param_count.restype = ctypes.c_size_t
        L.smg_model_samples.restype = ctypes.c_size_t
        L.smg_model_dataset.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t, ctypes.c_void_p]
        L.smg_model_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.smg_model_gradients.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t,
                                                                                      ctypes.c_void_p, ctypes.c_void_p]
        L.smg_model_train.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                      ctypes.c_uint64, ctypes.c_void_p]
        L.smg_model_collect.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.smg_model_set_policy.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_int32]
        for name in ("smg_device_close", "smg_arrangement_free", "smg_resident_free", "smg_mesh_free",
                     "smg_plan_free", "smg_model_free"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def _free(name, handle):
    """Frees a native object. The device outlives every object made on it
    (the library counts references), so objects may be freed in any order.
    Nothing is called while the interpreter shuts down: module globals may
    be gone by then, and the process's exit reclaims the device memory."""
    try:
        if not handle or _lib is None or sys.is_finalizing():
            return
        getattr(_lib, name)(handle)
    except Exception:
        pass


def _check(rc):
    if rc != 0:
        raise BackendError(rc, lib().smg_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class Device:
    """An open gfx950 device (reference hip::Device, crates/sesmesh-gpu/src/hip/mod.rs:170-234):
    opening runs the f32 and f64 arithmetic self-test."""

    def __init__(self, ordinal=0):
        self._h = _vp()
        info = DeviceInfo()
        _check(lib().smg_device_open(ctypes.c_int(ordinal), ctypes.byref(self._h), ctypes.byref(info)))
        self.ordinal = ordinal
        self.name = info.name.decode()
        self.arch = info.arch.decode()
        self.compute_units = info.compute_units
        self.clock_khz = info.clock_khz
        self.total_memory = info.total_memory
        self.lds_per_block = info.lds_per_block
        self.wavefront = info.wavefront

    @property
    def tuning(self):
        t = Tuning()
        _check(lib().smg_device_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    @tuning.setter
    def tuning(self, values):
        """Sets the launch shapes of the next runs (None: the defaults)."""
        if values is None:
            _check(lib().smg_device_set_tuning(self._h, None))
            return
        t = Tuning(**{**DEFAULT_TUNING, **dict(values)})
        _check(lib().smg_device_set_tuning(self._h, ctypes.byref(t)))

    def set_control(self, cancel=None, progress=None):
        """Cancellation and progress of the next runs (smg_device_set_control;
        reference Control, control.rs): `cancel` a ctypes.c_int32 the caller
        sets to non-zero to cancel, `progress` called as progress(stage name,
        done, total). Both None: no control."""

        class Control(ctypes.Structure):
            _fields_ = [("cancel", _vp), ("progress", _PROGRESS), ("user", _vp)]

        if cancel is None and progress is None:
            self._control = None
            _check(lib().smg_device_set_control(self._h, None))
            return
        cb = _PROGRESS(lambda user, stage, done, total: progress(STAGES.get(stage, stage), done, total)) \
            if progress else _PROGRESS()
        c = Control(ctypes.addressof(cancel) if cancel is not None else None, cb, None)
        self._control = (cancel, cb, c)  # kept alive while the library may call them
        _check(lib().smg_device_set_control(self._h, ctypes.byref(c)))

    def selftest(self, n=1 << 16, seed=1):
        m32, m64 = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().smg_selftest(self._h, ctypes.c_size_t(n), ctypes.c_uint64(seed), ctypes.byref(m32),
                                  ctypes.byref(m64)))
        return m32.value, m64.value

    def close(self):
        """Drops this handle; the device itself lives on until every object
        made on it is freed."""
        h = getattr(self, "_h", None)
        if h:
            self._h = _vp()
            _free("smg_device_close", h)

    def __del__(self):
        self.close()


KINDS = {"ses": 0, "sas": 1, "vdw": 2}


class Arrangement:
    """The host-built SAS arrangement (reference geom/arrangement.rs:437-645 and
    field/point.rs:213-274) of surface kind `kind`: "ses" (probe 0 gives
    "vdw"), "sas" or "vdw" (built with probe 0)."""

    def __init__(self, xyzr, probe=1.4, threads=0, kind="ses"):
        if kind not in KINDS:
            raise ValueError(f"unknown surface kind {kind!r} (ses, sas or vdw)")
        a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
        self._h = _vp()
        _check(lib().smg_arrangement_build_kind(_ptr(a), ctypes.c_size_t(len(a)), ctypes.c_double(probe),
                                                ctypes.c_int(KINDS[kind]), ctypes.c_int(threads),
                                                ctypes.byref(self._h)))
        self.ref = ArrangementRef()
        _check(lib().smg_arrangement_view(self._h, ctypes.byref(self.ref)))
        self.kind = {v: k for k, v in KINDS.items()}[self.ref.kind]
        self.probe = self.ref.probe

    @classmethod
    def _from_handle(cls, h):
        self = cls.__new__(cls)
        self._h = h
        self.ref = ArrangementRef()
        _check(lib().smg_arrangement_view(self._h, ctypes.byref(self.ref)))
        self.kind = {v: k for k, v in KINDS.items()}[self.ref.kind]
        self.probe = self.ref.probe
        return self

    def arrays(self):
        """Copies of the arrangement arrays (internal order)."""
        r = self.ref
        n, nc, nv = r.n_spheres, r.n_circles, r.n_vertices

        def arr(p, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(count,)).copy()

        nb = arr(r.nb_off, n + 1, np.uint32)
        cb = arr(r.cb_off, nc + 1, np.uint32)
        return {
            "centers": arr(r.center, 3 * n, np.float64).reshape(-1, 3), "radii": arr(r.radius, n, np.float64),
            "expanded": arr(r.expanded, n, np.float64), "status": arr(r.status, n, np.uint8),
            "orig": arr(r.orig, n, np.uint32), "neighbors": (nb, arr(r.nb_items, int(nb[-1]), np.uint32)),
            "circles_ij": arr(r.circle_ij, 2 * nc, np.uint32).reshape(-1, 2),
            "circles": arr(r.circle, 10 * nc, np.float64).reshape(-1, 10),
            "circle_blockers": (cb, arr(r.cb_items, int(cb[-1]), np.uint32)),
            "vertices": arr(r.vertex, 3 * nv, np.float64).reshape(-1, 3),
            "vertex_atoms": arr(r.vertex_atoms, 3 * nv, np.uint32).reshape(-1, 3),
            "max_expanded": r.max_expanded, "rho_max": r.rho_max,
            "grids": {k: self._grid(getattr(r, k)) for k in ("sphere_grid", "circle_grid", "vertex_grid")},
        }

    @staticmethod
    def _grid(g):
        dims = tuple(int(x) for x in g.dims)
        cells = dims[0] * dims[1] * dims[2]
        origin = np.ctypeslib.as_array(ctypes.cast(g.origin, ctypes.POINTER(ctypes.c_double)), shape=(3,)).copy()
        starts = np.ctypeslib.as_array(ctypes.cast(g.starts, ctypes.POINTER(ctypes.c_uint32)), shape=(cells + 1,)).copy()
        items = (np.ctypeslib.as_array(ctypes.cast(g.items, ctypes.POINTER(ctypes.c_uint32)), shape=(g.n_items,)).copy()
                 if g.n_items else np.zeros(0, np.uint32))
        return {"origin": origin, "cell": g.cell, "dims": dims, "starts": starts, "items": items}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = _vp()
            _free("smg_arrangement_free", h)


class ResidentInfo(ctypes.Structure):
    """smg_resident_info of include/sesmesh_gpu.h."""
    _fields_ = [("n_spheres", ctypes.c_size_t), ("n_circles", ctypes.c_size_t), ("n_vertices", ctypes.c_size_t),
                ("kind", ctypes.c_int32), ("built_on_device", ctypes.c_int32), ("probe", ctypes.c_double),
                ("build_ms", ctypes.c_double), ("bytes", ctypes.c_size_t)]


class Resident:
    """The arrangement resident in HBM (reference hip::Resident::new,
    crates/sesmesh-gpu/src/hip/resident.rs:343-371), from a host-built
    Arrangement, or (Resident.build) built on the device."""

    def __init__(self, device, arrangement=None, _handle=None):
        self.device = device
        self._h = _vp()
        if _handle is not None:
            self._h = _handle
        else:
            _check(lib().smg_resident_new(device._h, ctypes.byref(arrangement.ref), ctypes.byref(self._h)))

    @classmethod
    def build(cls, device, xyzr, probe=1.4, kind="ses"):
        """SesField::new + Resident::new with the arrangement built on the
        device (smg_resident_build)."""
        if kind not in KINDS:
            raise ValueError(f"unknown surface kind {kind!r} (ses, sas or vdw)")
        a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
        h = _vp()
        _check(lib().smg_resident_build(device._h, _ptr(a), ctypes.c_size_t(len(a)), ctypes.c_double(probe),
                                        ctypes.c_int(KINDS[kind]), ctypes.byref(h)))
        return cls(device, _handle=h)

    def info(self):
        i = ResidentInfo()
        _check(lib().smg_resident_info_get(self._h, ctypes.byref(i)))
        out = {k: getattr(i, k) for k, _ in ResidentInfo._fields_}
        out["kind"] = {v: k for k, v in KINDS.items()}[out["kind"]]
        out["built_on_device"] = bool(out["built_on_device"])
        return out

    def arrangement(self):
        """The resident's arrangement copied back to the host (an Arrangement)."""
        h = _vp()
        _check(lib().smg_resident_arrangement(self._h, ctypes.byref(h)))
        return Arrangement._from_handle(h)

    @property
    def bytes(self):
        return lib().smg_resident_bytes(self._h)

    def mesh(self, spacing=0.5, brick=10, refine="exact", sizing=0):
        return Mesh(self, spacing, brick, refine, sizing)

    def field_clamped(self, points, band):
        """Exact f64 clamped SES field at points (n x 3): probe minus the
        distance to the probe-position set clamped at probe + band
        (reference SesField::clamped_value, field/point.rs:759-765)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pts), np.float64)
        _check(lib().smg_field_clamped(self._h, _ptr(pts), len(pts), float(band), _ptr(out)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = _vp()
            _free("smg_resident_free", h)


def _mesh_params(spacing, brick, refine, sizing):
    return MeshParams(spacing, brick, 0 if refine == "exact" else 1, sizing, 0)


class Mesh:
    """The mesh stage of one or more molecules on the device (one part per
    molecule); the meshes stay in HBM until copied."""

    def __init__(self, resident, spacing, brick, refine, sizing, _batch=None, _plan=None, _handle=None):
        self._h = _vp()
        self.times = MeshTimes()
        self._plan = None
        self._full = _handle is None  # an uploaded mesh holds its own arrays only
        if _handle is not None:
            self._h = _handle
            self.residents = []
        elif _plan is not None:
            plan, residents, params = _plan
            handles = (ctypes.c_void_p * len(residents))(*[r._h.value for r in residents])
            arr = (MeshParams * len(params))(*params)
            _check(lib().smg_plan_run(plan._h, handles, arr, len(residents), ctypes.byref(self._h),
                                      ctypes.byref(self.times)))
            self.residents = list(residents)
            self._plan = plan  # a view of the plan's buffers: keep the plan alive
        elif _batch is None:
            p = _mesh_params(spacing, brick, refine, sizing)
            _check(lib().smg_mesh_new(resident._h, ctypes.byref(p), ctypes.byref(self._h),
                                      ctypes.byref(self.times)))
            self.residents = [resident]
        else:
            residents, params = _batch
            handles = (ctypes.c_void_p * len(residents))(*[r._h.value for r in residents])
            arr = (MeshParams * len(params))(*params)
            _check(lib().smg_mesh_many(handles, arr, len(residents), ctypes.byref(self._h),
                                       ctypes.byref(self.times)))
            self.residents = list(residents)
        self.brick = brick
        self.parts = [self._part_info(i) for i in range(lib().smg_mesh_count(self._h))]
        info = self.parts[0]
        self.info = info
        self.n_vertices, self.n_triangles, self.n_bricks = info.vertices, info.triangles, info.surface_bricks

    def _part_info(self, i):
        info = MeshInfo()
        _check(lib().smg_mesh_part_info(self._h, i, ctypes.byref(info)))
        return info

    def __len__(self):
        return len(self.parts)

    def arrays(self, values=False, part=0):
        info = self.parts[part]
        nv, nt, nb = info.vertices, info.triangles, info.surface_bricks
        out = {"positions": np.zeros((nv, 3), np.float32), "normals": np.zeros((nv, 3), np.float32),
               "atoms": np.zeros(nv, np.uint32), "patches": np.zeros(nv, np.uint8),
               "triangles": np.zeros((nt, 3), np.uint32), "bricks": np.zeros((nb, 3), np.int32),
               "t": np.zeros(nv, np.float64), "evals": np.zeros(nv, np.uint32)}
        vals = np.zeros((nb, (self.brick + 1) ** 3), np.float32) if values else None
        if not self._full:
            for k in ("normals", "bricks", "t", "evals"):
                out.pop(k)
        _check(lib().smg_mesh_part_copy(self._h, part, _ptr(out["positions"]), _ptr(out.get("normals")),
                                        _ptr(out["atoms"]), _ptr(out["patches"]), _ptr(out["triangles"]),
                                        _ptr(out.get("bricks")), _ptr(vals), _ptr(out.get("t")),
                                        _ptr(out.get("evals"))))
        if values:
            out["values"] = vals
        return out

    def measure(self):
        """The measure stage on the device (smg_mesh_measure): validation,
        components, triangles grouped by component (arrays() then returns
        the grouped order). Returns the device time in ms."""
        ms = ctypes.c_double()
        _check(lib().smg_mesh_measure(self._h, ctypes.byref(ms)))
        return ms.value

    def measure_info(self, part=0):
        mi = MeasureInfo()
        _check(lib().smg_mesh_part_measure(self._h, ctypes.c_size_t(part), ctypes.byref(mi)))
        return {"validation": mi.validation.as_dict(), "components": int(mi.components),
                "cavities": int(mi.cavities), "area": mi.area, "volume": mi.volume, "euler": int(mi.euler),
                "measured": bool(mi.measured)}

    def components(self, part=0):
        """The part's components (arrays of vertices, triangles, area,
        volume, euler, first_triangle) and per-vertex component ids."""
        nc = self.measure_info(part)["components"]
        arr = (Component * max(nc, 1))()
        vc = np.zeros(self.parts[part].vertices, np.uint32)
        _check(lib().smg_mesh_part_components(self._h, ctypes.c_size_t(part), arr, ctypes.c_size_t(nc), _ptr(vc)))
        comp = {k: np.array([getattr(arr[i], k) for i in range(nc)]) for k, _ in Component._fields_}
        return comp, vc

    def export(self, fmt, part=0, **options):
        """The file `fmt` ("obj", "ply", "msms_vert", "msms_face") of a part,
        formatted on the device (smg_export), as bytes. Options: precision,
        normals, object_name, comments, vertex_colors, groups (first
        triangle, name), description, spheres, density, probe_radius."""
        o, keep = _export_options(**options)
        L = lib()
        L.smg_export.argtypes = [_vp, ctypes.c_size_t, ctypes.c_int, _vp, _vp, ctypes.c_size_t, _vp]
        n = ctypes.c_size_t()
        _check(L.smg_export(self._h, part, FORMATS[fmt], ctypes.byref(o), None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(n.value, 1))
        _check(L.smg_export(self._h, part, FORMATS[fmt], ctypes.byref(o), buf, n.value, ctypes.byref(n)))
        del keep
        return buf.raw[:n.value]

    def write(self, path, fmt, part=0, **options):
        """Writes a part's file `fmt` to `path` (smg_export_file)."""
        o, keep = _export_options(**options)
        L = lib()
        L.smg_export_file.argtypes = [_vp, ctypes.c_size_t, ctypes.c_int, _vp, ctypes.c_char_p]
        _check(L.smg_export_file(self._h, part, FORMATS[fmt], ctypes.byref(o), str(path).encode()))
        del keep

    def copy_to(self, positions, normals, atoms, patches, triangles, part=0):
        """Copies a part's mesh into caller-owned host buffers given as raw
        addresses (e.g. pinned memory): positions and normals 3 f32 per
        vertex, atoms u32, patches u8, triangles 3 u32 each."""
        _check(lib().smg_mesh_part_copy(self._h, part, *[_vp(a) for a in (positions, normals, atoms, patches,
                                                                            triangles)], None, None, None, None))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = _vp()
            _free("smg_mesh_free", h)


def upload_mesh(device, positions, triangles, normals=None, atoms=None, patches=None):
    """A host mesh (one part) on the device (smg_mesh_upload), to measure
    or write it."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    nr = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    at = None if atoms is None else np.ascontiguousarray(atoms, np.uint32)
    pa = None if patches is None else np.ascontiguousarray(patches, np.uint8)
    h = _vp()
    L = lib()
    L.smg_mesh_upload.argtypes = [_vp] * 5 + [ctypes.c_size_t, _vp, ctypes.c_size_t, _vp]
    _check(L.smg_mesh_upload(device._h, _ptr(p), _ptr(nr), _ptr(at), _ptr(pa), len(p), _ptr(t), len(t),
                             ctypes.byref(h)))
    return Mesh(None, None, 10, "exact", 0, _handle=h)


def mesh_many(residents, spacings, brick=10, refine="exact", sizing=0):
    """The mesh stage of many resident molecules (of one device) in one
    batched pass (smg_mesh_many); part i equals residents[i].mesh(spacings[i])."""
    params = [_mesh_params(float(h), brick, refine, sizing) for h in spacings]
    if len(params) != len(residents):
        raise ValueError("one spacing per molecule")
    return Mesh(None, None, brick, refine, sizing, _batch=(residents, params))


class Plan:
    """A capture-ready batched mesh stage (smg_plan_*): buffers sized once by
    capacities, the whole pipeline replayed as a hipGraph; runs beyond the
    capacities fall back to smg_mesh_many and grow the plan."""

    def __init__(self, device, caps):
        self._h = _vp()
        self.device = device
        self.brick, self.refine = caps.brick, "exact" if caps.refine == 0 else "fast"
        _check(lib().smg_plan_new(device._h, ctypes.byref(caps), ctypes.byref(self._h)))

    @staticmethod
    def caps_for(residents, spacings, brick=10, refine="exact", mesh=None, margin=0.25):
        """Capacities for these molecules: exact on the table side; from
        `mesh` (their smg_mesh_many result) grown by `margin` on the device side."""
        params = (MeshParams * len(residents))(*[_mesh_params(float(h), brick, refine, 0) for h in spacings])
        handles = (ctypes.c_void_p * len(residents))(*[r._h.value for r in residents])
        caps = PlanCaps()
        _check(lib().smg_plan_caps_for(handles, params, len(residents), mesh._h if mesh is not None else None,
                                       ctypes.c_double(margin), ctypes.byref(caps)))
        return caps

    @staticmethod
    def caps_predict(residents, spacings, model, brick=10, refine="exact", margin=0.15):
        """Capacities for molecules no run has seen, from a sizing model
        (smg_plan_caps_predict): returns (caps, estimates)."""

        class Est(ctypes.Structure):
            _fields_ = [("touched_bricks", ctypes.c_uint64), ("surface_bricks", ctypes.c_double),
                        ("occupied_bricks", ctypes.c_double), ("vertices", ctypes.c_double),
                        ("triangles", ctypes.c_double), ("device_ms", ctypes.c_double)]

        params = (MeshParams * len(residents))(*[_mesh_params(float(h), brick, refine, 0) for h in spacings])
        handles = (ctypes.c_void_p * len(residents))(*[r._h.value for r in residents])
        caps, est = PlanCaps(), Est()
        L = lib()
        L.smg_plan_caps_predict.argtypes = [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_double, _vp, _vp]
        _check(L.smg_plan_caps_predict(handles, params, len(residents), model._h, ctypes.c_double(margin),
                                       ctypes.byref(caps), ctypes.byref(est)))
        return caps, {k: getattr(est, k) for k, _ in Est._fields_}

    def run(self, residents, spacings):
        """The meshes of `residents` (part i equals smg_mesh_many's part i). A
        mesh that views the plan must be dropped before the next run."""
        params = [_mesh_params(float(h), self.brick, self.refine, 0) for h in spacings]
        if len(params) != len(residents):
            raise ValueError("one spacing per molecule")
        return Mesh(None, None, self.brick, self.refine, 0, _plan=(self, residents, params))

    def stats(self):
        st = PlanStats()
        _check(lib().smg_plan_stats_get(self._h, ctypes.byref(st)))
        return st.as_dict()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = _vp()
            _free("smg_plan_free", h)


class Model:
    """The learned per-brick sizing model (smg_model_*): a 64-128-128-2 MLP on
    the BF16 matrix cores, trained on the device on bricks labelled by the
    exact classifier and gathers. It only chooses which LDS capacity a
    sub-block is tried at first; meshes do not depend on it."""

    N_PARAMS = 25090

    def __init__(self, device, seed=1):
        self._h = _vp()
        self.device = device
        _check(lib().smg_model_new(device._h, ctypes.c_uint64(seed), ctypes.byref(self._h)))

    @property
    def params(self):
        out = np.zeros(self.N_PARAMS, np.float32)
        _check(lib().smg_model_get_params(self._h, _ptr(out)))
        return out

    @params.setter
    def params(self, values):
        v = np.ascontiguousarray(values, np.float32)
        if v.shape != (self.N_PARAMS,):
            raise ValueError(f"{self.N_PARAMS} parameters")
        _check(lib().smg_model_set_params(self._h, _ptr(v)))

    def collect(self, residents, spacings, brick=10):
        """Appends the touched bricks of these molecules' exact mesh stage to
        the training set."""
        params = (MeshParams * len(residents))(*[_mesh_params(float(h), brick, "exact", 0) for h in spacings])
        handles = (ctypes.c_void_p * len(residents))(*[r._h.value for r in residents])
        _check(lib().smg_model_collect(self._h, handles, params, len(residents)))

    @property
    def samples(self):
        return lib().smg_model_samples(self._h)

    def dataset(self):
        n = self.samples
        x, occ, t = np.zeros((n, 64), np.float32), np.zeros(n, np.uint8), np.zeros(n, np.float32)
        got = ctypes.c_size_t()
        _check(lib().smg_model_dataset(self._h, _ptr(x), _ptr(occ), _ptr(t), n, ctypes.byref(got)))
        return x, occ, t

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 64)
        y = np.zeros((len(x), 2), np.float32)
        _check(lib().smg_model_forward(self._h, _ptr(x), len(x), _ptr(y)))
        return y

    def gradients(self, x, occ, target):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 64)
        occ = np.ascontiguousarray(occ, np.uint8)
        target = np.ascontiguousarray(target, np.float32)
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = np.zeros(2, np.float64)
        _check(lib().smg_model_gradients(self._h, _ptr(x), _ptr(occ), _ptr(target), len(x), _ptr(g), _ptr(loss)))
        return g, loss

    def train(self, epochs=10, batch=4096, lr=1e-3, seed=1):
        st = TrainStats()
        _check(lib().smg_model_train(self._h, epochs, batch, ctypes.c_float(lr), ctypes.c_uint64(seed),
                                     ctypes.byref(st)))
        return st.as_dict()

    def policy(self, scale=0.3, max_tier=1):
        """Placement: first tier holding scale x the predicted load, at most max_tier."""
        _check(lib().smg_model_set_policy(self._h, ctypes.c_float(scale), int(max_tier)))

    def use(self):
        """Runs with sizing 1 on the model's device use this model."""
        _check(lib().smg_device_set_model(self.device._h, self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = _vp()
            _free("smg_model_free", h)


def tune(residents, spacings, brick=10, refine="exact", repeats=2, bricks=None):
    """Autotuner over the launch shapes (smg_tune: coordinate descent on the
    device time of the mesh stage of `residents`), and over the brick size
    when `bricks` lists candidates (the meshes then equal the CPU back end's
    at the chosen brick size). Leaves the fastest tuning on the device and
    returns {"brick", "tuning", "ms", "trials"}."""
    dev = residents[0].device
    trials = []
    for b in (bricks or [brick]):
        params = (MeshParams * len(residents))(*[_mesh_params(float(h), int(b), refine, 0) for h in spacings])
        handles = (ctypes.c_void_p * len(residents))(*[r._h.value for r in residents])
        dev.tuning = None
        best, ms = Tuning(), ctypes.c_double()
        _check(lib().smg_tune(handles, params, len(residents), int(repeats), ctypes.byref(best), ctypes.byref(ms)))
        trials.append({"brick": int(b), "tuning": best.as_dict(), "ms": ms.value})
    win = min(trials, key=lambda t: t["ms"])
    dev.tuning = win["tuning"]
    return {**win, "trials": trials}


@dataclass
class SurfaceParams:
    """The parameters of the mesh stage (reference SurfaceParams,
    crates/sesmesh/src/surface.rs:195-300): SES with probe 1.4 A, spacing
    0.5 A, brick 10, exact refinement. `kind` is "ses", "sas" or "vdw"
    (SurfaceParams::ses / sas / vdw, surface.rs:245-285)."""
    probe_radius: float = 1.4
    spacing: float = 0.5
    brick_size: int = 10
    refine: str = "exact"
    sizing: int = 0
    kind: str = "ses"


def compute_surface(xyzr, params=None, device=None):
    """The mesh of spheres `xyzr` (n x 4: centre, radius) on the GPU for the
    surface kind of `params`: returns a dict of the mesh arrays (positions,
    normals, atoms, patches, triangles) and stats, equal to the reference
    CPU back end's mesh."""
    params = params or SurfaceParams()
    dev = device or Device(0)
    res = Resident.build(dev, xyzr, params.probe_radius, params.kind)  # the arrangement on the device
    m = res.mesh(params.spacing, params.brick_size, params.refine, params.sizing)
    out = m.arrays()
    out["stats"] = {"kind": res.info()["kind"], "touched_bricks": m.info.touched_bricks, "surface_bricks": m.info.surface_bricks,
                    "vertices": m.n_vertices, "triangles": m.n_triangles,
                    "refine_evaluations": m.info.refine_evaluations, "timings_ms": m.times.as_dict()}
    return out


# ------------------------------------------------------------ synthetic inputs
def synth_globule(seed, n_atoms, radius=None):
    """Seeded random-walk globule (n x 4 xyzr); ball radius 14 A at 330 atoms,
    scaled as n^(1/3) unless given."""
    if radius is None:
        radius = 14.0 * (n_atoms / 330.0) ** (1.0 / 3.0)
    out = np.zeros((n_atoms, 4), np.float64)
    lib().smg_synth_globule(ctypes.c_uint64(seed), ctypes.c_size_t(n_atoms), ctypes.c_double(radius), _ptr(out))
    return out


def synth_rings(seed=2, n_globule=4100, per_ring=7, ring_radius=55.0, separation=40.0):
    """Two stacked 7-fold rings of one 4.1k-atom globule (the 58k-atom assembly)."""
    radius = 14.0 * (n_globule / 330.0) ** (1.0 / 3.0)
    out = np.zeros((2 * per_ring * n_globule, 4), np.float64)
    lib().smg_synth_rings(ctypes.c_uint64(seed), ctypes.c_size_t(n_globule), ctypes.c_double(radius),
                          ctypes.c_int(per_ring), ctypes.c_double(ring_radius), ctypes.c_double(separation), _ptr(out))
    return out


def synth_shell(seed=3, n_globule=40000, copies=60, shell_radius=450.0):
    """60 icosahedrally placed copies of one seeded globule on a shell of
    radius 450 A (the 2.4M-atom precision-study assembly)."""
    radius = 14.0 * (n_globule / 330.0) ** (1.0 / 3.0)
    out = np.zeros((copies * n_globule, 4), np.float64)
    lib().smg_synth_shell(ctypes.c_uint64(seed), ctypes.c_size_t(n_globule), ctypes.c_double(radius),
                          ctypes.c_int(copies), ctypes.c_double(shell_radius), _ptr(out))
    return out


def mesh_area_volume(positions, triangles):
    """Area and enclosed volume of a triangle mesh, summed in f64 in
    triangle order (reference mesh/mod.rs area / volume)."""
    p = positions.astype(np.float64)
    a, b, c = p[triangles[:, 0]], p[triangles[:, 1]], p[triangles[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    volume = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    return float(area), float(volume)


def precision_report(resident, spacing, brick=10):
    """Fast (f32) against exact (f64) refinement of one
    molecule: |F| at every fast and exact vertex (F: the exact f64 clamped
    field), and the relative area and volume error of the fast mesh."""
    import time

    out = {}
    meshes = {}
    for mode in ("exact", "fast"):
        t = time.perf_counter()
        m = resident.mesh(spacing, brick, refine=mode)
        out[f"{mode}_s"] = time.perf_counter() - t
        out[f"{mode}_refine_ms"] = m.times.k_refine_ms
        g = m.arrays()
        t = time.perf_counter()
        f = np.abs(resident.field_clamped(g["positions"].astype(np.float64), 4.0))
        out[f"{mode}_field_s"] = time.perf_counter() - t
        m.measure()  # areas and volumes in the reference's summation order
        mi = m.measure_info()
        meshes[mode] = g
        out[mode] = {"vertices": int(m.n_vertices), "triangles": int(m.n_triangles),
                     "redone_exactly": int(((g["evals"] >> 30) & 1).sum()) if mode == "fast" else 0,
                     "max_abs_F": float(f.max()) if len(f) else 0.0,
                     "mean_abs_F": float(f.mean()) if len(f) else 0.0,
                     "area": mi["area"], "volume": mi["volume"], "components": mi["components"]}
        del m
    e, f = out["exact"], out["fast"]
    out["rel_area_error"] = abs(f["area"] - e["area"]) / e["area"] if e["area"] else 0.0
    out["rel_volume_error"] = abs(f["volume"] - e["volume"]) / abs(e["volume"]) if e["volume"] else 0.0
    same = (e["vertices"] == f["vertices"]) and (e["triangles"] == f["triangles"])
    out["same_topology"] = bool(same and np.array_equal(meshes["exact"]["triangles"], meshes["fast"]["triangles"]))
    if e["vertices"] == f["vertices"]:
        d = np.linalg.norm(meshes["exact"]["positions"].astype(np.float64) -
                           meshes["fast"]["positions"].astype(np.float64), axis=1)
        out["max_vertex_shift"] = float(d.max()) if len(d) else 0.0
        out["mean_vertex_shift"] = float(d.mean()) if len(d) else 0.0
    return out


def synth_batch_spec(seed, count, lo=20, hi=5000):
    atoms = np.zeros(count, np.uint32)
    probe = np.zeros(count, np.float64)
    spacing = np.zeros(count, np.float64)
    seeds = np.zeros(count, np.uint64)
    lib().smg_synth_batch_spec(ctypes.c_uint64(seed), ctypes.c_size_t(count), ctypes.c_uint32(lo),
                               ctypes.c_uint32(hi), _ptr(atoms), _ptr(probe), _ptr(spacing), _ptr(seeds))
    return atoms, probe, spacing, seeds


def mc_table():
    off = np.zeros(257, np.uint16)
    edges = np.zeros(3 * 1024, np.uint8)
    _check(lib().smg_mc_table(_ptr(off), _ptr(edges), ctypes.c_size_t(len(edges))))
    return off, edges[:3 * int(off[256])].reshape(-1, 3)
