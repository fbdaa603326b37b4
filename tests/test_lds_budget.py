"""LDS budget of the mesh-stage kernels, on the CPU.

The static LDS of each kernel (its __shared__ arrays) is read from the
gfx950 code object inside the built library (the `.group_segment_fixed_size`
of its kernel descriptor), and every even brick 4..32 is walked through the
library's own LDS plan: a staged edge or cell pass, or an LDS list tier that
runs, must fit dynamic plus static bytes in the 160 KiB a gfx950 workgroup
may hold, and staging is not given up while it fits."""
import ctypes
import re
import shutil
import subprocess

import numpy as np
import pytest

import sesmesh_amd

LDS_MAX = 163840
KINDS = ["k_edges", "k_cell_emit", "k_uniform", "k_sample", "k_refine", "k_refine_wide", "k_refine_fast",
         "k_refine_fast_wide"]
# Variants the pipeline launches per kind (the GLOBAL=true scratch variants take no dynamic LDS).
LAUNCHED = {"k_uniform": ("ILb0E", "ILb1E"), "k_sample": ("ILb0ELb0E", "ILb0ELb1E"),
            "k_refine": ("ILb0ELb0E", "ILb0ELb1E"), "k_refine_wide": ("ILb0ELb0E", "ILb0ELb1E"),
            "k_refine_fast": ("ILb0ELb0E", "ILb0ELb1E"), "k_refine_fast_wide": ("ILb0ELb0E", "ILb0ELb1E")}
# Other kernels that share a kind's LDS tiers (the touched-brick classification
# runs the sub-block classification's tiers).
SHARING = {"k_uniform": (("k_classify_touched", ("ILb0E", "ILb1E")),)}
TOOLS = ["objcopy", "/opt/rocm/llvm/bin/clang-offload-bundler", "/opt/rocm/llvm/bin/llvm-readelf"]


def kernel_statics(tmp_path):
    """Static LDS per kernel kind, from the library's gfx950 code object."""
    if any(shutil.which(t) is None for t in TOOLS):
        pytest.skip("binutils / ROCm LLVM tools not found")
    fat, co = tmp_path / "fatbin", tmp_path / "gfx950.co"
    subprocess.run([TOOLS[0], "-O", "binary", "--only-section=.hip_fatbin", str(sesmesh_amd.LIB_PATH), str(fat)],
                   check=True)
    subprocess.run([TOOLS[1], "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
    notes = subprocess.run([TOOLS[2], "--notes", str(co)], check=True, capture_output=True, text=True).stdout
    sizes, pending = {}, None
    for line in notes.splitlines():
        m = re.search(r"\.group_segment_fixed_size:\s+(\d+)", line)
        if m:
            pending = int(m.group(1))
            continue
        m = re.search(r"\.name:\s+(_Z\S+)", line)
        if m and pending is not None:
            sizes[m.group(1)] = pending
            pending = None
    out = []
    for kind in KINDS:
        prefix = f"_ZN3smg{len(kind)}{kind}"
        got = [v for name, v in sizes.items() if name.startswith(prefix) and
               (kind not in LAUNCHED or any(name[len(prefix):].startswith(t) for t in LAUNCHED[kind]))]
        assert got, kind
        for other, variants in SHARING.get(kind, ()):
            pre = f"_ZN3smg{len(other)}{other}"
            shared = [v for name, v in sizes.items() if any(name.startswith(pre + t) for t in variants)]
            assert shared, other
            got += shared
        out.append(max(got))
    return out


def lds_plan(brick, fast, statics):
    arr = (ctypes.c_uint32 * 8)(*statics)
    out = np.zeros(30, np.uint32)
    rc = sesmesh_amd.lib().smg_lds_plan(ctypes.c_uint32(brick), ctypes.c_int(fast), ctypes.c_uint64(LDS_MAX), arr,
                                        out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    p = {"edges_dyn": out[0], "cells_dyn": out[1], "edges_staged": out[2], "cells_staged": out[3],
         "uniform_dyn": out[4:7], "uniform_fits": out[7:10], "sample_dyn": out[10:15], "sample_fits": out[15:20],
         "refine_dyn": out[20:25], "refine_fits": out[25:30]}
    return p


def test_every_brick_fits_its_lds(tmp_path):
    st = dict(zip(KINDS, kernel_statics(tmp_path)))
    statics = [st[k] for k in KINDS]
    staged_cells = []
    for brick in range(4, 33, 2):
        for fast in (0, 1):
            p = lds_plan(brick, fast, statics)
            for dyn, flag, kind in ((p["edges_dyn"], p["edges_staged"], "k_edges"),
                                    (p["cells_dyn"], p["cells_staged"], "k_cell_emit")):
                # staged exactly when it fits
                assert bool(flag) == (int(dyn) + st[kind] <= LDS_MAX), (brick, kind)
            for u in range(3):
                assert bool(p["uniform_fits"][u]) == (int(p["uniform_dyn"][u]) + st["k_uniform"] <= LDS_MAX)
            for t in range(4):  # the LDS tiers (the fifth is global scratch)
                assert bool(p["sample_fits"][t]) == (int(p["sample_dyn"][t]) + st["k_sample"] <= LDS_MAX)
                kind = ("k_refine_fast" if fast else "k_refine") + ("_wide" if t >= 2 else "")
                assert bool(p["refine_fits"][t]) == (int(p["refine_dyn"][t]) + st[kind] <= LDS_MAX), (brick, t)
            assert p["sample_fits"][4] == 1 and p["refine_fits"][4] == 1
            # the first tiers always run
            assert p["sample_fits"][0] and p["refine_fits"][0] and p["uniform_fits"][0]
        if p["cells_staged"]:
            staged_cells.append(brick)
    # the cell pass keeps the samples as sign bits: brick 32 asks for 156,056
    # dynamic bytes, which with the kernel's static tables fit, so every
    # brick stages.
    assert staged_cells == list(range(4, 33, 2))
    assert lds_plan(32, 0, statics)["cells_dyn"] == 8 * (562 + 1) + 96 * 1536 + 4 * 1024
    assert lds_plan(32, 0, statics)["edges_staged"] == 1
