"""ctypes binding of libdcx.so (the C ABI declared in include/dcx.h).

There is no fallback: if the library is missing or a call fails, an exception
is raised.  Arguments are passed as raw device pointers (``Tensor.data_ptr()``)
and the current HIP stream, exactly like the reference's TF kernels hand
tensor pointers to their CUDA launchers.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# DCX_LIBRARY names another build of the same library (e.g. to time two builds
# of the kernels side by side on one device); the default is the in-tree one.
LIB_PATH = os.environ.get("DCX_LIBRARY") or os.path.join(_HERE, "libdcx.so")

# Signature codes: p = pointer, i = int, z = size_t, s = stream, f = float
# Every function returns dcx_status (int) unless listed in _SPECIAL.
SIGNATURES: dict[str, str] = {
    "dcx_dqs_fwd": "pppppiiiippppS",
    "dcx_dqs_bwd": "pppppppiiiipppiS",
    "dcx_dof_grad": "pppppiiiiipS",
    "dcx_lbs_fwd": "ppppppiiiippppS",
    "dcx_lbs_bwd": "ppppppiiiipppiS",
    "dcx_topo_create": "piipp",
    "dcx_topo_destroy": "p",
    "dcx_render_fwd": "ppipppiippppppppzS",
    "dcx_render_fwd_noalloc": "ppipppiipppppppS",
    "dcx_render_bwd": "ppippppiippppppppppppzS",
    "dcx_edgraph_create": "piippppipppp",
    "dcx_edgraph_destroy": "p",
    "dcx_ed_fwd": "pippppppppppppzS",
    "dcx_ed_bwd": "pippppppppppzS",
    "dcx_mvsil_fwd": "pppppiiiiiipppppzS",
    "dcx_mvsil_bwd": "pppppiiiiipS",
    "dcx_camproj_fwd": "ppppiiiipS",
    "dcx_camproj_bwd": "ppppiiipS",
    "dcx_mesh_boundary": "ppiiiiipS",
    "dcx_pipeline_create": "iplp",
    "dcx_pipeline_destroy": "p",
    "dcx_pipeline_sizes": "pp",
    "dcx_pipeline_step": "ppS",
    "dcx_pipeline_capture": "ppS",
    "dcx_pipeline_replay": "pS",
    "dcx_pipeline_profile": "pppiS", "dcx_pipeline_set_tile": "pi", "dcx_pipeline_set_grid": "pi", "dcx_projected_mesh_boundary": "piiiippppppzS", "dcx_pipeline_autotune": "ppiSppp",
    "dcx_skeleton_create": "iipppppipppp",
    "dcx_skeleton_destroy": "p",
    "dcx_fk": "pippppS", "dcx_fk_transforms": "pippppS", "dcx_pose_to_graph": "ppppiiiiippS", "dcx_arap_fwd": "ppppppiiipppppS",
    "dcx_arap_bwd": "pppppiiippppS",
    "dcx_dof_influence": "piippipp",
    "dcx_topo_texel_table": "piiS",
    "dcx_render_normal_map": "ppiipiippppppzS",
    "dcx_edt": "piiippzS",
    "dcx_dt_encode": "piiippzS",
    "dcx_dtloss_fwd": "ppppiiiipppzS",
    "dcx_dtloss_bwd": "ppppiiiippS",
    "dcx_probe_atomics": "ipllS",
    "dcx_probe_copy": "ppliiS",
}

# Functions whose return type is not dcx_status.
_RESTYPES = {"dcx_render_workspace_size": ctypes.c_size_t, "dcx_ed_workspace_size": ctypes.c_size_t,
             "dcx_mvsil_workspace_size": ctypes.c_size_t, "dcx_pipeline_num_stages": ctypes.c_int,
             "dcx_pipeline_stage_name": ctypes.c_char_p, "dcx_normal_map_workspace_size": ctypes.c_size_t,
             "dcx_edt_workspace_size": ctypes.c_size_t, "dcx_dtloss_workspace_size": ctypes.c_size_t,
             "dcx_render_bwd_workspace_size": ctypes.c_size_t,
             "dcx_pipeline_tile_width": ctypes.c_int, "dcx_pipeline_grid_cap": ctypes.c_int,
             "dcx_projected_mesh_boundary_workspace_size": ctypes.c_size_t}
_EXTRA_ARGTYPES = {"dcx_render_workspace_size": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
                   "dcx_ed_workspace_size": [ctypes.c_void_p, ctypes.c_int],
                   "dcx_mvsil_workspace_size": [ctypes.c_int] * 5, "dcx_pipeline_num_stages": [],
                   "dcx_pipeline_stage_name": [ctypes.c_int],
                   "dcx_normal_map_workspace_size": [ctypes.c_void_p, ctypes.c_int],
                   "dcx_edt_workspace_size": [ctypes.c_int] * 3, "dcx_dtloss_workspace_size": [ctypes.c_int] * 2,
                   "dcx_render_bwd_workspace_size": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
                   "dcx_pipeline_tile_width": [ctypes.c_void_p], "dcx_pipeline_grid_cap": [ctypes.c_void_p],
                   "dcx_projected_mesh_boundary_workspace_size": [ctypes.c_int] * 5}

_CODES = {
    "p": ctypes.c_void_p,
    "i": ctypes.c_int,
    "z": ctypes.c_size_t,
    "S": ctypes.c_void_p,
    "f": ctypes.c_float,
    "l": ctypes.c_longlong,
}

_lock = threading.Lock()
_lib = None


class DcxError(RuntimeError):
    """A dcx entry point returned a non-zero status."""

    def __init__(self, name: str, status: int, message: str):
        super().__init__(f"{name} failed with status {status}: {message}")
        self.status = status


def load() -> ctypes.CDLL:
    """Loads libdcx.so once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m deepcharacters_amd.build` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        lib.dcx_version.restype = ctypes.c_int
        lib.dcx_last_error.restype = ctypes.c_char_p
        for name, sig in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [_CODES[c] for c in sig]
        for name, rt in _RESTYPES.items():
            fn = getattr(lib, name)
            fn.restype = rt
            fn.argtypes = _EXTRA_ARGTYPES[name]
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Calls a dcx entry point and raises DcxError on a non-zero status."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        msg = lib.dcx_last_error().decode(errors="replace")
        raise DcxError(name, status, msg)


def exported_symbols() -> list[str]:
    return list(SIGNATURES.keys()) + list(_RESTYPES.keys()) + ["dcx_version", "dcx_last_error"]
