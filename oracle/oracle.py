"""ORACLE — test infrastructure only.

ctypes wrapper around oracle/liboracle.so, the CPU restatement of the
reference's hot-path algorithms.  Imported by tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py as the checker; the product path never
imports it.

Every function takes numpy arrays and a dtype (np.float64 = FP64 oracle,
np.float32 = single-precision restatement) and returns numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _suf(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _c(a, dtype):
    """Contiguous array of `dtype` (or None)."""
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _call(name: str, *args):
    fn = getattr(lib(), name)
    fn.restype = None
    conv = []
    for a in args:
        if isinstance(a, np.ndarray) or a is None:
            conv.append(_p(a))
        elif isinstance(a, (int, np.integer)):
            conv.append(ctypes.c_int(int(a)))
        elif isinstance(a, float):
            conv.append(ctypes.c_double(a))
        else:
            conv.append(a)
    fn(*conv)


# ----------------------------------------------------------------------------
# Skinning
# ----------------------------------------------------------------------------

def dqs_fwd(dq, w, jidx, rest_pos, rest_nrm=None, dtype=np.float64):
    B, J = dq.shape[0], dq.shape[1]
    N, K = jidx.shape
    dq, w, rp, rn = (_c(x, dtype) for x in (dq, w, rest_pos, rest_nrm))
    ji = _c(jidx, np.int32)
    pos = np.zeros((B, N, 3), dtype)
    nrm = np.zeros((B, N, 3), dtype) if rn is not None else None
    _call(f"orc_dqs_fwd_{_suf(dtype)}", dq, w, ji, rp, rn, B, N, K, J, pos, nrm)
    return pos, nrm


def dqs_bwd(g_pos, dq, w, jidx, rest_pos, rest_nrm=None, g_nrm=None, ref_compat=False,
            dtype=np.float64):
    B, J = dq.shape[0], dq.shape[1]
    N, K = jidx.shape
    gp, gn, dq, w, rp, rn = (_c(x, dtype) for x in (g_pos, g_nrm, dq, w, rest_pos, rest_nrm))
    ji = _c(jidx, np.int32)
    gdq = np.zeros((B, J, 8), dtype)
    gw = np.zeros((B, N, K), dtype)
    gr = np.zeros((B, N, 3), dtype)
    _call(f"orc_dqs_bwd_{_suf(dtype)}", gp, gn, dq, w, ji, rp, rn, B, N, K, J, gdq, gw, gr,
          1 if ref_compat else 0)
    return gdq, gw, gr


def dof_grad(g_pos, skinned, joint_pos, joint_axis, influence, dtype=np.float64):
    B, N = g_pos.shape[0], g_pos.shape[1]
    Jall = joint_pos.shape[1]
    D, E = influence.shape[0], influence.shape[1]
    args = [_c(x, dtype) for x in (g_pos, skinned, joint_pos, joint_axis, influence)]
    out = np.zeros((B, D), dtype)
    _call(f"orc_dof_grad_{_suf(dtype)}", *args, B, N, Jall, D, E, out)
    return out


def lbs_fwd(trans, w, jidx, rest_pos, rest_nrm, disp, dtype=np.float64):
    B, J = trans.shape[0], trans.shape[1]
    N, K = jidx.shape
    tr, w, rp, rn, dp = (_c(x, dtype) for x in (trans, w, rest_pos, rest_nrm, disp))
    ji = _c(jidx, np.int32)
    pos = np.zeros((B, N, 3), dtype)
    nrm = np.zeros((B, N, 3), dtype) if rn is not None else None
    _call(f"orc_lbs_fwd_{_suf(dtype)}", tr, w, ji, rp, rn, dp, B, N, K, J, pos, nrm)
    return pos, nrm


def lbs_bwd(g_pos, trans, w, jidx, rest_pos, disp, ref_compat=False, dtype=np.float64):
    B, J = trans.shape[0], trans.shape[1]
    N, K = jidx.shape
    gp, tr, w, rp, dp = (_c(x, dtype) for x in (g_pos, trans, w, rest_pos, disp))
    ji = _c(jidx, np.int32)
    gw = np.zeros((B, N, K), dtype)
    gd = np.zeros((B, N, 3), dtype)
    gt = np.zeros((B, J, 12), dtype)
    _call(f"orc_lbs_bwd_{_suf(dtype)}", gp, tr, w, ji, rp, dp, B, N, K, J, gw, gd, gt,
          1 if ref_compat else 0)
    return gw, gd, gt
