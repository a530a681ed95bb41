"""The C ABI library loads (no GPU needed) and exports every symbol include/*.h declares."""
import ctypes
import glob
import os
import re

import pytest

from deepcharacters_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(dcx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "dcx_version" in names and "dcx_last_error" in names
    assert len(names) >= 5


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libdcx.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    declared = set(declared_functions())
    bound = set(_lib.exported_symbols())
    assert declared == bound, (declared - bound, bound - declared)


def test_version_and_error_calls_need_no_gpu():
    lib = _lib.load()
    assert lib.dcx_version() >= 100
    assert isinstance(lib.dcx_last_error(), bytes)


def test_invalid_arguments_return_error_status_without_gpu():
    """Argument validation happens on the host before any HIP call."""
    lib = _lib.load()
    st = lib.dcx_dqs_fwd(None, None, None, None, None, 0, 0, 0, 0, None, None, None, None, None)
    assert st == -1
    assert b"bad sizes" in lib.dcx_last_error()
    with pytest.raises(_lib.DcxError):
        _lib.call("dcx_lbs_fwd", None, None, None, None, None, None, 1, 1, 9, 1, None, None, None,
                  None, None)
    with pytest.raises(_lib.DcxError, match="null pipeline"):
        _lib.call("dcx_pipeline_set_grid", None, 32)
    assert lib.dcx_pipeline_grid_cap(None) == -1


def declared_params():
    """{function: [kind per parameter]}, kind in p (pointer), i (int), z (size_t), f (float)."""
    out = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"\b(dcx_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
            args = m.group(2).strip()
            kinds = []
            if args not in ("", "void"):
                for a in args.split(","):
                    if "*" in a:
                        kinds.append("p")
                    elif "size_t" in a:
                        kinds.append("z")
                    elif "float" in a:
                        kinds.append("f")
                    elif "long long" in a:
                        kinds.append("l")
                    else:
                        kinds.append("i")
            out[m.group(1)] = kinds
    return out


def test_python_signatures_match_header_parameter_kinds():
    decl = declared_params()
    for name, sig in _lib.SIGNATURES.items():
        got = ["p" if c == "S" else c for c in sig]
        assert got == decl[name], (name, "".join(got), "".join(decl[name]))
    for name, args in _lib._EXTRA_ARGTYPES.items():
        assert len(args) == len(decl[name]), name
