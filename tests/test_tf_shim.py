"""The TensorFlow op shim (deepcharacters_amd/tf/), checked as text and, for
its host glue, by running it -- TensorFlow is not installed here, so the shim
itself is not compiled:

* every REGISTER_OP of dcx_tf_ops.cc declares exactly the reference's op
  (name, inputs and outputs with dtypes in order, attributes with types and
  defaults), against tests/golden/tf_op_signatures.json (written from the
  reference's REGISTER_OP blocks by tests/golden/make_tf_op_signatures.py),
  and the fixture against the reference itself where it is present;
* the shim covers every op of the hot path and registers a GPU kernel for each;
* every dcx_* function the shim calls is declared in include/dcx.h;
* gradients.py registers a gradient for each forward op and calls each
  gradient op with its own input and attribute names;
* the CPython glue the file-reading constructors use returns the package's
  tables unchanged.
"""
import ast
import ctypes
import glob
import json
import os
import re
import subprocess
import sysconfig

import numpy as np
import pytest

from tf_op_text import parse_register_ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TF_DIR = os.path.join(ROOT, "deepcharacters_amd", "tf")
SHIM = os.path.join(TF_DIR, "dcx_tf_ops.cc")
REF = "/root/reference"

HOT_PATH_OPS = {
    "CudaRendererGpu", "CudaRendererGradGpu", "DqSkinningGpu", "DqSkinningGpuGrad", "LbSkinningGpu",
    "LbSkinningGpuGrad", "EmbeddedGraphGpu", "EmbeddedGraphGpuGrad", "EmbeddedGraphArapGpu",
    "EmbeddedGraphArapGpuGrad", "MultiViewSilhouetteLossGpu", "MultiViewSilhouetteLossGpuGrad",
    "CameraProjectionGpu", "CameraProjectionGpuGrad", "ProjectedMeshBoundaryGpu", "Pose2EmbeddedGraphCpu"}
GRAD_OF = {"CudaRendererGpu": "CudaRendererGradGpu", "DqSkinningGpu": "DqSkinningGpuGrad",
           "LbSkinningGpu": "LbSkinningGpuGrad", "EmbeddedGraphGpu": "EmbeddedGraphGpuGrad",
           "EmbeddedGraphArapGpu": "EmbeddedGraphArapGpuGrad",
           "MultiViewSilhouetteLossGpu": "MultiViewSilhouetteLossGpuGrad",
           "CameraProjectionGpu": "CameraProjectionGpuGrad"}


def snake(name):
    return re.sub(r"(?<!^)(?=[A-Z])", "_", name).lower()


@pytest.fixture(scope="module")
def reference_ops():
    return json.load(open(os.path.join(ROOT, "tests", "golden", "tf_op_signatures.json")))


@pytest.fixture(scope="module")
def shim_ops():
    return parse_register_ops(open(SHIM).read())


def test_shim_ops_match_the_reference_registrations(reference_ops, shim_ops):
    assert set(shim_ops) == HOT_PATH_OPS
    for name, sig in shim_ops.items():
        ref = reference_ops[name]
        for key in ("inputs", "outputs", "attrs"):
            assert sig[key] == ref[key], (name, key, sig[key], ref[key])


def test_parser_reads_inputs_outputs_and_defaults(reference_ops):
    r = reference_ops["CudaRendererGpu"]
    assert [n for n, _ in r["inputs"]] == ["vertex_pos", "vertex_color", "texture", "sh_coeff", "target_image",
                                           "extrinsics", "intrinsics"]
    assert r["outputs"][1] == ["face_buffer", "int32"]
    assert ["render_resolution_u", "int", "512"] in r["attrs"]
    assert ["compute_normal_map", "string", None] in r["attrs"]
    # a registration that is commented out in the reference is not an op
    assert "EmbeddedGraphGpuGrad" in reference_ops
    text = '//REGISTER_OP("X")\n// .Input("a: float");\nREGISTER_OP("Y")\n.Input("a: float")\n.Attr("k: int = 3");'
    ops = parse_register_ops(text)
    assert set(ops) == {"Y"} and ops["Y"]["attrs"] == [["k", "int", "3"]]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not present")
def test_fixture_matches_the_reference_sources(reference_ops):
    live = {}
    for path in glob.glob(os.path.join(REF, "**", "*.cpp"), recursive=True):
        for name, sig in parse_register_ops(open(path, errors="replace").read()).items():
            live[name] = sig
    assert set(live) == set(reference_ops)
    for name, sig in live.items():
        for key in ("inputs", "outputs", "attrs"):
            assert sig[key] == reference_ops[name][key], (name, key)


def test_every_op_has_a_gpu_kernel(shim_ops):
    text = open(SHIM).read()
    kernels = set(re.findall(r'REGISTER_KERNEL_BUILDER\(\s*Name\("([A-Za-z0-9]+)"\)\.Device\(DEVICE_GPU\)', text))
    assert kernels == set(shim_ops)


def test_shim_calls_only_declared_c_abi_functions():
    header = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "dcx.h")).read(), flags=re.S)
    declared = set(re.findall(r"\b(dcx_[a-z0-9_]+)\s*\(", header))
    text = re.sub(r"//[^\n]*", "", open(SHIM).read())
    called = set(re.findall(r"\b(dcx_[a-z0-9_]+)\s*\(", text))
    assert called, "the shim calls no C-ABI function"
    assert called <= declared, called - declared
    assert {"dcx_render_fwd", "dcx_render_bwd", "dcx_dqs_fwd", "dcx_dqs_bwd", "dcx_fk", "dcx_ed_fwd", "dcx_ed_bwd",
            "dcx_mvsil_fwd", "dcx_mvsil_bwd", "dcx_camproj_fwd", "dcx_camproj_bwd"} <= called


def _gradient_functions():
    tree = ast.parse(open(os.path.join(TF_DIR, "gradients.py")).read())
    out = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef):
            for dec in node.decorator_list:
                if isinstance(dec, ast.Call) and getattr(dec.func, "attr", "") == "RegisterGradient":
                    out[dec.args[0].value] = node
    return tree, out


def test_gradients_are_registered_with_the_grad_ops_signatures(reference_ops):
    tree, grads = _gradient_functions()
    assert set(GRAD_OF) | {"Pose2EmbeddedGraphCpu"} <= set(grads)
    not_diff = [n.args[0].value for n in ast.walk(tree) if isinstance(n, ast.Call)
                and getattr(n.func, "attr", "") == "NotDifferentiable"]
    assert not_diff == ["ProjectedMeshBoundaryGpu"]
    for fwd, grad in GRAD_OF.items():
        fn = grads[fwd]
        # the gradient function takes the op and one gradient per forward output
        assert len(fn.args.args) == 1 + len(reference_ops[fwd]["outputs"]), fwd
        calls = [c for c in ast.walk(fn) if isinstance(c, ast.Call)
                 and isinstance(c.func, ast.Attribute) and c.func.attr == snake(grad)]
        assert len(calls) == 1, (fwd, grad)
        call = calls[0]
        inputs = [n for n, _ in reference_ops[grad]["inputs"]]
        attrs = [a[0] for a in reference_ops[grad]["attrs"]]
        kw = [k.arg for k in call.keywords]
        assert set(kw) <= set(inputs) | set(attrs), (grad, set(kw) - set(inputs) - set(attrs))
        given = inputs[:len(call.args)] + [k for k in kw if k in inputs]
        assert sorted(given) == sorted(inputs), (grad, given)
        # every attribute without a default is passed
        assert {a[0] for a in reference_ops[grad]["attrs"] if a[2] is None} <= set(kw), grad


@pytest.fixture(scope="module")
def host_glue(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("tfhost") / "libdcx_tf_host.so")
    inc = sysconfig.get_paths()["include"]
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{inc}",
                    os.path.join(TF_DIR, "dcx_tf_host.cc"), "-o", out], check=True)
    return ctypes.PyDLL(out)


def _summary(lib, fn, *args):
    buf = ctypes.create_string_buffer(1 << 16)
    a = [x.encode() if x is not None else None for x in (list(args) + [None, None])[:2]]
    rc = lib.dcx_tf_host_selftest(fn.encode(), a[0], a[1], buf, len(buf))
    return rc, buf.value.decode()


def test_host_glue_returns_the_package_tables(host_glue):
    from deepcharacters_amd.tf import character_io
    rc, text = _summary(host_glue, "synthetic_skinning_character", "1500", "3")
    assert rc == 0, text
    want = character_io.synthetic_skinning_character("1500", "3")
    got = {}
    for item in text.strip(";").split(";"):
        name, kind, shape, csum = item.split(":")
        got[name] = (kind, tuple(int(d) for d in shape.split("x")) if shape else (), float(csum))
    assert set(got) == set(want)
    for name, a in want.items():
        kind, shape, csum = got[name]
        assert kind == ("f" if a.dtype == np.float32 else "i"), name
        assert shape == a.shape, name
        if a.dtype == np.float32:
            ref = float(a.astype(np.float64).sum())
        else:
            ref = float((a.reshape(-1).astype(np.float64) * (np.arange(a.size) % 7 + 1)).sum())
        assert abs(csum - ref) <= 1e-6 * max(1.0, abs(ref)), name


def test_host_glue_reports_reader_errors(host_glue):
    rc, text = _summary(host_glue, "skinning_character", "/nonexistent/character.character")
    assert rc == -1 and "skinning_character" in text
    rc, text = _summary(host_glue, "no_such_reader", "x")
    assert rc == -1 and "no_such_reader" in text
