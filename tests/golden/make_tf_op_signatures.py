"""Writes tf_op_signatures.json: the TensorFlow op interface (op name, inputs
and outputs with their types in order, attributes with types and defaults)
of every op the reference registers with REGISTER_OP, read as text from the
reference's .cpp files.  The drop-in's TF op shim must declare exactly these.

    python tests/golden/make_tf_op_signatures.py /root/reference
"""
import glob
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from tf_op_text import parse_register_ops  # noqa: E402


def main(ref_root):
    ops = {}
    for path in sorted(glob.glob(os.path.join(ref_root, "**", "*.cpp"), recursive=True)):
        text = open(path, errors="replace").read()
        for name, sig in parse_register_ops(text).items():
            sig["source"] = os.path.relpath(path, ref_root)
            ops[name] = sig
    out = os.path.join(HERE, "tf_op_signatures.json")
    with open(out, "w") as f:
        json.dump(ops, f, indent=1, sort_keys=True)
    print(f"{len(ops)} ops -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
