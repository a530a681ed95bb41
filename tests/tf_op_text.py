"""Parser of TensorFlow REGISTER_OP(...) builder chains in C++ source text:
op name -> inputs / outputs [(name, type)] in order and attrs
[(name, type, default or None)].  Comments are stripped first, so an op
whose registration is commented out is not seen."""
import re

_STR = r'"((?:[^"\\]|\\.)*)"'


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _split(spec):
    name, _, rest = spec.partition(":")
    return name.strip(), rest.strip()


def parse_register_ops(text):
    text = _strip_comments(text)
    ops = {}
    for m in re.finditer(r'REGISTER_OP\(\s*' + _STR + r'\s*\)(.*?);', text, flags=re.S):
        name, chain = m.group(1), m.group(2)
        sig = {"inputs": [], "outputs": [], "attrs": []}
        for kind, spec in re.findall(r"\.(Input|Output|Attr)\(\s*" + _STR + r"\s*\)", chain):
            n, rest = _split(spec)
            if kind == "Attr":
                t, _, d = rest.partition("=")
                sig["attrs"].append([n, t.strip(), d.strip() if d else None])
            else:
                sig[kind.lower() + "s"].append([n, rest])
        ops[name] = sig
    return ops
