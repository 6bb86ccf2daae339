"""Extract the golden vectors that the reference's own test data holds into JSON fixtures.

The reference keeps its known-answer contraction cases as numeric tables inside
tests/gen/random-tensors.hpp (70 single contractions, grouped at :569-577) and
tests/gen/random-tn.hpp (20 small networks, grouped at :324-327).  This script
reads those tables as text and writes only the numbers (dims, elements, wires,
bonds, expected results) to:

    tests/golden/random_tensors.json
    tests/golden/random_tn.json

Run it in a checkout that has the reference mounted:

    python tests/golden/extract_fixtures.py [/root/reference]

Nothing under tests/ reads the reference at test time; the JSON files are the
committed fixtures.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

_NUM = r"[-+]?\s*\d+\.\d+"
_EL = re.compile(r"(" + _NUM + r")\s*([-+])\s*(\d+\.\d+)i")
_TI = re.compile(r"tensor_info\s*\{\{([^}]*)\}\s*,\s*\{([^}]*)\}\}")
_WIRE = re.compile(r"\{\s*(\d+)\s*,\s*(\d+)\s*\}")


def _parse_tensor(dims_txt, els_txt):
    dims = [int(x) for x in dims_txt.replace(" ", "").split(",") if x]
    els = []
    for m in _EL.finditer(els_txt):
        re_ = float(m.group(1).replace(" ", ""))
        im = float(m.group(3)) * (-1.0 if m.group(2) == "-" else 1.0)
        els.append([re_, im])
    n = 1
    for d in dims:
        n *= d
    if len(els) != n:
        raise ValueError("element count %d != prod(dims) %d" % (len(els), n))
    return {"dims": dims, "els": els}


def _blocks(text, kind):
    """Yield (index, body) for each `const <kind> vN { ... };` block."""
    pat = re.compile(r"const\s+" + kind + r"\s+v(\d+)\s*\{(.*?)\n  \};", re.S)
    for m in pat.finditer(text):
        yield int(m.group(1)), m.group(2)


def _groups(text, kind):
    pat = re.compile(r"const\s+std::vector<" + kind + r">\s+(\w+)\s*\{([^}]*)\};")
    out = {}
    for m in pat.finditer(text):
        out[m.group(1)] = [int(v.strip()[1:]) for v in m.group(2).split(",") if v.strip()]
    return out


def extract_contractions(path):
    text = open(path, encoding="utf-8").read()
    cases = {}
    for idx, body in _blocks(text, "contraction_validation"):
        tis = _TI.findall(body)
        if len(tis) != 3:
            raise ValueError("v%d: expected 3 tensors, got %d" % (idx, len(tis)))
        wtxt = body[body.index("std::vector<qtnh::wire>"):]
        wires = [[int(a), int(b)] for a, b in _WIRE.findall(wtxt)]
        t1, t2, t3 = (_parse_tensor(*t) for t in tis)
        cases["v%d" % idx] = {"t1": t1, "t2": t2, "t3": t3, "wires": wires}
    return {"cases": cases, "groups": _groups(text, "contraction_validation")}


def extract_networks(path):
    text = open(path, encoding="utf-8").read()
    nets = {}
    for idx, body in _blocks(text, "tn_validation"):
        tpart, rest = body.split("std::vector<bond_info>", 1)
        tensors = [_parse_tensor(*t) for t in _TI.findall(tpart)]
        bpart, rpart = rest.rsplit("tensor_info", 1)
        bonds = []
        for m in re.finditer(r"bond_info\s*\{\s*(\d+)\s*,\s*(\d+)\s*,\s*\{(.*?)\}\}", bpart):
            bonds.append({"t1": int(m.group(1)), "t2": int(m.group(2)),
                          "wires": [[int(a), int(b)] for a, b in _WIRE.findall("{" + m.group(3) + "}")]})
        res = _TI.findall("tensor_info" + rpart)
        if len(res) != 1:
            raise ValueError("v%d: result tensor not found" % idx)
        nets["v%d" % idx] = {"tensors": tensors, "bonds": bonds, "result": _parse_tensor(*res[0])}
    return {"networks": nets, "groups": _groups(text, "tn_validation")}


def main(argv):
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    gen = os.path.join(ref, "tests", "gen")
    con = extract_contractions(os.path.join(gen, "random-tensors.hpp"))
    tns = extract_networks(os.path.join(gen, "random-tn.hpp"))
    assert len(con["cases"]) == 70, len(con["cases"])
    assert len(tns["networks"]) == 20, len(tns["networks"])
    with open(os.path.join(HERE, "random_tensors.json"), "w") as f:
        json.dump(con, f, separators=(",", ":"))
    with open(os.path.join(HERE, "random_tn.json"), "w") as f:
        json.dump(tns, f, separators=(",", ":"))
    print("wrote %d contraction cases, %d networks" % (len(con["cases"]), len(tns["networks"])))


if __name__ == "__main__":
    main(sys.argv)
