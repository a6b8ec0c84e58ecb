"""The Julia extension (julia/PotentialFrameWorkMI355XExt.jl) cannot run here
(no Julia), so its binding is checked statically: every C symbol it calls is
declared in include/pfw_mi355x.h, and its isbits structs list the fields of
the header's structs in the same order (their C layout then matches)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = open(os.path.join(ROOT, "julia", "PotentialFrameWorkMI355XExt.jl")).read()
HDR = open(os.path.join(ROOT, "include", "pfw_mi355x.h")).read()


def header_struct(name):
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", HDR).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            fields.append(re.findall(r"\w+", part)[-1])
    return fields


def julia_struct(name):
    body = re.search(r"struct " + name + r"\b.*?\n(.*?)\nend", JL, re.S).group(1)
    return [m.group(1) for m in re.finditer(r"^\s+(\w+)::", body, re.M)]


def test_called_symbols_are_declared():
    called = set(re.findall(r"ccall\(\(:(\w+), LIB\)", JL))
    assert len(called) >= 15
    missing = [c for c in called if not re.search(r"\b" + c + r"\(", HDR)]
    assert not missing, missing


def test_struct_layouts_match_the_header():
    for jl, c in (("PfwRadial", "pfw_ace_radial"), ("PfwModel", "pfw_ace_model"),
                  ("PfwFeatureLayer", "pfw_feature_layer")):
        assert julia_struct(jl) == header_struct(c), jl
