"""Extract the StructGen known-answer vectors from the reference's unit tests.

Reads tests/geometry/mesh/implementation/source/MeshStructGenSourceTests.cpp
under the reference checkout as text and writes the expected arrays it holds
(counts, cell centres, cell→face labels, isBoundary, face→boundary labels,
cell1/cell2 labels) to structgen_kat.json beside this script. Only data
values are kept: the mesh sizes each test constructs and its expected arrays.

Usage: python tests/golden/make_structgen_kat.py [/root/reference]
"""
import json
import os
import re
import sys

SRC = "tests/geometry/mesh/implementation/source/MeshStructGenSourceTests.cpp"
CASES = ["getCellCount_test1", "getFaceCount_test1", "getBoundaryCount_test1", "getRegionCount_test1",
         "getVertexCount_test1", "getCellCenter_test1", "getCellFaceLabels_test1", "getFaceIsBoundary_test1",
         "getFaceBoundaryLabels_test1", "getFaceCell1Labels_test1", "getFaceCell2Labels_test1"]


def block(text, name):
    i = text.index(f"BOOST_AUTO_TEST_CASE({name})")
    j = text.find("BOOST_AUTO_TEST_CASE(", i + 10)
    return text[i:j if j > 0 else len(text)]


def numbers(s):
    return [float(x) if "." in x else int(x) for x in re.findall(r"-?\d+\.?\d*", s)]


def array(b, name):
    m = re.search(name + r"\[\d*\]\s*=\s*\{(.*?)\};", b, re.S)
    body = m.group(1)
    if "true" in body or "false" in body:
        return [t == "true" for t in re.findall(r"true|false", body)]
    if "EuclideanPoint" in body:
        return [numbers(t) for t in re.findall(r"EuclideanPoint<double,\s*3>\s*\(([^)]*)\)", body)]
    return numbers(body)


def source_args(b):
    m = re.search(r"MeshStructGenSource<int,\s*double>\s*source\(([^)]*)\)", b)
    return numbers(m.group(1))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, SRC)).read()
    out = {}
    for c in CASES:
        b = block(text, c)
        case = {"source": source_args(b)}
        m = re.search(r"BOOST_CHECK_EQUAL\((\w+),\s*(\d+)\);", b.split("E_SUCCESS);", 1)[-1])
        if m:
            case["expected"] = int(m.group(2))
        for nm in ("cmpIndices", "cmpData", "cmp", "faceLabels"):
            if re.search(nm + r"\[\d*\]\s*=", b):
                case[nm] = array(b, nm)
        out[c] = case
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "structgen_kat.json")
    with open(dst, "w") as f:
        json.dump({"source_file": SRC, "cases": out}, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
