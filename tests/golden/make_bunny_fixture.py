"""Writes tests/golden/bunny_closed.json: the vertex and face arrays of the
reference's closed bunny test mesh (tests/Meshes.hh:182-341, function
bunny_closed_mesh), used by the Laplacian known-answer test
(tests/DynamicElementsTest.cc:60-91).  Only the numbers are stored.

Run in the build container (reads /root/reference):  python tests/golden/make_bunny_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/tests/Meshes.hh"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bunny_closed.json")


def numbers(block):
    return [float(t) for t in re.findall(r"-?\d+\.?\d*(?:e-?\d+)?", block)]


def main():
    text = open(SRC).read()
    body = text[text.index("void bunny_closed_mesh("):]
    v_hdr = re.search(r"_V = Eigen::MatrixX<PassiveT>\((\d+), 3\);\s*_V <<(.*?);", body, re.S)
    f_hdr = re.search(r"_F = Eigen::MatrixX<int>\((\d+), 3\);\s*_F <<(.*?);", body, re.S)
    nv, nf = int(v_hdr.group(1)), int(f_hdr.group(1))
    V = numbers(v_hdr.group(2))
    F = [int(v) for v in numbers(f_hdr.group(2))]
    assert len(V) == 3 * nv and len(F) == 3 * nf
    json.dump({"source": "tests/Meshes.hh:182-341 (bunny_closed_mesh)",
               "V": [V[3 * i:3 * i + 3] for i in range(nv)],
               "F": [F[3 * i:3 * i + 3] for i in range(nf)]}, open(OUT, "w"))
    print(f"wrote {OUT}: {nv} vertices, {nf} faces")


if __name__ == "__main__":
    main()
