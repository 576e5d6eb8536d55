"""Writes tests/golden/node_data_rules.json: the quadrature rules of the
reference's node_data.h (the tables tet_cubature / tri_cubature load,
operators.cpp:1870-2100) for the degrees 2N that the compiled kernel instances
use (triangles of degree 2-10: 2D N = 1-5 and the faces of 3D N = 1-4;
tetrahedra of degree 2-8: 3D N = 1-4), and the interpolation nodes of degree
1-6 (Nodes3D; degree N + 1 serves the Kopriva geofacs of degree N), as read
from the header text.

    python tests/golden/make_node_data_rules.py /root/reference/node_data.h

The fixture holds data only (points and weights, in the header's order and
digits); the package's rule table carries the same numbers
(scratch_wadge_amd/quadrature_rules.json), which a CPU test checks against it.
"""

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

KEYS = ["tri2", "tri4", "tri6", "tri8", "tri10", "tet2", "tet4", "tet6", "tet8",
        "nodes1", "nodes2", "nodes3", "nodes4", "nodes5", "nodes6"]


def main(path):
    from scratch_wadge_amd.operators import parse_node_data
    rules = parse_node_data(path)
    out = {k: rules[k] for k in KEYS}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "node_data_rules.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")
    print(dst, {k: len(v["points"]) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/node_data.h")
