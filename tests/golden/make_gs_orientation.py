"""Writes gs_orientation_cases.json from the reference's gather-scatter unit
test (tests/unit/gather_scatter/test_gather_scatter_parallel.pf:38-903).

Each of its 7 cases builds two hex elements sharing one face, with element
1's vertices listed in a different orientation, and asserts that after
gs%op the value at an element-1 face point equals the value at the
corresponding element-2 point.  The fixture keeps the data only: point ids,
the two vertex lists and the expected (element-1 index, element-2 index)
pairs of the shared face at lx = 5, 0-based linear indices.

Run in a container that has the reference tree; the JSON is committed.
"""
import json
import os
import re

SRC = "/root/reference/tests/unit/gather_scatter/test_gather_scatter_parallel.pf"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "gs_orientation_cases.json")


def main():
    txt = open(SRC).read()
    subs = re.split(r"\n\s*@test\(npes=\[1\]\)\s*\n", txt)[1:]
    cases = []
    for sub in subs:
        name = re.search(r"subroutine (test_\w+)", sub).group(1)
        ids = {int(a): int(b) for a, b in
               re.findall(r"call p\((\d+)\)%set_id\((\d+)\)", sub)}
        els = re.findall(r"call m%add_element\((\d+), (\d+), ([^\n]*?)&\s*\n\s*([^\n]*?)\)\s*\n", sub)
        verts = []
        for _, _, l1, l2 in els:
            ps = [int(q) for q in re.findall(r"p\((\d+)\)", l1 + l2)]
            verts.append(ps)
        lx = int(re.search(r"lx = (\d+)", sub).group(1))
        expr = re.search(r"@assertEqual\(x%x\(([^)]*)\), x%x\(1,j,i,2\)", sub).group(1)
        pairs = []
        for i in range(1, lx + 1):
            for j in range(1, lx + 1):
                env = {"i": i, "j": j, "lx": lx}
                a = [eval(t.strip(), {}, env) for t in expr.split(",")]
                e1 = (a[0] - 1) + lx * ((a[1] - 1) + lx * (a[2] - 1))
                e2 = 0 + lx * ((j - 1) + lx * (i - 1)) + lx ** 3
                pairs.append([e1, e2])
        cases.append({"name": name, "lx": lx, "point_ids": ids,
                      "elements": verts, "expr_elem1": expr, "pairs": pairs})
    with open(OUT, "w") as f:
        json.dump({"source": "tests/unit/gather_scatter/test_gather_scatter_parallel.pf",
                   "cases": cases}, f, indent=1)
    print("wrote", OUT, len(cases), "cases")


if __name__ == "__main__":
    main()
