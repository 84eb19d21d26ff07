"""Write tests/golden/mbe_zip_nodes.json: maximal-biclique counts AND
processing-node counts of the reference's own ImbeaAdvFinder and
PmbeAdvFinder (baselines/MBE.zip -s 1 and -s 2, built by oracle/build_ref.sh
into oracle/_ref/mbe_ref) on small seeded graphs.

Every graph has at most 16 vertices on each side.  Both finders order
vertices with std::sort by a key that can tie (PMBE: the degree of the
vertices one step of its containment order releases; iMBEA: the overlap of
candidates, and equal neighbour lists in its vertex order), and std::sort
leaves ties in an unspecified order; libstdc++ sorts 16 or fewer elements by
insertion, which keeps them in their incoming order.  With at most 16
vertices on the root side the reference's traversal is therefore fixed, and
the oracle's node counts must equal it exactly.  The graphs are regenerated
by tests/test_oracle.py from the recipe stored next to each count.

Run in the development container (it needs the reference checkout to build
mbe_ref); the JSON it writes is the committed fixture.

    python tests/golden/make_mbe_zip_nodes.py
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import graphs  # noqa: E402
from gmbe_amd import BiGraph  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "mbe_ref")


def recipes():
    out = {}
    for i in range(60):
        out[f"random_{i}"] = {"kind": "random", "seed": 20_000 + i}
        out[f"nested_{i}"] = {"kind": "nested", "seed": 30_000 + i}
    return out


def run_ref(path, sel):
    text = subprocess.run([REF, "-i", path, "-s", str(sel)], check=True, capture_output=True,
                          text=True, timeout=600).stdout
    m = re.search(r"maximal nodes/processing nodes : (\d+)/(\d+)", text)
    return int(m.group(1)), int(m.group(2))


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/mbe_ref missing: run oracle/build_ref.sh")
    res = {}
    with tempfile.TemporaryDirectory() as d:
        for name, rec in recipes().items():
            nU, rows = graphs.pmbe_recipe_rows(rec)
            path = os.path.join(d, name + ".adj")
            BiGraph.from_rows(nU, rows).write_adj(path)
            ent = {"recipe": rec}
            for sel in (1, 2):
                maximal, nodes = run_ref(path, sel)
                ent[f"s{sel}"] = {"maximal": maximal, "nodes": nodes}
            res[name] = ent
    with open(os.path.join(HERE, "mbe_zip_nodes.json"), "w") as f:
        json.dump({"source": "baselines/MBE.zip ImbeaAdvFinder (-s 1) and PmbeAdvFinder (-s 2) "
                             "built by oracle/build_ref.sh",
                   "graphs": res}, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {len(res)} graphs")


if __name__ == "__main__":
    main()
