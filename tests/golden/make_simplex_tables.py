"""Extract the reference's literal simplex quadrature tables
(src/unstructured/dg/master.h:263-483: Quadrature<N,Q> sizes, nodes and
weights for N = 1, 2, 3) into tests/golden/simplex_quadrature.json.  The
tables are data (numbers), kept as a fixture so that tests can run where the
reference tree is absent.

Usage: python tests/golden/make_simplex_tables.py [/path/to/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "src/unstructured/dg/master.h")).read()
out = {}
for N, Q, body in re.findall(r"Quadrature<(\d),(\d+)>::nodes\s*=\s*\{(.*?)\};\n", src):
    nums = [float(v) for v in re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?", body)]
    out.setdefault(N, {}).setdefault(Q, {})["nodes"] = nums
for N, Q, body in re.findall(r"Quadrature<(\d),(\d+)>::weights\s*=\s*\{(.*?)\};", src):
    nums = [float(v) for v in re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?", body)]
    out.setdefault(N, {}).setdefault(Q, {})["weights"] = nums
json.dump({"source": "src/unstructured/dg/master.h:263-483", "rules": out},
          open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "simplex_quadrature.json"), "w"))
print({N: {Q: (len(v.get("nodes", [])), len(v.get("weights", []))) for Q, v in d.items()} for N, d in out.items()})
