"""Extract the literal basis tables of the reference (src/cartesian/dg/master.h:150-214)
into tests/golden/master_tables.json.  The tables are data (numbers), kept as a
fixture so that tests can run where the reference tree is absent.

Usage: python tests/golden/make_master_tables.py [/path/to/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "src/cartesian/dg/master.h")).read()
out = {}
pat_mat = re.compile(r"template<> const KronMat<1,(\d)> (Master<1,\d>::mass|Master<1,\d>::invmass|GaussLobatto<\d>::diff) = \(KronMat<1,\d>\(\) << ([^)]*)\)\.finished\(\);")
for P, name, body in pat_mat.findall(src):
    kind = "mass" if name.endswith("::mass") else "invmass" if name.endswith("invmass") else "diff"
    out.setdefault(kind, {})[P] = [float(v) for v in body.split(",")]
pat_arr = re.compile(r"template<> const double (GaussLobatto|LagrangePoly|Quadrature)<(\d+)>::(nodes|denom|weights)\[\]\s*=\s*\{([^}]*)\};")
for cls, P, field, body in pat_arr.findall(src):
    key = {"GaussLobatto": "gl_nodes", "LagrangePoly": "denom"}.get(cls, "quad_" + field)
    out.setdefault(key, {})[P] = [float(v) for v in body.split(",")]
# the 32-digit literals as written, for the record of what was parsed
json.dump({"source": "src/cartesian/dg/master.h:150-214", "tables": out},
          open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "master_tables.json"), "w"), indent=0)
print({k: sorted(v.keys(), key=int) for k, v in out.items()})
