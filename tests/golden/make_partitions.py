"""Writes tests/golden/partitions.json: the agglomeration hierarchy that the
reference's MGridGen 1.0 (compiled from its sources by `make -C oracle ref`)
produces for the reference's meshes (data/*.mesh), with the reference's
parameters (src/unstructured/mesh/agglomeration.h:96-116: minsize 1,
maxsize 2^N, options {4, 6, 1, N}, at most 10 agglomerated meshes).

Each entry lists the partitions level by level: level l elements -> level l+1
agglomerates.  The dual graph is the oracle's restatement of
agglomeration.h:57-93 / 133-185 (face order: first appearance)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402

MESHES = ["circle", "circle2", "square_4", "square_8", "square_16", "square_32", "ball"]


def main():
    if not po.mgridgen_available():
        raise SystemExit("oracle/_ref/libmgridgen.so missing: run make -C oracle ref")
    out = {}
    for m in MESHES:
        pts, tris = po.read_mesh(os.path.join(HERE, "meshes", m + ".mesh"))
        pr = po.SxProblem(pts, tris, 1)
        out[m] = [[int(v) for v in q] for q in po.agglomerate(pr)]
    with open(os.path.join(HERE, "partitions.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
