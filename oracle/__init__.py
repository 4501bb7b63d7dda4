"""CPU FP64 oracle for multigrid-ldg's LDG-Poisson MG-PCG path, on quadtrees
and on simplex meshes.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only the test suite,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` load
it.  The product library (``libmgldg.so``) never links or calls it.

The C restatement lives in ``oracle/cart.c`` (quadtrees, block-sparse
arithmetic, multigrid, PCG) and ``oracle/simplex.c`` (simplex meshes,
agglomerated hierarchy); this module is a ctypes binding with numpy
conveniences.  ``mgridgen``/``agglomerate`` call the reference's MGridGen as
compiled from its sources into ``oracle/_ref`` (``make -C oracle ref``).
"""
from .pyoracle import (  # noqa: F401
    lib, build, Quadtree, Problem, Multigrid, pcg, basis_tables, bsr_spmv, bsr_to_dense,
    MGParams, read_mesh, sx_master, sx_face_rule, SxProblem, SxMultigrid, mgridgen, mgridgen_available, agglomerate,
)
