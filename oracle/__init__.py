"""CPU FP64 oracle for multigrid-ldg's Cartesian LDG-Poisson MG-PCG path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only the test suite,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` load
it.  The product library (``libmgldg.so``) never links or calls it.

The C restatement lives in ``oracle/cart.c``; this module is a ctypes binding
with numpy conveniences.
"""
from .pyoracle import (  # noqa: F401
    lib, build, Quadtree, Problem, Multigrid, pcg, basis_tables, bsr_spmv,
    MGParams,
)
