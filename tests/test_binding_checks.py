"""The Python binding rejects device arguments of the wrong type, length or
shape before any call reaches the library (CPU-only: every case raises before
a launch, so no GPU is touched).  A float32 right-hand side handed to an FP64
entry point, or a vector of the wrong length, would otherwise be read or
written out of bounds on the device."""
import types

import numpy as np
import pytest
import torch

import mgldg
from mgldg import MGLDGError


def _fake_mg(ndof=40, nqb=12, rows=(10, 3), B=(4, 4)):
    mg = object.__new__(mgldg.Multigrid)
    mg.h = None
    mg.ctx = None
    mg.prob = types.SimpleNamespace(ndof=ndof, nqb=nqb)
    mg.nlevels = len(rows)
    mg.level_rows, mg.level_B = list(rows), list(B)
    return mg


def test_pcg_rejects_wrong_dtype_and_length():
    mg = _fake_mg()
    with pytest.raises(MGLDGError, match="dtype float64"):
        mgldg.pcg(mg, torch.zeros(40, dtype=torch.float32))
    with pytest.raises(MGLDGError, match="shape"):
        mgldg.pcg(mg, torch.zeros(39, dtype=torch.float64))
    with pytest.raises(MGLDGError, match="x0: expected 80 elements"):
        mgldg._ptr(torch.zeros(40, dtype=torch.float64), 80, name="x0")
    # host path: shape checked before the library is called
    with pytest.raises(MGLDGError, match="shape"):
        mgldg.pcg(mg, np.zeros(41))
    with pytest.raises(MGLDGError, match="x0"):
        mgldg.pcg(mg, np.zeros(40), np.zeros(41))


def test_device_tensor_required():
    mg = _fake_mg()
    with pytest.raises(MGLDGError, match="device tensor"):
        mgldg.pcg(mg, torch.zeros(40, dtype=torch.float64))
    with pytest.raises(MGLDGError, match="contiguous"):
        mgldg.pcg(mg, torch.zeros(2, 40, dtype=torch.float64).t())


def test_unit_hooks_check_level_length_and_precision():
    mg = _fake_mg()
    x64 = torch.zeros(40, dtype=torch.float64)
    x32 = torch.zeros(40, dtype=torch.float32)
    with pytest.raises(MGLDGError, match="dtype float32"):
        mg.spmv(0, x64, x64, precision="mixed")
    with pytest.raises(MGLDGError, match="dtype float64"):
        mg.smooth(0, x32, x32)
    with pytest.raises(MGLDGError, match="expected 12 elements"):
        mg.spmv(1, x64, x64)
    with pytest.raises(MGLDGError, match="level 2 out of range"):
        mg.spmv(2, x64, x64)
    with pytest.raises(MGLDGError, match="dtype float64"):
        mg.vcycle(x32, x32, precision="mixed")


def test_rhs_adjoint_meanzero_checks():
    mg = _fake_mg()
    with pytest.raises(MGLDGError, match="both None"):
        mgldg.rhs_device(mg)
    with pytest.raises(MGLDGError, match="f: expected shape"):
        mgldg.rhs_device(mg, torch.zeros(12, dtype=torch.float64))
    with pytest.raises(MGLDGError, match="g: expected shape"):
        mgldg.rhs_device(mg, None, torch.zeros(40, dtype=torch.float64))
    with pytest.raises(MGLDGError, match="dtype float64"):
        mgldg.rhs_device(mg, torch.zeros(40, dtype=torch.float32))
    with pytest.raises(MGLDGError, match="dLdu"):
        mgldg.adjoint(mg, torch.zeros(41, dtype=torch.float64))
    with pytest.raises(MGLDGError, match="lam"):
        mgldg.adjoint_maps(mg, torch.zeros(40, dtype=torch.float32))
    with pytest.raises(MGLDGError, match="dtype float64"):
        mgldg.mean_zero(None, torch.zeros(40, dtype=torch.float32))


def test_require_converged():
    ok = dict(status="converged", iterations=3, relres_true=1e-11)
    bad = dict(status="maxit", iterations=200, relres_true=1e-3)
    mgldg._require_converged("x", ok)
    mgldg._require_converged("x", [ok, ok])
    with pytest.raises(MGLDGError, match="column 1 maxit"):
        mgldg._require_converged("forward solve", [ok, bad])
