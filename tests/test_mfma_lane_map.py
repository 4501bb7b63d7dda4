"""CPU: the index arithmetic of the row kernels' lane and workgroup mappings.

* xcd_block (multigrid-ldg_amd/csrc/devutil.hpp) must be a permutation of the
  grid for every grid size (a row kernel indexed by it covers every row once),
  with each XCD (b mod 8) running one contiguous range of logical blocks.
* The batched MFMA operand layout (mfma.hip: LaneOff, blk_load, the
  epilogues): lane l = 16 q + n supplies A[16 rt + n][S q + s] and
  X[S q + s][col0 + NT n + t] in K step s and writes output columns
  col0 + NT n + t.  Emulating the 16x16x4 MFMA on those operands must give
  Y = A X for every block size, tile count and column count the launchers
  accept, with every load inside its block; a wrong offset (planted below)
  must not.
* The smoothers' D^-1 product (RowTile, dinv_times) reads the residual from
  the accumulators, K steps ordered by accumulator register: emulated, it
  must give D R for every block size in both register layouts.

The emulation restates the kernels' address formulas; the GPU parity tests
check the kernels themselves against the oracle."""
import itertools

import numpy as np
import pytest


def xcd_block(b, nb):
    q, r, x = nb >> 3, nb & 7, b & 7
    return x * q + min(x, r) + (b >> 3)


@pytest.mark.parametrize("nb", list(range(1, 130)) + [4273, 10040, 65536, 123457])
def test_xcd_block_is_a_permutation_with_contiguous_xcd_ranges(nb):
    out = [xcd_block(b, nb) for b in range(nb)]
    assert sorted(out) == list(range(nb))
    for x in range(8):
        lb = [out[b] for b in range(nb) if b % 8 == x]
        assert lb == list(range(lb[0], lb[0] + len(lb))) if lb else True


def pv(ts, B):
    """vector width of the panel layout (mfma.hip pv)"""
    if ts == 8:
        return 2 if B % 2 == 0 else 1
    return 4 if B % 4 == 0 else (2 if B % 2 == 0 else 1)


def panel(A, V):
    B = A.shape[0]
    out = np.zeros(B * B)
    for r in range(B):
        for c in range(B):
            out[((c // V) * B + r) * V + c % V] = A[r, c]
    return out


def fit_tiles(nt, k):
    while nt > 1 and k % nt:
        nt >>= 1
    return nt


def emulate(ts, B, NT, k, A, X, al_bias=0):
    """Y = A X through the lane mapping of mfma.hip (one block, every column group)"""
    S, RT, V = (B + 3) // 4, (B + 15) // 16, pv(ts, B)
    Ap, Xf = panel(A, V), X.reshape(-1)
    Y = np.full((B, k), np.nan)
    for g in range((k + 16 * NT - 1) // (16 * NT)):
        col0 = g * 16 * NT
        acc = np.zeros((RT, NT, 16, 16))
        for s in range(S):
            aop, xop = np.zeros((RT, 64)), np.zeros((NT, 64))
            for lane in range(64):
                n, q = lane & 15, lane >> 4
                xl = S * q * k + min(col0 + NT * n, k - NT)
                if 3 * S + s < B or S * q + s < B:
                    for t in range(NT):
                        off = s * k + xl + t
                        assert 0 <= off < B * k
                        xop[t, lane] = Xf[off]
                al = ((S * q // V) * B * V if S % V == 0 else 0) + n * V + al_bias
                for rt in range(RT):
                    rin = 16 * rt + 15 < B or 16 * rt + n < B
                    if S % V == 0:
                        s0 = s - s % V
                        if rin and (3 * S + s0 < B or S * q + s0 < B):
                            off = (s0 // V) * B * V + 16 * rt * V + al + s % V
                            assert 0 <= off < B * B
                            aop[rt, lane] = Ap[off]
                    else:
                        c, r = S * q + s, 16 * rt + n
                        if rin and (3 * S + s < B or c < B):
                            aop[rt, lane] = Ap[((c // V) * B + r) * V + c % V]
            # v_mfma 16x16x4: D[m][j] += sum_kk A(lane 16 kk + m) B(lane 16 kk + j)
            for rt in range(RT):
                for t in range(NT):
                    acc[rt, t] += aop[rt].reshape(4, 16).T @ xop[t].reshape(4, 16)
        for n in range(16):
            gc = col0 + NT * n
            if gc >= k:
                continue
            for rt in range(RT):
                for m in range(16):
                    if 16 * rt + m < B:
                        for t in range(NT):
                            Y[16 * rt + m, gc + t] = acc[rt, t, m, n]
    return Y


CASES = [(ts, B, nt, k) for ts, B, nt, k in itertools.product((4, 8), (6, 8, 9, 10, 16, 27), (1, 2, 4),
                                                              (1, 2, 3, 4, 16, 20, 32, 48, 96))
         if not (B > 16 and nt > 2)]


@pytest.mark.parametrize("ts,B,nt,k", CASES)
def test_mfma_lane_mapping_gives_the_block_product(ts, B, nt, k):
    rng = np.random.default_rng(B * 1000 + k)
    A, X = rng.standard_normal((B, B)), rng.standard_normal((B, k))
    NT = fit_tiles(nt, k)
    assert k % NT == 0
    Y = emulate(ts, B, NT, k, A, X)
    assert np.allclose(Y, A @ X, rtol=1e-12, atol=1e-12)


def test_a_planted_offset_error_is_caught():
    rng = np.random.default_rng(0)
    A, X = rng.standard_normal((16, 16)), rng.standard_normal((16, 32))
    with pytest.raises(AssertionError):
        Y = emulate(4, 16, 2, 32, A, X, al_bias=1)
        assert np.allclose(Y, A @ X)


def acc_row(ts, q, e):
    """row of the 16x16 accumulator in register e of a lane in group q (mfma.hip Mf::row)"""
    return 4 * q + e if ts == 4 else 4 * e + q


def emulate_dinv(ts, B, Dm, R, row=acc_row):
    """D R through the smoothers' D^-1 product (mfma.hip RowTile / dinv_times):
    R (B x 16 columns of one tile) sits in the accumulator layout, lane (q, n)
    register e of row tile rt2 holding R[16 rt2 + row(q, e)][n]; in K step e
    of K tile rt2 that register is the B operand and dv[rt2][e][rt] =
    D[16 rt + n][16 rt2 + row(q, e)] the A operand."""
    RT = (B + 15) // 16
    out = np.zeros((RT, 16, 16))
    for rt2 in range(RT):
        for e in range(4):
            aop, bop = np.zeros((RT, 64)), np.zeros(64)
            for lane in range(64):
                n, q = lane & 15, lane >> 4
                c = 16 * rt2 + row(ts, q, e)
                bop[lane] = R[c, n] if c < B else 0.0
                for rt in range(RT):
                    r = 16 * rt + n
                    aop[rt, lane] = Dm[r, c] if (r < B and c < B) else 0.0
            for rt in range(RT):
                out[rt] += aop[rt].reshape(4, 16).T @ bop.reshape(4, 16)
    return np.concatenate(list(out))[:B]


@pytest.mark.parametrize("ts,B", list(itertools.product((4, 8), (6, 8, 9, 10, 16, 27))))
def test_dinv_product_reads_the_residual_in_the_accumulators(ts, B):
    rng = np.random.default_rng(B)
    Dm, R = rng.standard_normal((B, B)), rng.standard_normal((B, 16))
    assert np.allclose(emulate_dinv(ts, B, Dm, R), Dm @ R, rtol=1e-12, atol=1e-12)


def test_dinv_product_with_the_other_accumulator_layout_is_caught():
    """A operand indexed by the FP64 register layout while the FP32 residual
    sits in the FP32 one: the emulation must see the wrong product"""
    rng = np.random.default_rng(1)
    Dm, R = rng.standard_normal((16, 16)), rng.standard_normal((16, 16))
    out = np.zeros((16, 16))
    for e in range(4):
        aop, bop = np.zeros(64), np.zeros(64)
        for lane in range(64):
            n, q = lane & 15, lane >> 4
            bop[lane] = R[acc_row(4, q, e), n]
            aop[lane] = Dm[n, acc_row(8, q, e)]
        out += aop.reshape(4, 16).T @ bop.reshape(4, 16)
    assert not np.allclose(out, Dm @ R)
