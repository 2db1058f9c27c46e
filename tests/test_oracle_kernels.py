"""The CPU checker's building blocks pinned independently of the GPU:
Thomas against dense LU, the segment scan exhaustively against a literal
restatement of the reference loop, and the run rule the GPU uses."""
import itertools

import numpy as np
import pytest

import oracle


def test_thomas_matches_dense_lu():
    rng = np.random.default_rng(1)
    for n in list(range(1, 12)) + [40, 120]:
        A = rng.normal(size=n); C = rng.normal(size=n)
        B = 3.0 + np.abs(A) + np.abs(C) + rng.random(n)
        D = rng.normal(size=n)
        x = oracle.thomas(A, B, C, D)
        T = np.diag(B) + np.diag(A[1:], -1) + np.diag(C[:-1], 1)
        ref = np.linalg.solve(T, D)
        assert np.allclose(x, ref, rtol=1e-12, atol=1e-14), n


def test_thomas_known_answer_and_order():
    rng = np.random.default_rng(2)
    n = 50
    A = rng.normal(size=n); C = rng.normal(size=n); B = 4 + np.abs(A) + np.abs(C)
    T = np.diag(B) + np.diag(A[1:], -1) + np.diag(C[:-1], 1)
    D = T @ np.ones(n)
    assert np.allclose(oracle.thomas(A, B, C, D), 1.0, rtol=1e-13)
    # swapping the off-diagonals must break it (catches a transposed operand)
    assert not np.allclose(oracle.thomas(C, B, A, D), 1.0, rtol=1e-6)
    # a single row: x = D/B exactly
    assert oracle.thomas([0.0], [4.0], [0.0], [3.0])[0] == 0.75


def literal_scan(kh, lim):
    """support_funcs.cu:252-279, statement by statement."""
    dau, cuoi = [], []
    start, end = 2, 0
    while start < lim:
        if kh[start] != 0:
            while kh[start] and start < lim:
                start += 1
        if start + 1 == lim:
            start = lim
        if kh[start] == 0 and start + 1 < lim:
            dau.append(start)
            end = start
            while kh[end] == 0 and end < lim:
                end += 1
            if kh[end] != 0 and end <= lim:
                cuoi.append(end - 1)
                start = end
            else:
                cuoi.append(lim)
                start = lim
    return dau, cuoi


def run_rule(kh, lim):
    """The run rule the GPU kernel applies with ballots: a wet run inside
    [2, lim] starting at a <= lim-2 is a segment ending at its last wet cell."""
    wet = lambda x: 2 <= x <= lim and kh[x] == 0
    dau = [j for j in range(2, lim + 1) if wet(j) and (j == 2 or not wet(j - 1)) and j <= lim - 2]
    cuoi = [j for j in range(2, lim + 1) if wet(j) and not wet(j + 1) and
            (j <= lim - 2 or (j == lim - 1 and wet(lim - 2)) or (j == lim and wet(lim - 1) and wet(lim - 2)))]
    return dau, cuoi


@pytest.mark.parametrize("lim", [4, 5, 6, 7, 8, 9, 10, 11])
def test_segment_scan_exhaustive(lim):
    for bits in itertools.product((0, 2), repeat=lim + 2):
        kh = np.array(bits + (2,), dtype=np.int32)
        n, d, c = oracle.scan_line(kh, lim, SL=20)
        ref_d, ref_c = literal_scan(list(kh), lim)
        assert n == len(ref_d) and list(d[:n]) == ref_d and list(c[:n]) == ref_c, bits
        rd, rc = run_rule(list(kh), lim)
        assert rd == ref_d and rc == ref_c, bits


def test_segment_overflow_is_an_error():
    lim = 60
    kh = np.array([2, 2] + [0, 2] * 30 + [2, 2], dtype=np.int32)
    n, _, _ = oracle.scan_line(kh, lim, SL=5)
    assert n < 0
