"""The index arithmetic the hex-lines build relies on, checked on the CPU
against a direct computation of the level orders (numpy, small texts).

Hex level h holds the 4-bit digits (bits 4h .. 4h+3 of each code) in the
order of binary level 4h; the build derives them from the quad levels 2h
and 2h + 1 (2-bit digits in the orders of binary levels 4h and 4h + 2):
  * position p of node N with quad digit d0 goes to position
    e = qbase[4N + d0] + (d0's before p) of quad level 2h + 1, whose digit
    is the hex digit's low half;
  * the hex digit x = 4 d0 + d1 occurs before position P
    CHD[N][x] + (d1's before e_d0(P) at quad level 2h + 1)
              - (d1's before the start of child 4N + d0) times,
    with CHD[N][x] the x's before node N (a prefix of the symbol counts).
A mistake in any of the three tables (node starts, quad bases, CHD) or in
the order of the two quad digits fails one of these equalities.
"""
import numpy as np
import pytest


def _level_order(codes, L, l):
    key = codes >> (L - l) if l > 0 else np.zeros_like(codes)
    return codes[np.argsort(key, kind="stable")]


def _check(codes, L):
    n = len(codes)
    sigma = 1 << L
    C = np.concatenate([[0], np.cumsum(np.bincount(codes, minlength=sigma))]).astype(np.int64)
    for h in range(L // 4):
        qa = (_level_order(codes, L, 4 * h) >> (L - 4 * h - 2)) & 3
        qb = (_level_order(codes, L, 4 * h + 2) >> (L - 4 * h - 4)) & 3
        hexd = (_level_order(codes, L, 4 * h) >> (L - 4 * h - 4)) & 15
        nodes = 16 ** h
        S = np.array([C[min(N << (L - 4 * h), sigma)] for N in range(nodes + 1)])

        def size(level, M):  # codes of node M of binary level `level`
            sh = L - level
            return C[min((M + 1) << sh, sigma)] - C[min(M << sh, sigma)]

        def cdq(q, M, d):  # d's before node M of quad level q (binary level 2q)
            return sum(size(2 * q + 2, 4 * Mp + d) for Mp in range(M))

        qbase = {4 * N + d: C[(4 * N + d) << (L - 4 * h - 2)] - cdq(2 * h, N, d)
                 for N in range(nodes) for d in range(4)}
        CHD = np.zeros((nodes + 1, 16), dtype=np.int64)
        for N in range(1, nodes + 1):
            for x in range(16):
                CHD[N, x] = CHD[N - 1, x] + size(4 * h + 4, 16 * (N - 1) + x)
        pre_a = np.vstack([np.concatenate([[0], np.cumsum(qa == d)]) for d in range(4)])
        pre_b = np.vstack([np.concatenate([[0], np.cumsum(qb == d)]) for d in range(4)])
        pre_x = np.vstack([np.concatenate([[0], np.cumsum(hexd == x)]) for x in range(16)])
        N = 0
        for p in range(n):
            while N + 1 < nodes and S[N + 1] <= p:
                N += 1
            d0 = qa[p]
            e = qbase[4 * N + d0] + pre_a[d0, p]
            assert 4 * d0 + qb[e] == hexd[p], (h, p)
        for P in list(range(0, n + 1, 128)) + [n]:
            N0 = max(N for N in range(nodes) if S[N] <= P)
            for d0 in range(4):
                e = qbase[4 * N0 + d0] + pre_a[d0, P]
                for d1 in range(4):
                    got = CHD[N0, 4 * d0 + d1] + pre_b[d1, e] - cdq(2 * h + 1, 4 * N0 + d0, d1)
                    assert got == pre_x[4 * d0 + d1, P], (h, P, d0, d1)


@pytest.mark.parametrize("L,n,kind", [(4, 3000, "uniform"), (8, 6000, "uniform"),
                                      (8, 6011, "zipf"), (12, 5000, "uniform"),
                                      (16, 2000, "zipf")])
def test_hex_digits_and_counts_from_quad_levels(L, n, kind):
    rng = np.random.default_rng(L * 1000 + n)
    sigma = 1 << L
    if kind == "uniform":
        codes = rng.integers(0, sigma, n)
    else:
        codes = np.minimum(rng.zipf(1.2, n) - 1, sigma - 1)
    _check(codes.astype(np.int64), L)
