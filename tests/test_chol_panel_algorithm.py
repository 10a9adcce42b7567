"""CPU restatement of the one-workgroup Cholesky inverse (k_chol_inv, and
k_chol_inv_la, which applies the same row operations to each tile in the same
order, with the panel eliminations on waves of their own): the
panel LDL^T elimination that carries L^-1 in the strictly lower triangle,
with the kernel's tile ownership and skip rules (16 x 16 tiles, upper tiles
hold S, lower tiles M = L^-1, the diagonal tile both; rank-PB updates of the
tiles below each panel).  Checked against numpy's Cholesky at the sizes and
panel boundaries the GPU kernel meets; the GPU kernel itself is compared
with the oracle in tests/test_scf_gpu.py."""
import numpy as np
import pytest


def chol_inv_tiles(S, NT, PB):
    n = S.shape[0]
    NM = 16 * NT
    assert n <= NM and PB % 4 == 0 and 16 % PB == 0
    up = {}  # (I, J) for J >= I: upper copy; lo: (I, J) for J <= I: lower copy
    lo = {}
    for I in range(NT):
        for J in range(NT):
            if J >= I:
                t = np.zeros((16, 16))
                for r in range(16):
                    for c in range(16):
                        i, x = 16 * I + r, 16 * J + c
                        if i <= x < n:
                            t[r, c] = S[i, x]
                up[I, J] = t
            if J <= I:
                lo[I, J] = np.zeros((16, 16))

    def element(i, x):  # combined array: upper part from upper tiles
        I, J = i // 16, x // 16
        return up[I, J][i % 16, x % 16] if x >= i else lo[I, J][i % 16, x % 16]

    for k0 in range(0, n, PB):
        pn = min(PB, n - k0)
        raw = np.array([[element(k0 + p, x) if k0 + p < n and x < n else 0.0 for x in range(NM)] for p in range(PB)])
        R = raw.copy()
        d = np.ones(PB)
        for q in range(pn):  # elimination inside the panel
            d[q] = R[q, k0 + q]
            assert d[q] > 0
            for cc in range(q + 1, pn):
                f = R[q, k0 + cc] / d[q]
                kq, kc = k0 + q, k0 + cc
                for y in range(n):
                    e = R[q, y] if (y < kq or y >= kc) else (1.0 if y == kq else 0.0)
                    R[cc, y] -= f * e
        RM = np.zeros_like(R)  # lower-part view: R before k_q, 1 at k_q, 0 after
        for q in range(pn):
            for y in range(n):
                RM[q, y] = R[q, y] if y < k0 + q else (1.0 if y == k0 + q else 0.0)
        for (I, J), T in list(up.items()) + list(lo.items()):
            upper = (I, J) in up and T is up[I, J]
            if 16 * I + 15 < k0 + pn or 16 * I >= n or 16 * J >= n:
                continue
            if not upper and 16 * J > k0 + PB - 1:
                continue  # lower tiles right of the panel stay unchanged
            F = np.zeros((16, PB))
            for m in range(16):
                i = 16 * I + m
                for q in range(pn):
                    if k0 + pn <= i < n:
                        F[m, q] = -R[q, i] / d[q]
            E = (R if upper else RM)[:, 16 * J:16 * J + 16]
            T += F @ E
        for (I, J), T in list(up.items()) + list(lo.items()):
            if I != k0 // 16:
                continue
            for r in range(16):
                for c in range(16):
                    i, x = 16 * I + r, 16 * J + c
                    if 0 <= i - k0 < pn and x < n:
                        T[r, c] = R[i - k0, x]
    dg = np.array([up[i // 16, i // 16][i % 16, i % 16] for i in range(n)])
    X = np.zeros((n, n))
    for c in range(n):
        X[c, c] = 1.0 / np.sqrt(dg[c])
        for r in range(c):
            X[r, c] = lo[c // 16, r // 16][c % 16, r % 16] / np.sqrt(dg[c])
    return X


@pytest.mark.parametrize("NT,PB,n", [(2, 8, 1), (2, 8, 7), (2, 8, 9), (2, 8, 17), (2, 8, 32),
                                      (4, 8, 33), (4, 8, 64), (2, 4, 30), (2, 16, 31)])
def test_panel_cholesky_inverse_matches_numpy(NT, PB, n):
    rng = np.random.default_rng(n + 7 * PB)
    A = rng.standard_normal((n, n + 3))
    S = A @ A.T + 0.1 * np.eye(n)
    X = chol_inv_tiles(S, NT, PB)
    U = np.linalg.cholesky(S).T
    assert np.abs(X - np.linalg.inv(U)).max() < 1e-10
    assert np.abs(np.tril(X, -1)).max() == 0.0
