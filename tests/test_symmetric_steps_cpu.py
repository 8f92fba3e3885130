"""The algebra behind the symmetric Gauss-Jordan steps, on the oracle's operator.

The step kernel computes only the tiles I <= J of each step, and reads the
others through M[j,i] = s_i s_j (D_i / D_j) M[i,j]: D is built from the
couplings of block 0 (D_{y+1} / D_y = A[y, y+1] / A[y+1, y]), and s = -1 on
the pivots already taken. This replays the kernel's tile schedule (the same
index and sign rules, on small tiles) in numpy and checks every block
inverse against a dense inverse of its Schur complement.
"""

import numpy as np
import pytest

from oracle import construct_A

OMEGA = 2 * np.pi * 299792458.0 / 1.55e-6


def _tile(M, I, J, tb):
    return M[I * tb:(I + 1) * tb, J * tb:(J + 1) * tb]


def _symmetric_gj(S, D, tb, row_sign=-1.0):
    """In-place tiled Gauss-Jordan inverse of S computing only the tiles
    I <= J per step, as the kernel does (row_sign: the sign of the X[p,J]
    mirror, -1 in the kernel)."""
    nt = S.shape[0] // tb
    X = S.copy()
    for p in range(nt):
        P = np.linalg.inv(_tile(X, p, p, tb))
        Y = np.full_like(X, np.nan)
        Dp = D[p * tb:(p + 1) * tb]

        def row(J):          # X[p, J]; J < p: -(D_J / D_p) X[J, p]^T
            if J < p:
                return row_sign * (D[J * tb:(J + 1) * tb][None, :] / Dp[:, None]) * _tile(X, J, p, tb).T
            return _tile(X, p, J, tb)

        def col(I):          # X[I, p]; I > p: (D_p / D_I) X[p, I]^T
            if I > p:
                return (Dp[None, :] / D[I * tb:(I + 1) * tb][:, None]) * _tile(X, p, I, tb).T
            return _tile(X, I, p, tb)

        for I in range(nt):
            for J in range(I, nt):
                if I == p and J == p:
                    V = P
                elif I == p:
                    V = P @ row(J)
                elif J == p:
                    V = -col(I) @ P
                else:
                    V = _tile(X, I, J, tb) - col(I) @ (P @ row(J))
                Y[I * tb:(I + 1) * tb, J * tb:(J + 1) * tb] = V
                if p == nt - 1 and I != J:   # the last step writes the lower tiles too
                    Y[J * tb:(J + 1) * tb, I * tb:(I + 1) * tb] = (
                        D[I * tb:(I + 1) * tb][None, :] / D[J * tb:(J + 1) * tb][:, None]) * V.T
        X = Y
    return X


@pytest.mark.parametrize("pol,npml", [("Ez", [3, 4]), ("Hz", [3, 4]), ("Hz", [0, 5]), ("Ez", [4, 0])])
def test_symmetric_steps_invert_every_schur_complement(pol, npml):
    nx, ny, tb = 9, 12, 4
    rng = np.random.default_rng(1)
    eps = 1 + 11.25 * (rng.random((nx, ny)) < 0.5)
    A = construct_A(OMEGA, [0, nx * 0.02], [0, ny * 0.02], eps, npml, pol, 1e-6)
    A = (A[0] if isinstance(A, tuple) else A).toarray()
    blk = lambda i, j: A[i * ny:(i + 1) * ny, j * ny:(j + 1) * ny]
    D = np.ones(ny, complex)
    for y in range(ny - 1):
        D[y + 1] = D[y] * A[y, y + 1] / A[y + 1, y]
    G = None
    for k in range(nx):
        S = blk(k, k) - (blk(k, k - 1) @ G @ blk(k - 1, k) if k else 0)
        Ds = D[:, None] * S
        assert np.allclose(Ds, Ds.T, rtol=0, atol=1e-12 * abs(Ds).max())
        Gk = _symmetric_gj(S, D, tb)
        G = np.linalg.inv(S)
        assert np.abs(Gk - G).max() <= 1e-12 * np.abs(G).max()


def test_wrong_mirror_rules_fail():
    """The same schedule with the X[p,J] mirror's sign flipped, or with D
    inverted, does not invert S: the check above tells the rules apart."""
    nx, ny, tb = 4, 12, 4
    rng = np.random.default_rng(2)
    eps = 1 + 11.25 * (rng.random((nx, ny)) < 0.5)
    A = construct_A(OMEGA, [0, nx * 0.02], [0, ny * 0.02], eps, [2, 4], "Ez", 1e-6)
    A = (A[0] if isinstance(A, tuple) else A).toarray()
    S = A[:ny, :ny]
    D = np.ones(ny, complex)
    for y in range(ny - 1):
        D[y + 1] = D[y] * A[y, y + 1] / A[y + 1, y]
    Gi = np.linalg.inv(S)
    assert np.abs(_symmetric_gj(S, D, tb) - Gi).max() <= 1e-12 * np.abs(Gi).max()
    assert np.abs(_symmetric_gj(S, D, tb, row_sign=1.0) - Gi).max() > 1e-6 * np.abs(Gi).max()
    assert np.abs(_symmetric_gj(S, 1 / D, tb) - Gi).max() > 1e-6 * np.abs(Gi).max()
