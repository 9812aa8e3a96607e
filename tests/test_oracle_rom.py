"""The oracle's POD reconstruction (oracle/rom.py) against the algebra it implements."""
import numpy as np
import pytest

import oracle.rom as rom


@pytest.mark.parametrize("gsum", [0, 1])
def test_full_rank_basis_gives_the_snapshots_back(gsum):
    rng = np.random.default_rng(2 + gsum)
    ng, glen, nt = 3, 20, 6
    X = rng.standard_normal((ng * glen, nt)) + 5.0
    C, S, U, V, rk = rom.basis_from_snapshots(X, ng, gsum)
    for t in range(1, nt + 1):
        f = rom.pod_reconstruct(C, S, U, V, rk, ng * glen, ng, nt, t, gsum)
        assert np.allclose(f, X[:, t - 1], rtol=0, atol=1e-12), t
    # a wrong storage order is caught: V read r-major instead of t-major
    Vt = V.reshape(-1)
    if gsum == 1:
        Vbad = Vt.reshape(nt, rk[0]).T.ravel()
        assert not np.allclose(rom.pod_reconstruct(C, S, U, Vbad, rk, ng * glen, ng, nt, 2, gsum), X[:, 1])


def test_truncation_and_interpolation():
    rng = np.random.default_rng(9)
    ng, glen, nt = 2, 30, 8
    X = rng.standard_normal((ng * glen, nt))
    C, S, U, V, rk = rom.basis_from_snapshots(X, ng, 0, rank=2)
    assert rk.tolist() == [2, 2]
    # the rank-2 reconstruction is the best rank-2 approximation of each centred group block
    for g in range(ng):
        B = X[g * glen:(g + 1) * glen] - C[g * glen:(g + 1) * glen, None]
        u, s, vt = np.linalg.svd(B, full_matrices=False)
        B2 = (u[:, :2] * s[:2]) @ vt[:2]
        f = rom.pod_reconstruct(C, S, U, V, rk, ng * glen, ng, nt, 3, 0)
        assert np.allclose(f[g * glen:(g + 1) * glen] - C[g * glen:(g + 1) * glen], B2[:, 2], atol=1e-12)
    f3 = rom.pod_reconstruct(C, S, U, V, rk, ng * glen, ng, nt, 3, 0)
    f4 = rom.pod_reconstruct(C, S, U, V, rk, ng * glen, ng, nt, 4, 0)
    mid = rom.linear_interpolation(f3, f4, 0.25, 0.2, 0.3)
    assert np.allclose(mid, 0.5 * (f3 + f4), atol=1e-14)


@pytest.mark.parametrize("gsum", [0, 1])
def test_dmd_of_a_linear_system_reproduces_expm(gsum):
    """For dx/dt = A x with A = W diag(L) W^-1 (real A, so complex-conjugate modes), the
    DMD reconstruction with amplitudes B = W^-1 x0 is expm(A t) x0, group by group."""
    from scipy.linalg import expm
    rng = np.random.default_rng(11 + gsum)
    ng, glen = (1, 12) if gsum else (3, 6)
    C = rng.uniform(0.1, 0.2, ng * glen)
    Ls, Bs, Ws, ranks, want = [], [], [], [], {}
    As, x0s = [], []
    for g in range(ng):
        A = rng.standard_normal((glen, glen)) * 0.5
        x0 = rng.standard_normal(glen)
        lam, vec = np.linalg.eig(A)
        b = np.linalg.solve(vec, x0)
        Ls.append(lam)
        Bs.append(b)
        Ws.append(vec.T.ravel())                     # rank columns of glen values
        ranks.append(glen)
        As.append(A)
        x0s.append(x0)
    L, B, W = np.concatenate(Ls), np.concatenate(Bs), np.concatenate(Ws)
    for t in (0.0, 0.3, 1.1):
        f = rom.dmd_reconstruct(C, L, B, W, np.array(ranks), ng * glen, ng, t, gsum)
        ref = np.concatenate([expm(A * t) @ x0 for A, x0 in zip(As, x0s)]) + C
        assert np.allclose(f, ref, rtol=1e-10, atol=1e-10), t
