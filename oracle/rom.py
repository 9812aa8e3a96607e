"""POD reconstruction restated in numpy (test infrastructure only).

POD_RECONSTRUCT (MainSrc/POD_ROUTINES.f08:372-432): dat = sum over the rank of
S * U * V at data time t, then the centering vector C added; PODgsum 0 keeps one basis
per group (U: per group rank(g) columns of glen = len/N_g values, concatenated; S and V
concatenated over the groups, V of a group laid out t-major, rank(g) values per data
time), PODgsum 1 one basis for the whole vector.  LINEAR_INTERPOLATION
(MainSrc/GRID_FUNCTIONS.f08:819-836): y1 + (x - x1)(y2 - y1)/(x2 - x1) per element.

DMD_RECONSTRUCT (MainSrc/DMD_ROUTINES.f08:28-103): the real part of sum_r B_r W_r
exp(L_r t), accumulated as (B W) exp(L t) in the loop order of :48-92, plus C.

Parity pinning: tests/test_oracle_rom.py checks that a full-rank basis built from a
snapshot matrix by numpy's SVD gives every snapshot back (which fixes the storage orders
of U, S and V against the algebra X - C = U diag(S) V^T) and that interpolation at a
midpoint gives the mean of the two snapshots; and that DMD modes and amplitudes of a
linear system dx/dt = A x (the eigendecomposition of A) reproduce scipy's expm(A t) x0.
"""
from __future__ import annotations

import numpy as np


def pod_reconstruct(C, S, U, V, rank, len_, ng, nt, t, gsum):
    """Field at data time t (1-based), the loop order of POD_ROUTINES.f08:386-431."""
    dat = np.zeros(len_)
    if gsum == 0:
        glen = len_ // ng
        p2 = 0
        p3 = 0
        for g in range(ng):
            rk = int(rank[g])
            p4 = nt * p3 + (t - 1) * rk
            for r in range(rk):
                blk = slice(g * glen, (g + 1) * glen)
                dat[blk] = dat[blk] + (S[p3 + r] * U[p2:p2 + glen]) * V[p4 + r]
                p2 += glen
            p3 += rk
    else:
        rk = int(rank[0])
        for r in range(rk):
            dat = dat + (S[r] * U[r * len_:(r + 1) * len_]) * V[r + (t - 1) * rk]
    return dat + C


def linear_interpolation(y1, y2, x, x1, x2):
    return y1 + (x - x1) * (y2 - y1) / (x2 - x1)


def basis_from_snapshots(X, ng, gsum, rank=None):
    """(C, S, U, V, ranks) in the reference's storage from snapshots X (len x N_t): the
    mean over time as C and the SVD of X - C, per group or whole."""
    len_, nt = X.shape
    C = X.mean(axis=1)
    blocks = [X - C[:, None]] if gsum else np.split(X - C[:, None], ng, axis=0)
    Ss, Us, Vs, ranks = [], [], [], []
    for B in blocks:
        u, s, vt = np.linalg.svd(B, full_matrices=False)
        rk = len(s) if rank is None else min(rank, len(s))
        Ss.append(s[:rk])
        Us.append(u[:, :rk].T.ravel())              # rank columns of the block, one after the other
        Vs.append(vt[:rk, :].T.ravel())             # t-major: rk values per data time
        ranks.append(rk)
    return C, np.concatenate(Ss), np.concatenate(Us), np.concatenate(Vs), np.array(ranks, dtype=np.int32)


def dmd_reconstruct(C, L, B, W, rank, len_, ng, t, gsum):
    """Field at time t (DMD_ROUTINES.f08:48-96)."""
    dat = np.zeros(len_, dtype=np.complex128)
    if gsum == 0:
        glen = len_ // ng
        for g in range(ng):
            rk = int(rank[g])
            p3 = int(np.sum(rank[:g]))
            p2 = p3 * glen
            for r in range(rk):
                e = np.exp(L[p3 + r] * t)
                blk = slice(g * glen, (g + 1) * glen)
                dat[blk] = dat[blk] + (B[p3 + r] * W[p2:p2 + glen]) * e
                p2 += glen
    else:
        for r in range(int(rank[0])):
            e = np.exp(L[r] * t)
            dat = dat + (B[r] * W[r * len_:(r + 1) * len_]) * e
    return dat.real + C


def split_bcg(bc, nx, ny, ng):
    """Cg_L, Cg_B, Cg_R, Cg_T (Fortran shapes) from the boundary data vector of
    POD_RECONSTRUCT_BCg (POD_ROUTINES.f08:279-352): per group L (N_y), B (N_x), R (N_y),
    T (N_x), on the simulation grid."""
    per = 2 * (nx + ny)
    L, B, R, T = (np.zeros((ny, ng)), np.zeros((nx, ng)), np.zeros((ny, ng)), np.zeros((nx, ng)))
    for g in range(ng):
        v = bc[g * per:(g + 1) * per]
        L[:, g], B[:, g], R[:, g], T[:, g] = v[:ny], v[ny:ny + nx], v[ny + nx:2 * ny + nx], v[2 * ny + nx:]
    return L, B, R, T
