"""Oracle of the multicolour ILU(0) preconditioner (MSG123_PRECON_MCILU32).

Test infrastructure only: the product never imports this module.

The reference's preconditioner is D-ILU(0) in natural order
(precon_dilu, solve_ilu; Linear_solution.f90:742-855): pivots
d_i = a_ii - sum_{k < i} a_ik a_ki / d_k and two serial sweeps.  This module
states the same ILU(0) in the red-black ordering of the 7-point grid (red:
x+y+z even, ordered first), where every neighbour of a cell has the other
colour:
    red pivot    d_r = a_rr
    black pivot  d_b = a_bb - sum_k a_bk a_kb / a_kk    (k in the order up, N, W, E, S, down)
    black pass   z_b = (r_b - sum_k a_bk (r_k / a_kk)) / d_b
    red pass     z_r = (r_r - sum_k a_rk z_k) / a_rr
`factor` computes the pivots in FP64; `apply` runs the two passes either in
FP64 (the mathematical preconditioner, pinned by the dense checks in
tests/test_oracle.py: (D+L) D^-1 (D+U) equals A on A's pattern, and apply
equals a dense solve with that M) or in FP32 with the device's operation
order (1/pivot and couplings rounded to FP32, r rounded to FP32, each term
a * (r_k * inv_k) or a * z_k subtracted in neighbour order, then one
multiply by the FP32 inverse pivot), which the GPU reproduces bit for bit.

The matrix is given as seven diagonals [7][N] in CSR neighbour order (diag,
up, N, W, E, S, down), zero where a neighbour is missing.
"""
from __future__ import annotations

import numpy as np

# neighbour slot s (1..6) -> (index offset kind, opposite slot)
_OPP = {1: 6, 2: 5, 3: 4, 4: 3, 5: 2, 6: 1}


def _geometry(nx, ny, nz):
    n = nx * ny * nz
    c = np.arange(n)
    x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
    colour = (x + y + z) & 1
    nxy = nx * ny
    nbr = [(z > 0, c - nxy), (y > 0, c - nx), (x > 0, c - 1), (x < nx - 1, c + 1),
           (y < ny - 1, c + nx), (z < nz - 1, c + nxy)]
    return colour, nbr


def factor(A7, nx, ny, nz):
    """FP64 pivots: a_rr for red cells, a_bb - sum_k a_bk a_kb / a_kk for
    black cells, subtracted in the order up, N, W, E, S, down."""
    A7 = np.asarray(A7, dtype=np.float64)
    colour, nbr = _geometry(nx, ny, nz)
    piv = A7[0].copy()
    blk = colour == 1
    for s, (ok, k) in enumerate(nbr, start=1):
        m = blk & ok
        kk = k[m]
        piv[m] = piv[m] - A7[s][m] * A7[_OPP[s]][kk] / A7[0][kk]
    return piv


def apply(A7, piv, r, nx, ny, nz, fp32=True):
    """z = M^-1 r by the black pass then the red pass."""
    A7 = np.asarray(A7, dtype=np.float64)
    colour, nbr = _geometry(nx, ny, nz)
    if fp32:
        ft = np.float32
        inv = (1.0 / piv).astype(ft)
        a = A7.astype(ft)
        rr = np.asarray(r, dtype=np.float64).astype(ft)
    else:
        ft = np.float64
        inv = 1.0 / piv
        a = A7
        rr = np.asarray(r, dtype=np.float64)
    blk = colour == 1
    red = ~blk
    t = rr.copy()
    for s, (ok, k) in enumerate(nbr, start=1):
        m = blk & ok
        kk = k[m]
        t[m] = t[m] - a[s][m] * (rr[kk] * inv[kk])
    zb = np.zeros_like(rr)
    zb[blk] = t[blk] * inv[blk]
    t = rr.copy()
    for s, (ok, k) in enumerate(nbr, start=1):
        m = red & ok
        kk = k[m]
        t[m] = t[m] - a[s][m] * zb[kk]
    z = zb.copy()
    z[red] = t[red] * inv[red]
    return z.astype(np.float64)


def dense_matrices(A7, piv, nx, ny, nz):
    """Dense A and M = (D+L) D^-1 (D+U) in natural indexing, L and U the
    couplings of A from a red row to nothing (reds come first) and from a
    black row to its red neighbours (L), from a red row to its black
    neighbours (U)."""
    n = nx * ny * nz
    colour, nbr = _geometry(nx, ny, nz)
    A = np.diag(np.asarray(A7[0], dtype=np.float64))
    L = np.zeros((n, n))
    U = np.zeros((n, n))
    for s, (ok, k) in enumerate(nbr, start=1):
        rows = np.nonzero(ok)[0]
        A[rows, k[rows]] = A7[s][rows]
        lo = rows[colour[rows] == 1]
        up = rows[colour[rows] == 0]
        L[lo, k[lo]] = A7[s][lo]
        U[up, k[up]] = A7[s][up]
    D = np.diag(piv)
    M = (D + L) @ np.diag(1.0 / piv) @ (D + U)
    return A, M


def from_csr(offind, offrow, dirc, dmat, lumat):
    """Seven diagonals [7][N] from the reference's CSR (offind row ends,
    offrow 1-based columns, dir_conn 1..6)."""
    n = len(dmat)
    A7 = np.zeros((7, n))
    A7[0] = dmat
    for i in range(n):
        for k in range(offind[i], offind[i + 1]):
            A7[dirc[k], i] = lumat[k]
    return A7
