"""Edge-element core of the oracle: local matrices, portrait, assembly, Dirichlet conditions.

Restates ``A2/T_Brick.cpp:142-540`` (closed-form 12x12 curl-curl ``B`` and mass ``C`` of an
axis-aligned brick), ``A2/T_Portrait.cpp:45-122`` (lower-triangular RSF pattern),
``A2/T_Global_SLAE.cpp:76-215`` (scatter of element matrices into ``di``/``ggl``) and
``A2/T_Bound_Cond.cpp:58-122`` (boundary edges and their symmetric elimination).

The local matrices are built here from 1D integrals of the basis (a tensor-product
derivation), not from the reference's unrolled expressions, so the two can check each other.
Local edge a has direction d=a//4 and transverse bits (see ``A2/T_Brick.h:203-218``):
x-edges carry (y,z) bits, y-edges (z,x) bits, z-edges (x,y) bits, low bit first.
The basis of edge a is (2/h_d) * prod_{axis != d} l_bit(s_axis) e_d with l0=1-s, l1=s on the
normalized coordinate s in [0,1] (``A2/T_Brick.cpp:741-836``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# transverse axes of each direction, low bit first
_TRANS = {0: (1, 2), 1: (2, 0), 2: (0, 1)}


def edge_bits(a: int) -> dict:
    """Map axis -> bit for the transverse axes of local edge a."""
    d = a // 4
    k = a % 4
    p, q = _TRANS[d]
    return {p: k & 1, q: (k >> 1) & 1}


_M1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])   # int l_i l_j
_DL = np.array([-1.0, 1.0])                        # l_i'


def _int1(fa, fb):
    """1D integral over [0,1] of two factors; a factor is ('c',), ('v',i) or ('d',i)."""
    ka, kb = fa[0], fb[0]
    if ka == "c" and kb == "c":
        return 1.0
    if ka == "c":
        fa, fb = fb, fa
        ka, kb = kb, ka
    if kb == "c":
        return 0.5 if ka == "v" else _DL[fa[1]]
    if ka == "v" and kb == "v":
        return _M1[fa[1], fb[1]]
    if ka == "d" and kb == "d":
        return _DL[fa[1]] * _DL[fb[1]]
    if ka == "v":
        return 0.5 * _DL[fb[1]]
    return 0.5 * _DL[fa[1]]


def _factors(a: int, deriv_axis: int | None):
    d = a // 4
    bits = edge_bits(a)
    f = []
    for ax in range(3):
        if ax == d:
            f.append(("c",))
        elif ax == deriv_axis:
            f.append(("d", bits[ax]))
        else:
            f.append(("v", bits[ax]))
    return f


_EPS = np.zeros((3, 3, 3))
_EPS[0, 1, 2] = _EPS[1, 2, 0] = _EPS[2, 0, 1] = 1
_EPS[0, 2, 1] = _EPS[2, 1, 0] = _EPS[1, 0, 2] = -1


def _coeff_tables():
    """Coefficients of C_ab and B_ab as sums of monomials in (hx, hy, hz).

    C_ab = cC[a,b] * 4*hx*hy*hz/h_d^2 (same direction only).
    B_ab = sum_{j,k} cB[a,b,j,k] * 4*hx*hy*hz/(h_da*h_db*h_j*h_k).
    """
    cC = np.zeros((12, 12))
    cB = np.zeros((12, 12, 3, 3))
    for a in range(12):
        for b in range(12):
            da, db = a // 4, b // 4
            if da == db:
                fa, fb = _factors(a, None), _factors(b, None)
                cC[a, b] = np.prod([_int1(fa[ax], fb[ax]) for ax in range(3)])
            for i in range(3):
                for j in range(3):
                    for k in range(3):
                        e = _EPS[i, j, da] * _EPS[i, k, db]
                        if e == 0 or j == da or k == db:
                            continue
                        fa, fb = _factors(a, j), _factors(b, k)
                        cB[a, b, j, k] += e * np.prod([_int1(fa[ax], fb[ax]) for ax in range(3)])
    return cC, cB


_CC, _CB = _coeff_tables()


def local_matrices(h: np.ndarray):
    """(B, C) for bricks of sizes h (nel, 3) -> two (nel, 12, 12) arrays, B without 1/mu."""
    h = np.atleast_2d(np.asarray(h, float))
    vol = h[:, 0] * h[:, 1] * h[:, 2]
    dirs = np.arange(12) // 4
    hd = h[:, dirs]                                           # (nel, 12)
    C = _CC[None] * (4 * vol[:, None, None] / (hd[:, :, None] * hd[:, None, :]))
    B = np.zeros_like(C)
    for j in range(3):
        for k in range(3):
            cb = _CB[:, :, j, k]
            if not cb.any():
                continue
            B += cb[None] * (4 * vol[:, None, None]
                             / (hd[:, :, None] * hd[:, None, :] * h[:, j, None, None] * h[:, k, None, None]))
    return B, C


def element_sizes(model) -> np.ndarray:
    """hx,hy,hz from vertices 0 and 7 (``A2/T_Brick.cpp:113-122``)."""
    p0 = model.xyz[model.nver[:, 0]]
    p7 = model.xyz[model.nver[:, 7]]
    return p7 - p0


def portrait(ed: np.ndarray, n: int):
    """Lower RSF pattern (ig, jg): strictly lower, sorted, unique (``T_Portrait.cpp:45-122``)."""
    a = np.repeat(ed, 12, axis=1).reshape(-1)
    b = np.tile(ed, (1, 12)).reshape(-1)
    m = b < a
    key = np.unique(a[m].astype(np.int64) * n + b[m])
    rows = key // n
    cols = key % n
    ig = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ig, rows + 1, 1)
    ig = np.cumsum(ig)
    return ig, cols.astype(np.int64)


def assemble(model, ig, jg, sigma=None):
    """B (curl-curl / mu) and C (sigma mass) in lower RSF: (b_ggl, b_di, c_ggl, c_di).

    ``sigma`` (per material) replaces the model's conductivity in C; sigma - sigma0 gives the
    mass matrix of the primary/secondary term (``T_Brick`` case 4, ``A2/T_Brick.cpp:518-526``).
    """
    n = len(ig) - 1
    h = element_sizes(model)
    B, C = local_matrices(h)
    B = B / model.mu[model.nvkat][:, None, None]
    sig = model.sigma if sigma is None else np.asarray(sigma, float)
    C = C * sig[model.nvkat][:, None, None]
    out = []
    for Mloc in (B, C):
        rows = np.repeat(model.ed, 12, axis=1).reshape(-1)
        cols = np.tile(model.ed, (1, 12)).reshape(-1)
        A = sp.coo_matrix((Mloc.reshape(-1), (rows, cols)), shape=(n, n)).tocsr()
        A.sum_duplicates()
        di = A.diagonal().copy()
        L = sp.tril(A, k=-1).tocsr()
        L.sort_indices()
        ggl = np.zeros(len(jg))
        # place L's entries into the given pattern
        Lc = L.tocoo()
        key_pat = (np.repeat(np.arange(n), np.diff(ig)) * n + jg)
        key_L = Lc.row.astype(np.int64) * n + Lc.col
        pos = np.searchsorted(key_pat, key_L)
        assert np.all(key_pat[pos] == key_L)
        ggl[pos] = Lc.data
        out += [ggl, di]
    return tuple(out)


def boundary_edges(model) -> np.ndarray:
    """Edges whose two nodes are both in ``l13d`` (``T_Bound_Cond.cpp:58-89``)."""
    isb = np.zeros(model.kuzlov, bool)
    isb[model.l13d] = True
    return isb[model.edges[:, 0]] & isb[model.edges[:, 1]]


def system_matrix(ig, jg, b_ggl, b_di, c_ggl, c_di, mult, bnd):
    """Decade matrix: mult*C + B, diagonal bump, Dirichlet (``A2/time_approx.cpp:987-996``)."""
    di = c_di * mult + b_di
    ggl = c_ggl * mult + b_ggl
    di = di + di * 1e-12
    di = np.where(bnd, 1.0, di)
    n = len(ig) - 1
    rows = np.repeat(np.arange(n), np.diff(ig))
    ggl = np.where(bnd[rows] | bnd[jg], 0.0, ggl)
    return ggl, di


def rsf_to_csc(ig, jg, ggl, di):
    """Full symmetric scipy matrix from lower RSF."""
    n = len(ig) - 1
    rows = np.repeat(np.arange(n), np.diff(ig))
    L = sp.coo_matrix((ggl, (rows, jg)), shape=(n, n))
    D = sp.diags(di)
    return (L + L.T + D).tocsc()


def sym_matvec(ig, jg, ggl, di, x):
    """``mult_symmetr`` (``A2/TaskSP.cpp:59-73``) on one or several columns."""
    n = len(ig) - 1
    rows = np.repeat(np.arange(n), np.diff(ig))
    L = sp.csr_matrix((ggl, (rows, jg)), shape=(n, n))
    return di[:, None] * x + L @ x + L.T @ x if x.ndim == 2 else di * x + L @ x + L.T @ x
