"""Pins the oracle's edge-element core (CPU only).

Known answers come from the reference's own unrolled local matrices (T_Brick.cpp), from
brute-force quadrature of the basis, from algebraic invariants of the Nedelec space, and from
the tensor-grid counts of the three named meshes.
"""
import itertools

import numpy as np
import pytest

from conftest import make_model
from oracle import fem, refio


def ref_local_b(hx, hy, hz):
    """Curl-curl matrix exactly as A2/T_Brick.cpp:142-310 writes it (known answer)."""
    mx, my, mz = 2 * hx / (hy * hz), 2 * hy / (hx * hz), 2 * hz / (hx * hy)
    t1 = my + mz; t2 = my / 3; t3 = 2 / 3 * mz; t4 = t2 - t3; t5 = 2 / 3 * my; t6 = mz / 3
    t7 = -t5 + t6; t8 = mx + mz; t9 = 2 / 3 * mx; t10 = -t9 + t6; t11 = mx / 3; t12 = t11 - t3
    t13 = mx + my; t14 = t11 - t5; t15 = -t9 + t2
    b = np.array([
        [2 / 3 * t1, t4, t7, -t1 / 3, -t3, -t6, t3, t6, -t5, t5, -t2, t2],
        [t4, 2 / 3 * t1, -t1 / 3, t7, t3, t6, -t3, -t6, -t2, t2, -t5, t5],
        [t7, -t1 / 3, 2 / 3 * t1, t4, -t6, -t3, t6, t3, t5, -t5, t2, -t2],
        [-t1 / 3, t7, t4, 2 / 3 * t1, t6, t3, -t6, -t3, t2, -t2, t5, -t5],
        [-t3, t3, -t6, t6, 2 / 3 * t8, t10, t12, -t8 / 3, -t9, -t11, t9, t11],
        [-t6, t6, -t3, t3, t10, 2 / 3 * t8, -t8 / 3, t12, t9, t11, -t9, -t11],
        [t3, -t3, t6, -t6, t12, -t8 / 3, 2 / 3 * t8, t10, -t11, -t9, t11, t9],
        [t6, -t6, t3, -t3, -t8 / 3, t12, t10, 2 / 3 * t8, t11, t9, -t11, -t9],
        [-t5, -t2, t5, t2, -t9, t9, -t11, t11, 2 / 3 * t13, t14, t15, -t13 / 3],
        [t5, t2, -t5, -t2, -t11, t11, -t9, t9, t14, 2 / 3 * t13, -t13 / 3, t15],
        [-t2, -t5, t2, t5, t9, -t9, t11, -t11, t15, -t13 / 3, 2 / 3 * t13, t14],
        [t2, t5, -t2, -t5, t11, -t11, t9, -t9, -t13 / 3, t15, t14, 2 / 3 * t13],
    ])
    return b


def ref_local_c(hx, hy, hz):
    """Mass matrix exactly as A2/T_Brick.cpp:311-474 writes it (known answer)."""
    c = np.zeros((12, 12))
    for d, mult in enumerate((hy * hz / (2 * hx), hx * hz / (2 * hy), hx * hy / (2 * hz))):
        t1, t2, t3 = 8 / 9 * mult, 4 / 9 * mult, 2 / 9 * mult
        blk = np.array([[t1, t2, t2, t3], [t2, t1, t3, t2], [t2, t3, t1, t2], [t3, t2, t2, t1]])
        c[4 * d:4 * d + 4, 4 * d:4 * d + 4] = blk
    return c


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (2.0, 3.0, 5.0), (10.0, 0.7, 123.0)])
def test_local_matrices_match_reference_closed_forms(h):
    B, C = fem.local_matrices(np.array([h]))
    np.testing.assert_allclose(B[0], ref_local_b(*h), rtol=1e-13, atol=1e-13 * np.abs(B[0]).max())
    np.testing.assert_allclose(C[0], ref_local_c(*h), rtol=1e-13, atol=1e-15)


def test_cube_known_entries():
    h = 2.0
    B, C = fem.local_matrices(np.array([[h, h, h]]))
    assert np.isclose(C[0, 0, 0], 4 * h / 9) and np.isclose(C[0, 0, 1], 2 * h / 9)
    assert np.isclose(C[0, 0, 3], h / 9) and C[0, 0, 4] == 0
    assert np.isclose(B[0, 0, 0], 8 / (3 * h)) and np.isclose(B[0, 0, 1], -2 / (3 * h))


def _basis(a, s, h):
    """Value and curl of local basis a at normalized point s (independent re-derivation)."""
    d = a // 4
    k = a % 4
    p, q = (d + 1) % 3, (d + 2) % 3
    bp, bq = k & 1, k >> 1
    l = lambda b, x: x if b else 1 - x  # noqa: E731
    dl = lambda b: 1.0 if b else -1.0  # noqa: E731
    val = np.zeros(3)
    val[d] = 2 / h[d] * l(bp, s[p]) * l(bq, s[q])
    grad = np.zeros(3)  # gradient of the scalar factor
    grad[p] = 2 / h[d] * dl(bp) / h[p] * l(bq, s[q])
    grad[q] = 2 / h[d] * l(bp, s[p]) * dl(bq) / h[q]
    e = np.zeros(3)
    e[d] = 1.0
    return val, np.cross(grad, e)


def test_local_matrices_by_gauss_quadrature():
    h = np.array([1.5, 0.8, 2.2])
    g, w = np.polynomial.legendre.leggauss(3)
    g = 0.5 * (g + 1)
    w = 0.5 * w
    Bq = np.zeros((12, 12))
    Cq = np.zeros((12, 12))
    for (i, a), (j, b), (k, c) in itertools.product(enumerate(g), repeat=3):
        s = (a, b, c)
        wt = w[i] * w[j] * w[k] * h.prod()
        vals = [_basis(e, s, h) for e in range(12)]
        for x in range(12):
            for y in range(12):
                Cq[x, y] += wt * vals[x][0] @ vals[y][0]
                Bq[x, y] += wt * vals[x][1] @ vals[y][1]
    B, C = fem.local_matrices(h[None])
    np.testing.assert_allclose(B[0], Bq, atol=1e-12)
    np.testing.assert_allclose(C[0], Cq, atol=1e-12)


def test_local_invariants():
    h = np.array([[1.3, 0.4, 2.0]])
    B, C = fem.local_matrices(h)
    B, C = B[0], C[0]
    assert np.allclose(B, B.T) and np.allclose(C, C.T)
    assert np.linalg.eigvalsh(C).min() > 0
    ev = np.linalg.eigvalsh(B)
    assert ev.min() > -1e-12 * ev.max()
    assert np.sum(ev > 1e-10 * ev.max()) == 5          # rank 5 = 12 edges - 7 gradients
    # discrete gradient: u_e = (phi(end) - phi(start)) / 2 with the edge DOF = tangential*len/2
    verts = [(i & 1, (i >> 1) & 1, (i >> 2) & 1) for i in range(8)]
    loc_edges = []
    for a in range(12):
        d = a // 4
        k = a % 4
        p, q = (d + 1) % 3, (d + 2) % 3
        v0 = [0, 0, 0]
        v0[p] = k & 1
        v0[q] = k >> 1
        v1 = list(v0)
        v1[d] = 1
        loc_edges.append((verts.index(tuple(v0)), verts.index(tuple(v1))))
    G = np.zeros((12, 8))
    for a, (s, e) in enumerate(loc_edges):
        G[a, e] += 0.5
        G[a, s] -= 0.5
    assert np.abs(B @ G).max() < 1e-12 * np.abs(B).max()


def test_edge_numbering_matches_std_set_order(mg):
    nx, ny, nz = 4, 3, 3
    edges, ed, nver = mg.grid_topology(nx, ny, nz)
    pairs = set()
    for e in range((nx - 1) * (ny - 1) * (nz - 1)):
        v = nver[e] + 1
        for a, b in ((0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (4, 6), (1, 3), (5, 7),
                     (0, 4), (1, 5), (2, 6), (3, 7)):
            pairs.add((int(v[a]), int(v[b])))
    order = sorted(pairs)
    assert [tuple(p) for p in (edges + 1).tolist()] == [tuple(p) for p in order]
    num = {p: i for i, p in enumerate(order)}
    loc = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (4, 6), (1, 3), (5, 7), (0, 4), (1, 5), (2, 6), (3, 7)]
    for e in range(len(ed)):
        v = nver[e] + 1
        assert [num[(int(v[a]), int(v[b]))] for a, b in loc] == ed[e].tolist()


def _dense_assembly(m):
    n = m.edges.shape[0]
    B, C = fem.local_matrices(fem.element_sizes(m))
    Bd = np.zeros((n, n))
    Cd = np.zeros((n, n))
    for e in range(m.kpar):
        idx = m.ed[e]
        Bd[np.ix_(idx, idx)] += B[e] / m.mu[m.nvkat[e]]
        Cd[np.ix_(idx, idx)] += C[e] * m.sigma[m.nvkat[e]]
    return Bd, Cd


def test_portrait_and_assembly_brute_force(tmp_path):
    d, _ = make_model(tmp_path, cfg=1, cells=(2, 3, 2))
    m = refio.read_model(d)
    n = m.edges.shape[0]
    ig, jg = fem.portrait(m.ed, n)
    Bd, Cd = _dense_assembly(m)
    pat = np.zeros((n, n), bool)
    for e in range(m.kpar):
        pat[np.ix_(m.ed[e], m.ed[e])] = True
    for i in range(n):
        assert jg[ig[i]:ig[i + 1]].tolist() == [j for j in range(i) if pat[i, j]]
    b_ggl, b_di, c_ggl, c_di = fem.assemble(m, ig, jg)
    rows = np.repeat(np.arange(n), np.diff(ig))
    np.testing.assert_allclose(b_ggl, Bd[rows, jg], atol=1e-13 * np.abs(Bd).max())
    np.testing.assert_allclose(c_ggl, Cd[rows, jg], atol=1e-13 * np.abs(Cd).max())
    np.testing.assert_allclose(b_di, np.diag(Bd), rtol=1e-13)
    np.testing.assert_allclose(c_di, np.diag(Cd), rtol=1e-13)


def test_config1_counts_known_answer(tmp_path, mg):
    """n, strictly-lower nnz and Dirichlet edges of the 40x40x30 mesh (152,110 / 2,328,000 / 16,000)."""
    spec = mg.config_spec(1)
    nxn, nyn, nzn = len(spec.xs), len(spec.ys), len(spec.zs)
    edges, ed, _ = mg.grid_topology(nxn, nyn, nzn)
    n = edges.shape[0]
    assert n == 152110
    ig, jg = fem.portrait(ed, n)
    assert len(jg) == 2328000
    bn = np.zeros(nxn * nyn * nzn, bool)
    bn[mg.boundary_nodes(nxn, nyn, nzn)] = True
    assert int((bn[edges[:, 0]] & bn[edges[:, 1]]).sum()) == 16000
    full = np.diff(ig) + np.bincount(jg, minlength=n) + 1
    assert full.max() == 33


def test_config3_and_2_sizes(mg):
    for cfg, n_expect in ((3, 989010), (2, 1844260)):
        spec = mg.config_spec(cfg)
        edges, _, _ = mg.grid_topology(len(spec.xs), len(spec.ys), len(spec.zs))
        assert edges.shape[0] == n_expect


def test_dirichlet_and_bump(tmp_path):
    d, _ = make_model(tmp_path, cfg=1, cells=(3, 3, 2))
    m = refio.read_model(d)
    n = m.edges.shape[0]
    ig, jg = fem.portrait(m.ed, n)
    b_ggl, b_di, c_ggl, c_di = fem.assemble(m, ig, jg)
    bnd = fem.boundary_edges(m)
    ggl, di = fem.system_matrix(ig, jg, b_ggl, b_di, c_ggl, c_di, 7.0, bnd)
    K = fem.rsf_to_csc(ig, jg, ggl, di).toarray()
    raw = fem.rsf_to_csc(ig, jg, c_ggl * 7.0 + b_ggl, c_di * 7.0 + b_di).toarray()
    for i in range(n):
        if bnd[i]:
            assert K[i, i] == 1.0 and np.count_nonzero(K[i]) == 1 and np.count_nonzero(K[:, i]) == 1
        else:
            assert K[i, i] == raw[i, i] + raw[i, i] * 1e-12
            for j in range(n):
                if j != i and not bnd[j]:
                    assert K[i, j] == raw[i, j]
    assert np.linalg.eigvalsh(K).min() > 0
