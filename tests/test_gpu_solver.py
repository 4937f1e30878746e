"""GPU parity of the assembly kernel and of the pardiso-seam Cholesky against the oracle."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from conftest import make_model
from oracle import fem, refio

pytestmark = pytest.mark.gpu


def _mesh(emtd, m):
    return emtd.Mesh(m.xyz, m.nver, m.ed, m.edges, m.nvkat, m.sigma, m.mu)


@pytest.mark.parametrize("cfg,cells", [(1, (8, 8, 6)), (3, (10, 9, 8))])
def test_assembly_matches_oracle(emtd, tmp_path, cfg, cells):
    d, _ = make_model(tmp_path, cfg=cfg, cells=cells)
    m = refio.read_model(d)
    n = m.edges.shape[0]
    ig, jg = fem.portrait(m.ed, n)
    ref = fem.assemble(m, ig, jg)
    mesh = _mesh(emtd, m)
    got = emtd.assemble(mesh, ig.astype(np.int32), jg.astype(np.int32))
    for r, g, name in zip(ref, got, ("b_ggl", "b_di", "c_ggl", "c_di")):
        scale = np.abs(r).max()
        err = np.abs(r - g).max() / scale
        print(name, "max rel err", err)
        assert err <= 1e-13, name


def _system(d, mult_scale=1.0):
    m = refio.read_model(d)
    n = m.edges.shape[0]
    ig, jg = fem.portrait(m.ed, n)
    b_ggl, b_di, c_ggl, c_di = fem.assemble(m, ig, jg)
    bnd = fem.boundary_edges(m)
    mult = mult_scale / (m.times[1] - m.times[0])
    ggl, di = fem.system_matrix(ig, jg, b_ggl, b_di, c_ggl, c_di, mult, bnd)
    return ig, jg, ggl, di


@pytest.mark.parametrize("cells", [(6, 6, 4), (10, 10, 8), (14, 12, 10)])
def test_solver_fp64_matches_superlu(emtd, tmp_path, cells):
    d, _ = make_model(tmp_path, cfg=1, cells=cells)
    ig, jg, ggl, di = _system(d)
    K = fem.rsf_to_csc(ig, jg, ggl, di)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((2, len(di)))
    xr = np.stack([spla.splu(K).solve(bb) for bb in b])
    s = emtd.Solver(precision=emtd.FP64)
    s.factorize(ig, jg, ggl, di)
    x = s.solve(b)
    st = s.stats()
    print("fronts", st.n_fronts, "levels", st.n_levels, "nnzL", st.nnz_l)
    res = np.linalg.norm(K @ x[0] - b[0]) / np.linalg.norm(b[0])
    err = np.abs(x - xr).max() / np.abs(xr).max()
    print("residual", res, "rel err vs SuperLU", err)
    assert res < 1e-12
    assert err < 1e-9


def test_solver_dense_small(emtd):
    """A random SPD matrix with a generic sparsity: graph dissection path vs dense numpy."""
    rng = np.random.default_rng(7)
    n = 300
    A = np.zeros((n, n))
    for i in range(n):
        for j in rng.choice(n, 4, replace=False):
            if j != i:
                v = rng.standard_normal()
                A[i, j] = A[j, i] = v
    A += np.diag(np.abs(A).sum(1) + 1.0)
    ig = [0]
    jg, ggl = [], []
    for i in range(n):
        cols = np.nonzero(A[i, :i])[0]
        jg += cols.tolist()
        ggl += A[i, cols].tolist()
        ig.append(len(jg))
    b = rng.standard_normal(n)
    s = emtd.Solver()
    s.factorize(np.array(ig), np.array(jg), np.array(ggl), np.diag(A).copy())
    x = s.solve(b)
    xr = np.linalg.solve(A, b)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-12
