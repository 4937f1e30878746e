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


def _system(d, mult_scale=1.0, sigma_floor=None):
    m = refio.read_model(d)
    if sigma_floor is not None:
        m.sigma[:] = np.maximum(m.sigma, sigma_floor)
    n = m.edges.shape[0]
    ig, jg = fem.portrait(m.ed, n)
    b_ggl, b_di, c_ggl, c_di = fem.assemble(m, ig, jg)
    bnd = fem.boundary_edges(m)
    mult = mult_scale / (m.times[1] - m.times[0])
    ggl, di = fem.system_matrix(ig, jg, b_ggl, b_di, c_ggl, c_di, mult, bnd)
    return ig, jg, ggl, di


def _solve_both(emtd, d, sigma_floor, nrhs=2):
    ig, jg, ggl, di = _system(d, sigma_floor=sigma_floor)
    K = fem.rsf_to_csc(ig, jg, ggl, di)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((nrhs, len(di)))
    lu = spla.splu(K)
    xr = np.stack([lu.solve(bb) for bb in b])
    s = emtd.Solver(precision=emtd.FP64)
    s.factorize(ig, jg, ggl, di)
    x = s.solve(b)
    st = s.stats()
    print("fronts", st.n_fronts, "levels", st.n_levels, "nnzL", st.nnz_l)
    res = max(np.linalg.norm(K @ x[k] - b[k]) / np.linalg.norm(b[k]) for k in range(nrhs))
    res_ref = max(np.linalg.norm(K @ xr[k] - b[k]) / np.linalg.norm(b[k]) for k in range(nrhs))
    err = np.abs(x - xr).max() / np.abs(xr).max()
    print("residual", res, "SuperLU residual", res_ref, "rel diff vs SuperLU", err)
    return res, res_ref, err


@pytest.mark.parametrize("cells", [(6, 6, 4), (10, 10, 8), (14, 12, 10)])
def test_solver_fp64_well_conditioned(emtd, tmp_path, cells, sweeps):
    """Conductivity floored at 1e-2 S/m: the system is well conditioned, tight bar."""
    d, _ = make_model(tmp_path, cfg=1, cells=cells)
    res, res_ref, err = _solve_both(emtd, d, 1e-2)
    assert res < 1e-13
    assert err < 1e-11


@pytest.mark.parametrize("cells", [(6, 6, 4), (14, 12, 10)])
def test_solver_fp64_with_air(emtd, tmp_path, cells, sweeps):
    """Real air (1e-8 S/m): kappa ~ 1e8-1e10; the residual must match a backward-stable
    direct solve (SuperLU's) and the solution SuperLU's to its conditioning."""
    d, _ = make_model(tmp_path, cfg=1, cells=cells)
    res, res_ref, err = _solve_both(emtd, d, None)
    assert res < 20 * max(res_ref, 1e-15) + 1e-14
    assert err < 1e-6


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
