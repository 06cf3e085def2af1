"""Mathematical pins of the oracle that do not rely on the reference build:
topology identities, symmetry, the DILU invariant, the closed-form uniform
grid operator, dense-solve agreement of the Krylov solvers, the adjoint
identity of the gradient map, and the 1-D halfspace impedance."""
import numpy as np
import pytest
import scipy.sparse as sp

import sp2oracle
from conftest import rel, small_problem


def S(csr, shape=None):
    n = csr["nrow"] if shape is None else shape[0]
    m = csr["ncol"] if shape is None else shape[1]
    return sp.csr_matrix((csr["val"], csr["col"], csr["row"]), shape=(n, m))


@pytest.fixture(scope="module")
def prob():
    dx, dy, dz, nza, lc = small_problem(21)
    return sp2oracle.Problem(dx, dy, dz, nza, lc)


def test_curl_grad_is_zero(prob):
    T, G = S(prob.csr("T")), S(prob.csr("G"))
    assert (T @ G).count_nonzero() == 0
    assert set(np.unique(T.data)) <= {-1.0, 1.0} and np.all(np.diff(T.indptr) == 4)
    assert np.all(np.diff(G.indptr) == 2)


def test_curlcurl_symmetric_and_annihilates_gradients(prob):
    CCii, CCib = S(prob.csr("CCii")), S(prob.csr("CCib"))
    d = (CCii - CCii.T)
    assert abs(d).max() <= 1e-12 * abs(CCii).max()
    # CC (G phi / EdgeL) = 0 for any node potential (T G = 0)
    m = prob.metric()
    G = S(prob.csr("G"))
    phi = np.random.default_rng(0).standard_normal(prob.nNode)
    e = (G @ phi) / m["EdgeL"]
    ei, eb, _, _ = prob.index()
    y = CCii @ e[ei] + CCib @ e[eb]
    assert np.linalg.norm(y) <= 1e-11 * abs(CCii).max() * np.linalg.norm(e)


def test_uniform_grid_operator_is_vector_laplacian():
    """Uniform spacing and conductivity, no air: CC + GD = -Laplacian per
    component, so cross-component couplings cancel and each row keeps the
    7-point pattern 6h, -h (h = cell size)."""
    h = 100.0
    n = 6
    p = sp2oracle.Problem(np.full(n, h), np.full(n, h), np.full(n, h), 0,
                          np.full((n, n, n), np.log(0.01)))
    A = S(p.csr("AAii")).tocoo()
    ei = p.index()[0]
    comp = np.searchsorted([p.nxInt, p.nxInt + p.nyInt], np.arange(p.nInt), side="right")
    cross = comp[A.row] != comp[A.col]
    big = abs(A.data).max()
    assert abs(A.data[cross]).max() <= 1e-12 * big
    same = ~cross & (abs(A.data) > 1e-12 * big)
    diag = A.row == A.col
    # rows whose 6 lattice neighbours are all interior: the full 7-point stencil
    full = np.bincount(A.row[same & ~diag], minlength=p.nInt) == 6
    sel = full[A.row] & same
    assert full.sum() > 0
    assert np.allclose(A.data[sel & diag], 6 * h, rtol=1e-12)
    assert np.allclose(A.data[sel & ~diag], -h, rtol=1e-12)


def test_dilu_reproduces_block_pattern(prob):
    """(L U)_ij = A_blk_ij on A_blk's pattern (no fill inside the 7-point
    lattice blocks), the defining property of Dilu_Cmplx_AS."""
    w = 2 * np.pi * 0.3
    prob.update_freq(w)
    L, U = S(prob.lu(0)), S(prob.lu(1))
    A = S(prob.csr("AAii")).astype(complex).tolil()
    v = prob.vms()
    for i in range(prob.nInt):
        A[i, i] = A[i, i] - 1j * v[i]
    A = A.tocsr()
    blk = np.searchsorted([prob.nxInt, prob.nxInt + prob.nyInt], np.arange(prob.nInt), side="right")
    Ac = A.tocoo()
    keep = blk[Ac.row] == blk[Ac.col]
    LU = (L @ U).tocsr()
    got = np.asarray(LU[Ac.row[keep], Ac.col[keep]]).ravel()
    assert np.max(abs(got - Ac.data[keep])) <= 1e-12 * abs(Ac.data).max()


def test_dilu_restatement_matches_oracle_in_natural_order(prob):
    """The numpy DILU-in-any-ordering restatement (tests/conftest.py, used to
    check the red-black tile DILU on the GPU) reproduces, with the natural
    (EDGEi) ordering, the oracle's U pivots and its PC_Lsolve/PC_Usolve: the
    oracle's M^-1 x is mapped back to x by the restated M."""
    from conftest import block_diag_operator, dilu_factor, dilu_mult
    w = 2 * np.pi * 0.7
    B = block_diag_operator(prob, w)
    piv, lower, upper = dilu_factor(B, np.arange(prob.nInt))
    U = S(prob.lu(1))
    assert rel(piv, U.diagonal()) <= 1e-13
    rng = np.random.default_rng(9)
    x = rng.standard_normal(prob.nInt) + 1j * rng.standard_normal(prob.nInt)
    z = prob.pc_usolve(prob.pc_lsolve(x))
    assert rel(dilu_mult(piv, lower, upper, z), x) <= 1e-12
    # a plausible slip -- a lower coupling dropped -- is caught
    bad = lower.tolil()
    r0 = bad.rows[prob.nInt // 2][0]
    bad[prob.nInt // 2, r0] = 0.0
    assert rel(dilu_mult(piv, bad.tocsr(), upper, z), x) > 1e-9


def test_bicg_matches_dense_solve():
    dx, dy, dz, nza, lc = small_problem(22, nx=6, ny=5, nzAir=3, nzEarth=4)
    p = sp2oracle.Problem(dx, dy, dz, nza, lc)
    for f in (1e-3, 10.0):
        w = 2 * np.pi * f
        p.update_freq(w)
        A = S(p.csr("AAii")).toarray().astype(complex) - 1j * np.diag(p.vms())
        rng = np.random.default_rng(1)
        b = rng.standard_normal(p.nInt) + 1j * rng.standard_normal(p.nInt)
        xd = np.linalg.solve(A, b)
        for gpu in (False, True):
            for pc in (0, 1, 2):
                x, info = p.bicg(b, None, maxit=3000, tol=1e-12, gpu_semantics=gpu, precond=pc)
                assert info["status"] == 0
                assert rel(x, xd) <= 1e-8 * np.linalg.cond(A) ** 0.5


def test_pcg_matches_dense_solve(prob):
    prob.update_freq(2 * np.pi)
    prob.divcor_setup()
    V = S(prob.csr("VDsG")).toarray()
    rng = np.random.default_rng(2)
    b = rng.standard_normal(prob.nNi) + 0j
    phi, info = prob.pcg(b, maxit=500, tol=1e-12)
    assert rel(phi, np.linalg.solve(V, b)) <= 1e-9


def test_gradient_map_adjoint_identity(prob):
    """<dsigma_E(dm), e> == <dm, dEdgeToModelParam(e)> (the P / P^T pair of
    SymmetryTest.f90), with the linearised sigma map taken by central
    differences of ModelParamToEdge."""
    rng = np.random.default_rng(3)
    dm = rng.standard_normal(prob.logcond.shape)
    e = rng.standard_normal(prob.nE)
    eps = 1e-5
    dsig = (prob.model2edge(prob.logcond + eps * dm) - prob.model2edge(prob.logcond - eps * dm)) / (2 * eps)
    lhs = np.dot(dsig, e)
    rhs = np.sum(dm * prob.edge2model(e))
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs)


def test_halfspace_impedance():
    """Uniform 100 ohm-m halfspace: apparent resistivity 100 ohm-m and phase
    45 degrees at the surface (Z = i w mu E / dE/dz for e^{-i w t})."""
    from mtsp2 import problem as P
    grid = P.make_grid(20, 20, 40, nair=8, dcore=400.0, npad=5)
    nze = len(grid["dz"]) - grid["nzAir"]
    lc = np.full((20, 20, nze), np.log(0.01))
    p = sp2oracle.Problem(grid["dx"], grid["dy"], grid["dz"], grid["nzAir"], lc)
    f = 1.0
    w = 2 * np.pi * f
    p.update_freq(w)
    e = P.plane_wave(grid, w, 0)
    b = p.rhs_fwd(e)
    x0 = e[p.index()[0]]
    x, info = p.bicg(b, x0, maxit=2000, tol=1e-10)
    assert info["status"] == 0
    full = e.copy()
    full[p.index()[0]] = x
    nx, ny, nz = 20, 20, len(grid["dz"])
    Ex = full[:nx * (ny + 1) * (nz + 1)].reshape(nz + 1, ny + 1, nx)
    k0 = grid["nzAir"]
    E0 = Ex[k0, ny // 2, nx // 2]
    E1 = Ex[k0 + 1, ny // 2, nx // 2]
    h = grid["dz"][k0]
    Z = 1j * w * P.MU0 * 0.5 * (E0 + E1) / ((E1 - E0) / h)
    rho_a = abs(Z) ** 2 / (w * P.MU0)
    phase = np.degrees(np.angle(Z))
    assert abs(rho_a - 100.0) / 100.0 <= 0.02
    assert abs(abs(phase) - 45.0) <= 1.0
