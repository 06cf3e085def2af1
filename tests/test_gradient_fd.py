"""Gradient assembly (mt_host_gradient / mt_gradient) against central finite
differences of a field misfit on a tiny grid, with dense solves as the
forward/adjoint oracle (numpy, exact to rounding):

  phi(m) = 1/2 sum_pol |d_pol - P e_pol(m)|^2,  A(m) e = b (plane-wave BC)
  lam   = A^-T P^T conj(d - P e)   ->  d phi/dm = Re sum_pol lam^T (dA/dm) e

The exact mode (the transpose adjoint plus the grad-div derivative) must match
finite differences; the reference mode (PmultT's diagonal term only) is
checked against the oracle's PmultT, which is bit-exact with the reference's
EdgeToModelParam."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtsp2
import sp2oracle
from mtsp2 import problem as P

GRID = P.make_grid(10, 9, 6, nair=3, npad=2)
NX, NY, NZ = len(GRID["dx"]), len(GRID["dy"]), len(GRID["dz"])
NZE = NZ - GRID["nzAir"]
W = 2 * np.pi * 1.0


def _system(lc):
    p = sp2oracle.Problem(GRID["dx"], GRID["dy"], GRID["dz"], GRID["nzAir"], lc)
    p.update_freq(W)
    a = p.csr("AAii")
    A = sp.csr_matrix((a["val"], a["col"], a["row"]), shape=(p.nInt, p.nInt)).toarray().astype(complex)
    A -= 1j * np.diag(p.vms())
    return p, A


def _fields(lc):
    p, A = _system(lc)
    E = [np.linalg.solve(A, p.rhs_fwd(P.plane_wave(GRID, W, pol))) for pol in (0, 1)]
    return p, A, E


@pytest.fixture(scope="module")
def setup():
    rng = np.random.default_rng(5)
    lc = np.log(0.01) + 0.5 * rng.standard_normal((NX, NY, NZE))
    p, A, E = _fields(lc)
    sites = rng.choice(p.nxInt + p.nyInt, 30, replace=False)
    _, _, Et = _fields(lc + 0.3 * rng.standard_normal(lc.shape))
    d = [e[sites] for e in Et]
    return lc, p, A, E, sites, d


def _phi(lc, sites, d):
    _, _, E = _fields(lc)
    return 0.5 * sum(np.sum(np.abs(d[k] - E[k][sites]) ** 2) for k in range(2))


def _adjoint(p, A, E, sites, d):
    ei = p.index()[0]
    e_all, lam = [], []
    for k in range(2):
        ea = P.plane_wave(GRID, W, k)
        ea[ei] = E[k]
        e_all.append(ea)
        s = np.zeros(p.nInt, complex)
        s[sites] = np.conj(d[k] - E[k][sites])
        lam.append(np.linalg.solve(A.T, s))
    return np.array(e_all), np.array(lam)


def test_exact_gradient_matches_central_differences(setup):
    lc, p, A, E, sites, d = setup
    e_all, lam = _adjoint(p, A, E, sites, d)
    g = mtsp2.host_gradient(GRID["dx"], GRID["dy"], GRID["dz"], GRID["nzAir"], lc, W, e_all, lam,
                            mode=mtsp2.GRAD_EXACT)
    rng = np.random.default_rng(8)
    cells = [tuple(rng.integers(0, s) for s in lc.shape) for _ in range(8)]
    h = 1e-4
    for c in cells:
        lp, lm = lc.copy(), lc.copy()
        lp[c] += h
        lm[c] -= h
        fd = (_phi(lp, sites, d) - _phi(lm, sites, d)) / (2 * h)
        assert abs(fd - g[c]) <= 1e-5 * abs(fd) + 1e-12 * np.abs(g).max(), (c, fd, g[c])


def test_reference_mode_is_pmultT(setup):
    """Mode 0 equals -PmultT(e, V_E lam) of the oracle (the reference's PmultT
    + dEdgeToModelParam, bit-exact), i.e. the diagonal term alone."""
    lc, p, A, E, sites, d = setup
    e_all, lam = _adjoint(p, A, E, sites, d)
    g0 = mtsp2.host_gradient(GRID["dx"], GRID["dy"], GRID["dz"], GRID["nzAir"], lc, W, e_all, lam,
                             mode=mtsp2.GRAD_REFERENCE)
    ei = p.index()[0]
    ve = p.metric()["Vedge"]
    lam_all = np.zeros((2, p.nE), complex)
    lam_all[:, ei] = lam * ve[ei]
    ref = -p.pmultT(W, e_all, lam_all)
    assert np.max(np.abs(g0 - ref)) <= 1e-12 * np.max(np.abs(ref))
