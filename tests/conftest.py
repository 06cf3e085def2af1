"""Shared fixtures.  `-m "not gpu"` runs on any machine (oracle, host logic,
C-ABI load); `-m gpu` tests need one MI355X and the built libmtsp2.so."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "modem-gpu_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libmtsp2.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    import mtsp2
    n = mtsp2.device_count()
    if n < 1:
        raise RuntimeError("gpu test selected but no HIP device is visible")
    ctx = mtsp2.Context(0)
    yield ctx
    ctx.close()


def small_problem(seed=0, nx=10, ny=9, nzAir=4, nzEarth=7, rho=(1.0, 1000.0)):
    """A small non-uniform grid with a random LOGE model (oracle-sized)."""
    rng = np.random.default_rng(seed)
    dx = rng.uniform(100, 300, nx)
    dy = rng.uniform(100, 300, ny)
    dz = np.concatenate([np.geomspace(30000, 100, nzAir), rng.uniform(50, 200, nzEarth)])
    lc = -np.log(rng.uniform(rho[0], rho[1], (nx, ny, nzEarth)))
    return dx, dy, dz, nzAir, lc


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ------------------------------------------------------------------ DILU
# A numpy restatement of block DILU in an arbitrary ordering (test
# infrastructure): pivots p_e = b_ee - sum_{l in N(e), key[l] < key[e]}
# b_el b_le / p_l over the x-x, y-y, z-z blocks B of A(omega), and
# M = (P + E) P^-1 (P + F) with E / F the couplings to lower / upper keys.
def block_diag_operator(o, w):
    """B = the same-component blocks of A(w) = AAii - i*w*mu0*sigma_E*V_E
    (ISIGN = -1), interior EDGEi indexing, as scipy CSR."""
    import scipy.sparse as sp
    o.update_freq(w)
    a = o.csr("AAii")
    n = o.nInt
    A = sp.csr_matrix((a["val"], a["col"], a["row"]), shape=(n, n)).astype(complex)
    A = (A - sp.diags(1j * o.vms())).tocoo()
    comp = np.zeros(n, int)
    comp[o.nxInt:o.nxInt + o.nyInt] = 1
    comp[o.nxInt + o.nyInt:] = 2
    same = comp[A.row] == comp[A.col]
    return sp.csr_matrix((A.data[same], (A.row[same], A.col[same])), shape=(n, n))


def dilu_factor(B, key):
    import scipy.sparse as sp
    C = B.tocoo()
    val = dict(zip(zip(C.row.tolist(), C.col.tolist()), C.data.tolist()))
    rows = [[] for _ in range(B.shape[0])]
    for r, c, v in zip(C.row.tolist(), C.col.tolist(), C.data.tolist()):
        if r != c and key[c] < key[r]:
            rows[r].append((c, v))
    piv = np.zeros(B.shape[0], complex)
    diag = B.diagonal()
    for e in np.argsort(key, kind="stable").tolist():
        p = diag[e]
        for l, bel in rows[e]:
            p -= bel * val[(l, e)] / piv[l]
        piv[e] = p
    lo = key[C.row] > key[C.col]
    up = key[C.row] < key[C.col]
    n = B.shape[0]
    lower = sp.csr_matrix((C.data[lo], (C.row[lo], C.col[lo])), shape=(n, n))
    upper = sp.csr_matrix((C.data[up], (C.row[up], C.col[up])), shape=(n, n))
    return piv, lower, upper


def dilu_mult(piv, lower, upper, z):
    """M z = (P + E) P^-1 (P + F) z."""
    u = piv * z + upper @ z
    return u + lower @ (u / piv)


def rb_tile_keys(nx, ny, nz, t=8):
    """Order key of every interior edge (EDGEi order) in the red-black tile
    ordering: t^3 tiles of each component's interior lattice, coloured by the
    parity of the tile index sum, all red tiles first; lexicographic order
    inside a tile.  Tiles of one colour never couple, so their relative order
    is immaterial."""
    lat = [(nx, ny - 1, nz - 1), (nx - 1, ny, nz - 1), (nx - 1, ny - 1, nz)]
    keys = []
    for c, (A, Bd, Cd) in enumerate(lat):
        k, j, i = np.meshgrid(np.arange(Cd), np.arange(Bd), np.arange(A), indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()   # Fortran order, i fastest
        ti, tj, tk = i // t, j // t, k // t
        colour = (ti + tj + tk) & 1
        tile = ti + 4096 * (tj + 4096 * tk)
        lex = (i % t) + t * ((j % t) + t * (k % t))
        keys.append(((colour * 3 + c) * 4096 ** 3 + tile) * t ** 3 + lex)
    return np.concatenate(keys)
