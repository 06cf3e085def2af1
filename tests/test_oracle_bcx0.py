"""The oracle's BC_x0_WS restatement (boundary_ws.f90:21-171 with the 2-D TE
code of 3D_MT/FWD/Mod2d/): bit for bit against the reference's own build
(committed fixture and, where the reference is mounted, live), and against
what the mathematics fixes: a laterally uniform model has the 1-D solution in
every column, a uniform halfspace decays as exp(-(1-i) z / skin depth), a
mirrored model gives a mirrored field, and the two modes are transposes of
each other on a transposed model."""
import os

import numpy as np
import pytest

import refdriver
import sp2oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_bcx0.npz")
AIR = np.log(1e-10)


def _split(e, nx, ny, nz):
    nxe, nye = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    ex = e[:nxe].reshape(nz + 1, ny + 1, nx)          # [k][j][i]
    ey = e[nxe:nxe + nye].reshape(nz + 1, ny, nx + 1)
    ez = e[nxe + nye:]
    return ex, ey, ez


def test_fixture_bit_exact():
    g = np.load(GOLD)
    dx, dy, dz, nza, lc = g["dx"], g["dy"], g["dz"], int(g["nzAir"]), g["logcond"]
    pb = O.Problem(dx, dy, dz, nza, lc)
    ei, eb, _, _ = pb.index()
    for s, (w, m) in enumerate(zip(g["omega"], g["mode"])):
        e0, nsw = O.bc_x0(dx, dy, dz, nza, lc, w, int(m))
        assert nsw == 0
        assert np.array_equal(e0, g["e0"][s]), s
        bc = O.bc_form(e0, len(dx), len(dy), len(dz))
        assert np.array_equal(bc[eb], g["bc"][s][eb]), s     # setBC of the zeroed-bottom BC
        assert not np.any(g["bc"][s][ei])
    # nz = 70: the reference's LAPACK takes its blocked factorisation there
    dx, dy, dz, nza, lc = (g["deep_dx"], g["deep_dy"], g["deep_dz"], int(g["deep_nzAir"]),
                           g["deep_logcond"])
    nx, ny, nz = len(dx), len(dy), len(dz)
    nxe, nye = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    e1, n1 = O.bc_x0(dx, dy, dz, nza, lc, g["deep_omega"][0], int(g["deep_mode"][0]))
    e2, n2 = O.bc_x0(dx, dy, dz, nza, lc, g["deep_omega"][1], int(g["deep_mode"][1]))
    assert n1 == 0 and n2 == 0
    assert np.array_equal(e1[nxe:nxe + nye], g["deep_ey"])
    assert np.array_equal(e2[:nxe], g["deep_ex"])


@pytest.mark.skipif(not refdriver.bcx0_available(), reason="reference build (oracle/_ref/ref_bcx0) absent")
def test_live_reference_bit_exact():
    rng = np.random.default_rng(5)
    for nx, ny, nza, nze in ((6, 9, 4, 7), (11, 5, 7, 12)):
        dx = rng.uniform(100, 900, nx)
        dy = rng.uniform(100, 900, ny)
        dz = np.concatenate([2000.0 * 1.5 ** np.arange(nza)[::-1], rng.uniform(30, 300, nze)])
        lc = np.log(10.0 ** rng.uniform(-3.5, 0.5, (nx, ny, nze)))
        sy = [(2 * np.pi * f, m) for f in (3e-4, 0.7, 150.0) for m in (1, 2)]
        e0, _ = refdriver.run_bcx0(dict(dx=dx, dy=dy, dz=dz, nzAir=nza), lc, AIR, sy)
        for s, (w, m) in enumerate(sy):
            eo, nsw = O.bc_x0(dx, dy, dz, nza, lc, w, m)
            assert nsw == 0
            assert np.array_equal(eo, e0[s]), (nx, s)


def _layered_grid(nx, ny, nza, dz_earth, rng):
    dx = rng.uniform(300, 800, nx)
    dy = rng.uniform(300, 800, ny)
    dz = np.concatenate([3000.0 * 1.5 ** np.arange(nza)[::-1], dz_earth])
    return dx, dy, dz


def test_layered_model_every_column_is_the_1d_solution():
    rng = np.random.default_rng(8)
    nx, ny, nza = 7, 5, 5
    dze = 60.0 * 1.15 ** np.arange(18)
    dx, dy, dz = _layered_grid(nx, ny, nza, dze, rng)
    nz = len(dz)
    layers = np.log(10.0 ** rng.uniform(-3, 0, len(dze)))
    lc = np.broadcast_to(layers, (nx, ny, len(dze))).copy()
    for m in (1, 2):
        e0, nsw = O.bc_x0(dx, dy, dz, nza, lc, 2 * np.pi * 0.3, m)
        assert nsw == 0
        ex, ey, ez = _split(e0, nx, ny, nz)
        assert not np.any(ez)
        f = ey if m == 1 else ex
        prof = f[:, :, 0] if m == 1 else f[:, 0, :]       # [k][slice]: the left column (a 1-D solve)
        cols = f if m == 1 else f.transpose(0, 2, 1)      # [k][slice][h]
        assert np.allclose(cols, prof[:, :, None], rtol=0, atol=1e-11 * np.abs(prof).max())
        other = ex if m == 1 else ey
        assert not np.any(other)


def test_halfspace_decays_with_the_skin_depth():
    """E'' = -i omega mu0 sigma E below the air: exp(-(1-i) z / delta) with
    delta = sqrt(2 / (omega mu0 sigma)), sampled on a fine vertical grid."""
    rng = np.random.default_rng(9)
    sig, f = 0.01, 1.0
    w = 2 * np.pi * f
    delta = np.sqrt(2.0 / (w * 4e-7 * np.pi * sig))
    dze = np.full(120, delta / 30.0)                  # 4 skin depths
    dx, dy, dz = _layered_grid(3, 3, 6, dze, rng)
    nza, nz = 6, len(dz)
    lc = np.full((3, 3, len(dze)), np.log(sig))
    e0, _ = O.bc_x0(dx, dy, dz, nza, lc, w, 1)
    _, ey, _ = _split(e0, 3, 3, nz)
    col = ey[nza:, 1, 1]                              # surface and below
    z = np.concatenate([[0.0], np.cumsum(dze)])
    keep = z <= 1.5 * delta
    want = np.exp(-(1 - 1j) * z[keep] / delta)
    got = col[keep] / col[0]
    assert np.max(np.abs(got - want)) < 3e-3
    # the conjugate convention would be off by far more
    assert np.max(np.abs(got - np.conj(want))) > 0.3


def test_mirrored_model_gives_a_mirrored_field():
    rng = np.random.default_rng(10)
    nx, ny, nza = 8, 4, 4
    half = rng.uniform(200, 900, nx // 2)
    dx = np.concatenate([half, half[::-1]])
    dy = rng.uniform(200, 900, ny)
    dz = np.concatenate([2500.0 * 1.5 ** np.arange(nza)[::-1], rng.uniform(40, 200, 9)])
    lh = np.log(10.0 ** rng.uniform(-3, 0, (nx // 2, ny, 9)))
    lc = np.concatenate([lh, lh[::-1]], axis=0)       # sigma(i) = sigma(nx+1-i)
    e0, _ = O.bc_x0(dx, dy, dz, nza, lc, 2 * np.pi * 2.0, 1)   # slices along y, horizontal x
    _, ey, _ = _split(e0, nx, ny, len(dz))
    assert np.allclose(ey, ey[:, :, ::-1], rtol=0, atol=1e-12 * np.abs(ey).max())
    assert np.abs(ey[:, :, 1] - ey[:, :, 2]).max() > 1e-3   # not trivially uniform


def test_modes_are_transposes_on_a_transposed_model():
    rng = np.random.default_rng(11)
    n, nza = 6, 3
    d = rng.uniform(200, 900, n)
    dz = np.concatenate([2500.0 * 1.5 ** np.arange(nza)[::-1], rng.uniform(40, 200, 7)])
    a = np.log(10.0 ** rng.uniform(-3, 0, (n, n, 7)))
    lc = 0.5 * (a + a.transpose(1, 0, 2))             # sigma(i, j) = sigma(j, i)
    w = 2 * np.pi * 0.5
    e1, _ = O.bc_x0(d, d, dz, nza, lc, w, 1)
    e2, _ = O.bc_x0(d, d, dz, nza, lc, w, 2)
    ex1, ey1, _ = _split(e1, n, n, len(dz))
    ex2, ey2, _ = _split(e2, n, n, len(dz))
    assert not np.any(ex1) and not np.any(ey2)
    # E0%y(ih, s, iz) of mode 1 = E0%x(s, ih, iz) of mode 2
    assert np.array_equal(ey1, ex2.transpose(0, 2, 1))
