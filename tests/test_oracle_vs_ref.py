"""The oracle (oracle/sp2_oracle.c) against the reference's own CPU SP2 path:
live against the flang build (oracle/_ref, where it was built) and always
against the committed fixture tests/golden/ref_small.npz written by it.
Every operator and factor must agree bit for bit."""
import os

import numpy as np
import pytest

import refdriver
import sp2oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_small.npz")


def _same_csr(o, row, col, val):
    return (np.array_equal(o["row"] + 1, row) and np.array_equal(o["col"] + 1, col)
            and np.array_equal(o["val"], val))


def _replay(p, g):
    """Replays the reference driver's call sequence on the oracle."""
    om = g["omegas"]
    p.update_freq(om[0])
    p.update_freq(om[1])
    return om


def test_fixture_operators_bit_exact():
    g = np.load(GOLD)
    p = sp2oracle.Problem(g["dx"], g["dy"], g["dz"], int(g["nzAir"]), g["logcond"])
    ei, eb, ni, _ = p.index()
    assert np.array_equal(ei + 1, g["EDGEi"]) and np.array_equal(eb + 1, g["EDGEb"])
    assert np.array_equal(ni + 1, g["NODEi"])
    se, sn = p.sigma()
    assert np.array_equal(se, g["SigEdge"]) and np.array_equal(sn, g["SigNode"])
    assert np.array_equal(p.metric()["Vedge"], g["Vedge"])
    for name in ("AAii", "CCib", "GDii"):
        assert _same_csr(p.csr(name), g[f"{name}_row"], g[f"{name}_col"], g[f"{name}_val"]), name
    om = g["omegas"]
    p.update_freq(om[0])
    assert np.array_equal(p.vms(), g["vms_0"])
    assert _same_csr(p.lu(0), g["L_0_row"], g["L_0_col"], g["L_0_val"])
    assert _same_csr(p.lu(1), g["U_0_row"], g["U_0_col"], g["U_0_val"])
    p.update_freq(om[1])
    assert np.array_equal(p.vms(), g["vms_1"])


def test_fixture_solvers_bit_exact():
    g = np.load(GOLD)
    p = sp2oracle.Problem(g["dx"], g["dy"], g["dz"], int(g["nzAir"]), g["logcond"])
    om = _replay(p, g)
    p.update_freq(om[0])
    x, info = p.bicg(g["b"], g["x0"], maxit=300, tol=1e-10, gpu_semantics=False)
    assert info["niter"] == int(g["bicg_niter"])
    assert np.array_equal(x, g["bicg_x"])
    assert np.array_equal(info["rerr"], g["bicg_rerr"])
    p.update_freq(om[1])
    assert np.array_equal(p.mult_aii(g["b"]), g["apply_Ax"])
    assert np.array_equal(p.pc_usolve(p.pc_lsolve(g["b"])), g["apply_LU"])
    p.divcor_setup()
    v = p.csr("VDsG")
    assert _same_csr(v, g["VDsG_row"], g["VDsG_col"], g["VDsG_val"])
    phi, pinfo = p.pcg(g["pcg_rhs"], maxit=200, tol=1e-10)
    assert pinfo["niter"] == int(g["pcg_niter"])
    assert np.array_equal(phi, g["pcg_phi"])
    assert np.array_equal(p.edge2model(g["e2m_in"]), g["e2m"])


@pytest.mark.skipif(not refdriver.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", [11, 12])
def test_live_reference_bit_exact(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nza, nze = 8 + seed % 3, 7, 3, 6
    dx = rng.uniform(50, 400, nx)
    dy = rng.uniform(50, 400, ny)
    dz = np.concatenate([[1e5, 1e4, 1e3], rng.uniform(20, 300, nze)])
    lc = np.log(10.0 ** rng.uniform(-4, 0, (nx, ny, nze)))
    grid = dict(dx=dx, dy=dy, dz=dz, nzAir=nza)
    p = sp2oracle.Problem(dx, dy, dz, nza, lc)
    n = p.nInt
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    om = [2 * np.pi * 3.0, 2 * np.pi * 1e-3]
    ref = refdriver.run(grid, lc, np.log(1e-10), om, solves=[(1, 400, 1e-9, b, np.zeros(n))],
                        applies=[(0, b)])
    for name in ("AAii", "CCii", "CCib", "GDii", "T", "G"):
        r = ref[name]
        assert _same_csr(p.csr(name), r["row"], r["col"], r["val"]), name
    for k, w in enumerate(om):
        p.update_freq(w)
        assert np.array_equal(p.vms(), ref[f"vms_{k}"])
        for which, nm in ((0, "L"), (1, "U")):
            r = ref[f"{nm}_{k}"]
            assert _same_csr(p.lu(which), r["row"], r["col"], r["val"]), nm
    p.update_freq(om[1])
    x, info = p.bicg(b, None, maxit=400, tol=1e-9, gpu_semantics=False)
    assert info["niter"] == ref["bicg_0"]["niter"]
    assert np.array_equal(x, ref["bicg_0"]["x"])
    p.update_freq(om[0])
    assert np.array_equal(p.mult_aii(b), ref["apply_0"]["Ax"])
    assert np.array_equal(p.pc_usolve(p.pc_lsolve(b)), ref["apply_0"]["LUsolve"])


QMR_GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_qmr.npz")


def _qmr_check(p, om, b, x0, maxit, tol, ref):
    """Replays the driver's UpdateFreq sequence (every frequency once, then
    one per QMR record, then one per adjoint record) and compares."""
    for w in om:
        p.update_freq(w)
    for k, w in enumerate(om):
        p.update_freq(w)
        x, info = p.qmr(b, x0, maxit=maxit, tol=tol)
        assert info["niter"] == ref[f"qmr_niter_{k}"], k
        assert np.array_equal(x, ref[f"qmr_x_{k}"]), k
        assert np.array_equal(info["rerr"], ref[f"qmr_rerr_{k}"]), k
    for k, (w, v) in enumerate(zip(om, (b, x0))):
        p.update_freq(w)
        assert np.array_equal(p.mult_aii(v, adjt=True), ref[f"adj_AHx_{k}"]), k
        assert np.array_equal(p.pc_lsolve_h(v), ref[f"adj_LH_{k}"]), k
        assert np.array_equal(p.pc_usolve_h(v), ref[f"adj_UH_{k}"]), k


def test_fixture_qmr_bit_exact():
    """QMR (solver.f90:140-370) and the adjoint applications it uses
    (Mult_Aii adjt, PC_Lsolve/PC_Usolve adjt) against the reference's."""
    g = np.load(QMR_GOLD)
    p = sp2oracle.Problem(g["dx"], g["dy"], g["dz"], int(g["nzAir"]), g["logcond"])
    _qmr_check(p, g["omegas"], g["b"], g["x0"], int(g["maxit"]), float(g["tol"]), g)
    # one of the two systems converges, the other is cut at maxit or converges
    assert int(g["qmr_niter_0"]) < int(g["maxit"])


@pytest.mark.skipif(not refdriver.available(), reason="reference build not present")
def test_live_qmr_bit_exact():
    rng = np.random.default_rng(17)
    nx, ny, nza, nze = 9, 8, 3, 6
    dx = rng.uniform(100, 400, nx)
    dy = rng.uniform(100, 400, ny)
    dz = np.concatenate([[8000.0, 1500.0, 300.0], rng.uniform(40, 250, nze)])
    lc = np.log(rng.uniform(0.002, 0.5, (nx, ny, nze)))
    grid = dict(dx=dx, dy=dy, dz=dz, nzAir=nza)
    p = sp2oracle.Problem(dx, dy, dz, nza, lc)
    n = p.nInt
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = np.zeros(n, complex)
    om = [2 * np.pi * 3.0, 2 * np.pi * 0.05]
    maxit, tol = 40, 1e-9
    out = refdriver.run(grid, lc, np.log(1e-10), om, qmr=[(0, maxit, tol, b, x0), (1, maxit, tol, b, x0)],
                        adj_applies=[(0, b), (1, x0 + b)], dump=False)
    ref = {}
    for k in range(2):
        ref.update({f"qmr_x_{k}": out[f"qmr_{k}"]["x"], f"qmr_niter_{k}": out[f"qmr_{k}"]["niter"],
                    f"qmr_rerr_{k}": out[f"qmr_{k}"]["rerr"],
                    f"adj_AHx_{k}": out[f"adj_{k}"]["AHx"], f"adj_LH_{k}": out[f"adj_{k}"]["LHsolve"],
                    f"adj_UH_{k}": out[f"adj_{k}"]["UHsolve"]})
    for w in om:
        p.update_freq(w)
    for k, w in enumerate(om):
        p.update_freq(w)
        x, info = p.qmr(b, x0, maxit=maxit, tol=tol)
        assert info["niter"] == ref[f"qmr_niter_{k}"] and np.array_equal(x, ref[f"qmr_x_{k}"])
        assert np.array_equal(info["rerr"], ref[f"qmr_rerr_{k}"])
    for k, (w, v) in enumerate(zip(om, (b, x0 + b))):
        p.update_freq(w)
        assert np.array_equal(p.mult_aii(v, adjt=True), ref[f"adj_AHx_{k}"])
        assert np.array_equal(p.pc_lsolve_h(v), ref[f"adj_LH_{k}"])
        assert np.array_equal(p.pc_usolve_h(v), ref[f"adj_UH_{k}"])
