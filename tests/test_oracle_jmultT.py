"""The oracle's sensitivity chain against the reference's own (the flang
build of SENS/SensComp.f90 JmultT_TX, driver oracle/ref/ref_jmultT.f90):
forward fields, LmultT sources, the TRN solves of sensSolve (which pin the
oracle's FWDsolve3D TRN right-hand side sp2o_rhs_trn and post-processing
sp2o_post_trn: a wrong sign of iOmegaMuInv, a dropped GDii term or a V_E
divide for a multiply changes the TRN solution) and PmultT, bit for bit;
against the committed fixture tests/golden/ref_jmultT.npz and, where the
reference build is present, live on another grid and model."""
import os

import numpy as np
import pytest

import refdriver
from cfg_oracle import jmult_chain
from mtsp2 import problem as P

HERE = os.path.dirname(os.path.abspath(__file__))


def _check(got, ref):
    for k in ("fwd", "comb", "trn", "dsig"):
        assert np.array_equal(got[k], ref[k]), k


def test_chain_matches_reference_fixture():
    z = np.load(os.path.join(HERE, "golden", "ref_jmultT.npz"))
    grid = {"dx": z["dx"], "dy": z["dy"], "dz": z["dz"], "nzAir": int(z["nzAir"])}
    for k, w in enumerate(z["omegas"]):
        got = jmult_chain(grid, z["logcond"], float(w), z["sites"], z["d"][k], tol=float(z["tol"]),
                          maxit=int(z["maxit"]))
        _check(got, {n: z[f"{n}_{k}"] for n in ("fwd", "comb", "trn", "dsig")})
        # JmultT_TX itself gives the step-by-step chain's gradient (QmultT = 0 for impedances)
        assert np.array_equal(z[f"dsig_jt_{k}"], z[f"dsig_{k}"])
        assert np.abs(z[f"dsig_{k}"]).max() > 0


@pytest.mark.skipif(not refdriver.jmultT_available(), reason="reference build not present")
def test_chain_matches_reference_live():
    grid = P.make_grid(12, 11, 7, nair=4, dcore=250.0, npad=3)
    nx, ny = len(grid["dx"]), len(grid["dy"])
    nze = len(grid["dz"]) - grid["nzAir"]
    rng = np.random.default_rng(31)
    lc = np.log(0.02) + 0.6 * rng.standard_normal((nx, ny, nze))
    sites = P.core_sites(grid, 4, npad=3)
    d = rng.standard_normal((len(sites), 8))
    w = 2 * np.pi * 3.0
    ref = refdriver.run_jmultT(grid, lc, np.log(1e-10), w, sites, d)
    _check(jmult_chain(grid, lc, w, sites, d), ref)
    assert np.array_equal(ref["dsig"], ref["dsig_jt"])
