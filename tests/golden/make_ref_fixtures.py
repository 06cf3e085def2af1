"""Writes tests/golden/ref_small.npz: what the reference's own CPU SP2 path
(the flang build of the reference sources, oracle/ref/Makefile) computes on a
small seeded grid.  Inputs and outputs only -- data, no reference code.

Run where the reference is mounted:  python tests/golden/make_ref_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import refdriver  # noqa: E402


def case():
    rng = np.random.default_rng(2024)
    nx, ny, nza, nze = 7, 6, 3, 5
    dx = rng.uniform(100, 300, nx)
    dy = rng.uniform(100, 300, ny)
    dz = np.concatenate([[5000.0, 1000.0, 200.0], rng.uniform(50, 200, nze)])
    lc = np.log(rng.uniform(0.001, 0.1, (nx, ny, nze)))
    return rng, dx, dy, dz, nza, lc


def main():
    rng, dx, dy, dz, nza, lc = case()
    grid = dict(dx=dx, dy=dy, dz=dz, nzAir=nza)
    om = [2 * np.pi * 1.0, 2 * np.pi * 0.01]
    first = refdriver.run(grid, lc, np.log(1e-10), om[:1])
    n, nni, ne = len(first["EDGEi"]), len(first["NODEi"]), len(first["EdgeL"])
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    rhs = rng.standard_normal(nni) + 0j
    e2m = rng.standard_normal(ne)
    out = refdriver.run(grid, lc, np.log(1e-10), om, solves=[(0, 300, 1e-10, b, x0)],
                        applies=[(1, b)], divcor=(200, 1e-10, rhs), e2m=e2m)
    d = dict(dx=dx, dy=dy, dz=dz, nzAir=nza, logcond=lc, omegas=np.array(om), b=b, x0=x0,
             pcg_rhs=rhs, e2m_in=e2m, EDGEi=out["EDGEi"], EDGEb=out["EDGEb"],
             NODEi=out["NODEi"], SigEdge=out["SigEdge"], SigNode=out["SigNode"],
             Vedge=out["Vedge"], vms_0=out["vms_0"], vms_1=out["vms_1"],
             bicg_x=out["bicg_0"]["x"], bicg_niter=out["bicg_0"]["niter"],
             bicg_rerr=out["bicg_0"]["rerr"], apply_Ax=out["apply_0"]["Ax"],
             apply_LU=out["apply_0"]["LUsolve"], pcg_phi=out["pcg"]["phi"],
             pcg_niter=out["pcg"]["niter"], e2m=out["e2m"])
    for name in ("AAii", "CCib", "GDii", "L_0", "U_0", "VDsG"):
        for k in ("row", "col", "val"):
            d[f"{name}_{k}"] = out[name][k]
    np.savez_compressed(os.path.join(HERE, "ref_small.npz"), **d)
    print("wrote", os.path.join(HERE, "ref_small.npz"))


if __name__ == "__main__":
    main()
