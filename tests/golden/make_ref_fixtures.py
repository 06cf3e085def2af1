"""Writes tests/golden/ref_small.npz: what the reference's own CPU SP2 path
(the flang build of the reference sources, oracle/ref/Makefile) computes on a
small seeded grid, tests/golden/ref_datafunc.npz: its dataResp and LmultT
for every data type on the same grid, and tests/golden/ref_bcx0.npz: the 2-D
TE boundary fields of BC_x0_WS (through ComputeBC) on that grid and on a deep
one whose bandwidth takes LAPACK's blocked factorisation, and
tests/golden/io/: small WS model files with what the reference's reader reads
from them (ref_io.npz), and a solution file written by the reference's
writer with its inputs, and tests/golden/ref_jmultT.npz: the reference's
sensitivity chain for one MT transmitter at two frequencies (forward fields,
LmultT sources, the TRN solves of sensSolve, PmultT and JmultT_TX) on a small
seeded grid, and tests/golden/ref_qmr.npz: its CPU QMR solves and the
adjoint applications (A^H, L^-H, U^-H) they use.  Inputs and outputs only --
data, no reference code.

Run where the reference is mounted:
    python tests/golden/make_ref_fixtures.py [small] [datafunc] [bcx0] [io] [jmultT] [qmr]
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


DF_NCOMP = {1: 8, 2: 4, 3: 4, 4: 8, 5: 4, 6: 4, 7: 2, 8: 2, 9: 2, 10: 2, 11: 2}


def datafunc_case():
    """Sites on the case() grid: one exactly on an x/y node (minNode ties go
    through clean()), one at the Earth surface, one with non-orthogonal
    channel azimuths; reference sites for the interstation type."""
    rng, dx, dy, dz, nza, lc = case()
    rng = np.random.default_rng(77)
    origin = np.array([-dx.sum() / 2, -dy.sum() / 2, 0.0])
    xe = origin[0] + np.concatenate([[0.0], np.cumsum(dx)])
    ye = origin[1] + np.concatenate([[0.0], np.cumsum(dy)])
    x = np.array([[xe[3], ye[3], 20.0],
                  [0.5 * (xe[2] + xe[3]), 0.3 * ye[2] + 0.7 * ye[3], 0.0],
                  [0.2 * xe[4] + 0.8 * xe[3], 0.6 * ye[4] + 0.4 * ye[3], 75.0],
                  [0.9 * xe[2] + 0.1 * xe[3], 0.5 * (ye[2] + ye[3]), 130.0]])
    r = x[[2, 3, 0, 1]].copy()
    az = np.tile([0.0, 90.0, 0.0, 0.0, 90.0, 0.0, 90.0], (4, 1))
    az[2] = [12.0, 97.0, 0.0, -8.0, 85.0, 5.0, 100.0]
    az[3] = [30.0, 120.0, 0.0, 30.0, 120.0, 30.0, 120.0]
    ne = 7 * 7 * 9 + 8 * 6 * 9 + 8 * 7 * 8
    e1 = rng.standard_normal(ne) + 1j * rng.standard_normal(ne)
    e2 = rng.standard_normal(ne) + 1j * rng.standard_normal(ne)
    return dict(dx=dx, dy=dy, dz=dz, nzAir=nza, origin=origin, omega=2 * np.pi * 0.05,
                x=x, r=r, az=az, e1=e1, e2=e2), rng


def datafunc_main():
    c, rng = datafunc_case()
    grid = dict(dx=c["dx"], dy=c["dy"], dz=c["dz"], nzAir=c["nzAir"])
    out = dict(c)
    for dt, nc in DF_NCOMP.items():
        d = rng.standard_normal((4, nc))
        ex = np.ones((4, nc), np.int32)
        ex[1, 0] = 0
        resp, c1, c2 = refdriver.run_datafunc(grid, c["origin"], c["omega"], dt, c["x"], c["e1"],
                                              c["e2"], d, r=c["r"], az=c["az"], exist=ex)
        out.update({f"d_{dt}": d, f"exist_{dt}": ex, f"resp_{dt}": resp, f"comb1_{dt}": c1,
                    f"comb2_{dt}": c2})
    np.savez_compressed(os.path.join(HERE, "ref_datafunc.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_datafunc.npz"))


BC_SYSTEMS = [(2 * np.pi * 0.05, 1), (2 * np.pi * 0.05, 2), (2 * np.pi * 20.0, 1),
              (2 * np.pi * 1e-3, 2)]


def bcx0_deep_case():
    """nz = 70 (10 air + 60 earth): bandwidth 69 > 64, LAPACK's blocked ZGBTRF."""
    rng = np.random.default_rng(31)
    nx, ny, nza, nze = 4, 3, 10, 60
    dx = rng.uniform(200, 2000, nx)
    dy = rng.uniform(200, 2000, ny)
    dz = np.concatenate([1000.0 * 1.6 ** np.arange(nza)[::-1], 40.0 * 1.08 ** np.arange(nze)])
    lc = np.log(rng.uniform(0.001, 1.0, (nx, ny, nze)))
    return dx, dy, dz, nza, lc


def bcx0_main():
    _, dx, dy, dz, nza, lc = case()
    air = np.log(1e-10)
    e0, bc = refdriver.run_bcx0(dict(dx=dx, dy=dy, dz=dz, nzAir=nza), lc, air, BC_SYSTEMS)
    out = dict(dx=dx, dy=dy, dz=dz, nzAir=nza, logcond=lc, air=air,
               omega=np.array([w for w, _ in BC_SYSTEMS]), mode=np.array([m for _, m in BC_SYSTEMS]),
               e0=e0, bc=bc)
    ddx, ddy, ddz, dnza, dlc = bcx0_deep_case()
    sy = [(2 * np.pi * 1e-2, 1), (2 * np.pi * 3.0, 2)]
    de0, _ = refdriver.run_bcx0(dict(dx=ddx, dy=ddy, dz=ddz, nzAir=dnza), dlc, air, sy)
    nx, ny, nz = len(ddx), len(ddy), len(ddz)
    nxe, nye = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    out.update(deep_dx=ddx, deep_dy=ddy, deep_dz=ddz, deep_nzAir=dnza, deep_logcond=dlc,
               deep_omega=np.array([w for w, _ in sy]), deep_mode=np.array([m for _, m in sy]),
               deep_ey=de0[0][nxe:nxe + nye], deep_ex=de0[1][:nxe])   # the only non-zero parts
    np.savez_compressed(os.path.join(HERE, "ref_bcx0.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_bcx0.npz"))


# a WS file written by hand with list-directed freedoms: a value list split
# over lines, commas, r*c repeats, D exponents, surplus values at the end of
# a record, and no origin or rotation lines
WS_QUIRKS = """# hand-written test model (10 earth layers: the default air layers mirror them)
2 2 10 0 LOG10
 1.0D+03, 500.0
 250.5
 300. 200. 999.0 888.0
 2*10.0, 12.5 15.0
 20.0 25.0 3*30.0
 40.0 99.0
 1.5 , 2.5, 3.5 7.7
 2*1.0 0.5
 0.25D0 -0.5
 2*-1.0
 1.0 2.0
 3.0 4.0
 0.5 0.5
 0.1 0.2
 1.25 2.75
 3*2.0
 -1.5 1.5
 2.2 3.3 4.4
 1.0
 0.0
 5.0 6.0
 7.0 8.0
 9.0 1.0
"""

AIR_FIXED = ("fixed height", 12, 1000.0, 30.0, 3.0)


def zlist_text():
    """A list data file in five blocks exercising read_Z_list's MT branches:
    azimuth columns (12-15 and 10-14), both sign conventions, Ohm and
    [mV/km]/[nT] units, a geographic orientation, transmitters and sites
    shared across blocks."""
    rng = np.random.default_rng(44)
    sites = [("S01", 45.123, -97.5, 100.0, -2000.0, 0.0), ("S02", 45.2, -97.61, 2100.5, 350.0, 12.0),
             ("LONGSITENAME-0003", 44.9, -98.0, -1500.0, 800.0, 0.0)]
    hdr = "# Period(s) Code GG_Lat GG_Lon X(m) Y(m) Z(m) Component Real Imag Error"

    def cplx_block(name, comps, sign, units, orient, periods, extra):
        out = [f"# synthetic {name}", hdr, f"> {name}", f"> exp({sign}i\\omega t)", f"> {units}",
               f">   {orient:.2f}", ">   45.000  -97.000", f">  {len(periods)}  {len(sites)}"]
        for p in periods:
            for si, (c, la, lo, x, y, z) in enumerate(sites):
                for ci, cp in enumerate(comps):
                    re_, im_, er = rng.standard_normal(2).tolist() + [abs(rng.standard_normal()) + 0.1]
                    az = " ".join(f"{a:.2f}" for a in extra[(si + ci) % len(extra)])
                    out.append(f"{p:.6E} {c} {la:.3f} {lo:.3f} {x:.3f} {y:.3f} {z:.3f} {cp} "
                               f"{re_:.6E} {im_:.6E} {er:.6E} {az}".rstrip())
        return out

    def real_block(name, comps, sign, orient, periods, extra):
        out = [f"# synthetic {name}", hdr, f"> {name}", f"> exp({sign}i\\omega t)", "> []",
               f">   {orient:.2f}", ">   45.000  -97.000", f">  {len(periods)}  {len(sites)}"]
        for p in periods:
            for si, (c, la, lo, x, y, z) in enumerate(sites):
                for ci, cp in enumerate(comps):
                    v = 10.0 ** rng.uniform(0, 3) if cp.startswith("RHO") else rng.uniform(5, 80)
                    er = 0.05 * v
                    az = " ".join(f"{a:.2f}" for a in extra[(si + ci) % len(extra)])
                    out.append(f"{p:.6E} {c} {la:.3f} {lo:.3f} {x:.3f} {y:.3f} {z:.3f} {cp} "
                               f"{v:.6E} {er:.6E} {az}".rstrip())
        return out
    ext = [(), (12.0,), (12.0, 98.0), (12.0, 98.0, 8.0), (12.0, 98.0, 8.0, 101.0)]
    lines = []
    lines += cplx_block("Full_Impedance", ["ZXX", "ZXY", "ZYX", "ZYY"], "-", "[V/m]/[A/m]", 10.0,
                        [1.0, 100.0], ext)
    lines += real_block("Off_Diagonal_Rho_Phase", ["RHOXY", "PHSXY", "RHOYX", "PHSYX"], "+", 0.0,
                        [100.0, 3.16], ext)
    lines += cplx_block("Full_Vertical_Components", ["TX", "TY"], "+", "[]", 0.0, [1.0], ext[:1])
    lines += real_block("Phase_Tensor", ["PTXX", "PTXY", "PTYX", "PTYY"], "-", -5.0, [3.16], ext)
    lines += cplx_block("Off_Diagonal_Impedance", ["ZXY", "ZYX"], "+", "[mV/km]/[nT]", 0.0,
                        [1000.0], ext[:1])
    return "\n".join(lines) + "\n"


def _pack_zlist(prefix, z, out):
    out[prefix + "periods"] = np.asarray(z["periods"])
    out[prefix + "rx_ids"] = np.array([r[0] for r in z["receivers"]])
    out[prefix + "rx_x"] = np.array([r[1] for r in z["receivers"]])
    out[prefix + "nd"] = np.array(len(z["data"]))
    for i, d in enumerate(z["data"]):
        out[f"{prefix}d{i}_tx"] = np.array(d["tx"])
        out[f"{prefix}d{i}_nb"] = np.array(len(d["blocks"]))
        for j, b in enumerate(d["blocks"]):
            for k in ("type", "ncomp", "rx", "value", "error", "exist", "azimuth"):
                out[f"{prefix}d{i}_b{j}_{k}"] = np.asarray(b[k])


def io_main():
    d = os.path.join(HERE, "io")
    os.makedirs(d, exist_ok=True)
    sys.path.insert(0, os.path.join(ROOT, "modem-gpu_amd"))
    from mtsp2 import formats
    rng = np.random.default_rng(42)
    nx, ny, nze = 5, 4, 12     # the default air layers mirror the top 10 earth layers
    dx = np.round(rng.uniform(100, 900, nx), 3)
    dy = np.round(rng.uniform(100, 900, ny), 3)
    dze = np.round(25.0 * 1.3 ** np.arange(nze), 3)
    lc = np.log(10.0 ** rng.uniform(-3, 0, (nx, ny, nze)))
    files = {}
    for pt in ("LOGE", "LOG10", "LINEAR"):
        name = f"ws_{pt.lower()}.rho"
        formats.write_ws_model(os.path.join(d, name), dx, dy, dze, lc, pt,
                               origin=(-1500.0, -1200.0, 0.0), rotdeg=7.5)
        files[name] = None
    with open(os.path.join(d, "ws_quirks.rho"), "w") as f:
        f.write(WS_QUIRKS)
    files["ws_quirks.rho"] = None
    out = {}
    for name in files:
        for tag, air in (("default", None), ("fixed", AIR_FIXED)):
            r = refdriver.run_io_ws(os.path.join(d, name), air=air)
            key = f"{name[:-4]}_{tag}"
            out.update({f"{key}_{k}": np.asarray(v) for k, v in r.items()})
    # a solution file from the reference's own writer
    srng = np.random.default_rng(43)
    sdx, sdy = srng.uniform(100, 500, 4), srng.uniform(100, 500, 3)
    sdz = np.concatenate([[3000.0, 800.0], srng.uniform(50, 300, 3)])
    ne = sum(formats.edge_counts(4, 3, 5))
    om = np.array([2 * np.pi * 0.1, 2 * np.pi * 7.0])
    F = srng.standard_normal((2, 2, ne)) + 1j * srng.standard_normal((2, 2, ne))
    refdriver.run_io_soln_write(os.path.join(d, "ref_small.soln"), sdx, sdy, sdz, 2, om, F,
                                origin=(-600.0, -450.0, 0.0), rotdeg=12.5)
    out.update(soln_dx=sdx, soln_dy=sdy, soln_dz=sdz, soln_omegas=om, soln_fields=F)
    # a list data file read by the reference's read_Z_list
    zpath = os.path.join(d, "zlist_mixed.dat")
    with open(zpath, "w") as f:
        f.write(zlist_text())
    _pack_zlist("zl_", refdriver.run_io_zlist(zpath), out)
    np.savez_compressed(os.path.join(d, "ref_io.npz"), **out)
    print("wrote", d)


def qmr_main():
    """The reference's CPU QMR (solver.f90:140-370) and the adjoint
    applications it uses, on the small grid of case(), two frequencies."""
    rng, dx, dy, dz, nza, lc = case()
    grid = dict(dx=dx, dy=dy, dz=dz, nzAir=nza)
    om = [2 * np.pi * 1.0, 2 * np.pi * 0.01]
    n = len(refdriver.run(grid, lc, np.log(1e-10), om[:1], dump=True)["EDGEi"])
    rng = np.random.default_rng(99)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    out = refdriver.run(grid, lc, np.log(1e-10), om, qmr=[(0, 300, 1e-10, b, x0), (1, 300, 1e-10, b, x0)],
                        adj_applies=[(0, b), (1, x0)], dump=False)
    d = dict(dx=dx, dy=dy, dz=dz, nzAir=nza, logcond=lc, omegas=np.array(om), b=b, x0=x0, maxit=300,
             tol=1e-10)
    for k in range(2):
        q = out[f"qmr_{k}"]
        d.update({f"qmr_x_{k}": q["x"], f"qmr_niter_{k}": q["niter"], f"qmr_rerr_{k}": q["rerr"],
                  f"qmr_failed_{k}": q["failed"]})
        a = out[f"adj_{k}"]
        d.update({f"adj_AHx_{k}": a["AHx"], f"adj_LH_{k}": a["LHsolve"], f"adj_UH_{k}": a["UHsolve"]})
    np.savez_compressed(os.path.join(HERE, "ref_qmr.npz"), **d)
    print("wrote", os.path.join(HERE, "ref_qmr.npz"))


def jmult_case():
    """Grid, model, sites and full-impedance data of the sensitivity fixture."""
    sys.path.insert(0, os.path.join(ROOT, "modem-gpu_amd"))
    from mtsp2 import problem as P
    rng = np.random.default_rng(77)
    grid = P.make_grid(10, 9, 6, nair=3, dcore=300.0, npad=2)
    nx, ny = len(grid["dx"]), len(grid["dy"])
    nze = len(grid["dz"]) - grid["nzAir"]
    lc = np.log(0.01) + 0.8 * rng.standard_normal((nx, ny, nze))
    sites = P.core_sites(grid, 3, npad=2)
    omegas = 2 * np.pi * np.array([0.05, 20.0])
    d = rng.standard_normal((len(omegas), len(sites), 8))
    return grid, lc, sites, omegas, d


def jmult_main():
    grid, lc, sites, omegas, d = jmult_case()
    out = dict(dx=grid["dx"], dy=grid["dy"], dz=grid["dz"], nzAir=grid["nzAir"], logcond=lc,
               sites=sites, omegas=omegas, d=d, tol=1e-10, maxit=2000)
    for k, w in enumerate(omegas):
        r = refdriver.run_jmultT(grid, lc, np.log(1e-10), w, sites, d[k], maxit=2000, tol=1e-10)
        for name, v in r.items():
            out[f"{name}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "ref_jmultT.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_jmultT.npz"))


if __name__ == "__main__":
    which = sys.argv[1:] or ["small", "datafunc", "bcx0", "io", "jmultT", "qmr"]
    if "qmr" in which:
        qmr_main()
    if "jmultT" in which:
        jmult_main()
    if "small" in which:
        main()
    if "datafunc" in which:
        datafunc_main()
    if "bcx0" in which:
        bcx0_main()
    if "io" in which:
        io_main()
