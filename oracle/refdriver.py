"""Runs the reference's own CPU SP2 path (built by oracle/ref/Makefile into
oracle/_ref/modem_ref) on one grid/model and reads back what it computed.

Test infrastructure only: tests use it as a checker for the oracle
restatement and, through the golden fixtures, for the HIP library.  The
driver program is oracle/ref/ref_driver.f90; its input layout is documented
there.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "modem_ref")
DF_BIN = os.path.join(HERE, "_ref", "ref_datafunc")
DF_DROPIN_BIN = os.path.join(HERE, "_ref", "ref_dropin_datafunc")
BCX0_BIN = os.path.join(HERE, "_ref", "ref_bcx0")
BC_DROPIN_BIN = os.path.join(HERE, "_ref", "ref_dropin_bc")


def available() -> bool:
    return os.path.isfile(REF_BIN) and os.access(REF_BIN, os.X_OK)


def _rd_csr(path, cplx=False):
    with open(path, "rb") as f:
        nrow, ncol, nnz = np.fromfile(f, np.int32, 3)
        row = np.fromfile(f, np.int32, nrow + 1)
        col = np.fromfile(f, np.int32, nnz)
        val = np.fromfile(f, np.complex128 if cplx else np.float64, nnz)
    return {"nrow": int(nrow), "ncol": int(ncol), "row": row, "col": col, "val": val}


def run(grid, logcond, air_log, omegas, solves=(), applies=(), divcor=None,
        e2m=None, workdir=None, timeout=600, dump=True, qmr=(), adj_applies=()):
    """grid: dict(dx, dy, dz, nzAir); logcond: (nx, ny, nzEarth) natural-log
    conductivities.  solves: list of (iomega0, maxit, tol, b, x0) for BICG,
    qmr: the same for QMR.  applies: list of (iomega0, x) (A x, L^-1 x,
    U^-1 L^-1 x); adj_applies: the same for the adjoint forms (A^H x, L^-H x,
    U^-H x).  divcor: (maxit, tol, rhs) or None.  e2m: real all-edge vector or
    None.  Returns a dict of numpy arrays."""
    dx = np.asarray(grid["dx"], np.float64)
    dy = np.asarray(grid["dy"], np.float64)
    dz = np.asarray(grid["dz"], np.float64)
    nx, ny, nz = len(dx), len(dy), len(dz)
    nza = int(grid["nzAir"])
    nze = nz - nza
    lc = np.asarray(logcond, np.float64)
    assert lc.shape == (nx, ny, nze)
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="mtref_") if own else workdir
    inp = os.path.join(wd, "input.bin")
    with open(inp, "wb") as f:
        np.array([1 if dump else 0], np.int32).tofile(f)
        np.array([nx, ny, nza, nze], np.int32).tofile(f)
        dx.tofile(f); dy.tofile(f); dz.tofile(f)
        lc.ravel(order="F").tofile(f)
        np.array([air_log], np.float64).tofile(f)
        np.array([len(omegas)], np.int32).tofile(f)
        np.asarray(omegas, np.float64).tofile(f)
        np.array([len(solves)], np.int32).tofile(f)
        for (iom, maxit, tol, b, x0) in solves:
            np.array([iom + 1, maxit], np.int32).tofile(f)
            np.array([tol], np.float64).tofile(f)
            np.asarray(b, np.complex128).tofile(f)
            np.asarray(x0, np.complex128).tofile(f)
        np.array([len(applies)], np.int32).tofile(f)
        for (iom, x) in applies:
            np.array([iom + 1], np.int32).tofile(f)
            np.asarray(x, np.complex128).tofile(f)
        if divcor is None:
            np.array([0], np.int32).tofile(f)
        else:
            maxit, tol, rhs = divcor
            np.array([1, maxit], np.int32).tofile(f)
            np.array([tol], np.float64).tofile(f)
            np.asarray(rhs, np.complex128).tofile(f)
        if e2m is None:
            np.array([0], np.int32).tofile(f)
        else:
            np.array([1], np.int32).tofile(f)
            np.asarray(e2m, np.float64).tofile(f)
        np.array([len(qmr)], np.int32).tofile(f)
        for (iom, maxit, tol, b, x0) in qmr:
            np.array([iom + 1, maxit], np.int32).tofile(f)
            np.array([tol], np.float64).tofile(f)
            np.asarray(b, np.complex128).tofile(f)
            np.asarray(x0, np.complex128).tofile(f)
        np.array([len(adj_applies)], np.int32).tofile(f)
        for (iom, x) in adj_applies:
            np.array([iom + 1], np.int32).tofile(f)
            np.asarray(x, np.complex128).tofile(f)
    subprocess.run([REF_BIN, inp, wd], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    out = {}
    nall, nint, nb, nnall, nni, nnb = np.fromfile(os.path.join(wd, "sizes.bin"), np.int32, 6)
    out["sizes"] = (int(nall), int(nint), int(nb), int(nnall), int(nni), int(nnb))
    nfx, nfy, nfz = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    nface = nfx + nfy + nfz
    for k in range(len(solves)):
        with open(os.path.join(wd, f"bicg_{k+1}.bin"), "rb") as f:
            niter, failed, maxit = np.fromfile(f, np.int32, 3)
            secs = float(np.fromfile(f, np.float64, 1)[0])
            rerr = np.fromfile(f, np.float64, maxit)
            x = np.fromfile(f, np.complex128, nint)
        out[f"bicg_{k}"] = {"niter": int(niter), "failed": bool(failed), "rerr": rerr, "x": x,
                            "seconds": secs}
    for k in range(len(qmr)):
        with open(os.path.join(wd, f"qmr_{k+1}.bin"), "rb") as f:
            niter, failed, maxit = np.fromfile(f, np.int32, 3)
            rerr = np.fromfile(f, np.float64, maxit)
            x = np.fromfile(f, np.complex128, nint)
        out[f"qmr_{k}"] = {"niter": int(niter), "failed": bool(failed), "rerr": rerr, "x": x}
    for k in range(len(adj_applies)):
        a = np.fromfile(os.path.join(wd, f"adj_{k+1}.bin"), np.complex128)
        out[f"adj_{k}"] = {"AHx": a[:nint], "LHsolve": a[nint:2 * nint], "UHsolve": a[2 * nint:]}
    if not dump:
        _cleanup(wd, own)
        return out
    with open(os.path.join(wd, "index.bin"), "rb") as f:
        np.fromfile(f, np.int32, 6)
        out["EDGEi"] = np.fromfile(f, np.int32, nint)
        out["EDGEb"] = np.fromfile(f, np.int32, nb)
        out["NODEi"] = np.fromfile(f, np.int32, nni)
        out["NODEb"] = np.fromfile(f, np.int32, nnb)
    with open(os.path.join(wd, "metric.bin"), "rb") as f:
        out["FaceA"] = np.fromfile(f, np.float64, nface)
        out["EdgeL"] = np.fromfile(f, np.float64, nall)
        out["DualFaceA"] = np.fromfile(f, np.float64, nall)
        out["DualEdgeL"] = np.fromfile(f, np.float64, nface)
        out["Vnode"] = np.fromfile(f, np.float64, nnall)
        out["Vedge"] = np.fromfile(f, np.float64, nall)
    with open(os.path.join(wd, "sigma.bin"), "rb") as f:
        out["SigEdge"] = np.fromfile(f, np.float64, nall)
        out["SigNode"] = np.fromfile(f, np.float64, nnall)
    for name in ("AAii", "CCii", "CCib", "GDii", "T", "G"):
        out[name] = _rd_csr(os.path.join(wd, name + ".csr"))
    for k in range(len(omegas)):
        out[f"vms_{k}"] = np.fromfile(os.path.join(wd, f"vms_{k+1}.bin"), np.float64)
        out[f"L_{k}"] = _rd_csr(os.path.join(wd, f"L_{k+1}.csr"), cplx=True)
        out[f"U_{k}"] = _rd_csr(os.path.join(wd, f"U_{k+1}.csr"), cplx=True)
    for k in range(len(applies)):
        a = np.fromfile(os.path.join(wd, f"apply_{k+1}.bin"), np.complex128)
        out[f"apply_{k}"] = {"Ax": a[:nint], "Lsolve": a[nint:2 * nint], "LUsolve": a[2 * nint:]}
    if divcor is not None:
        for name in ("VDiv", "VDs", "Gai", "VDsG", "VDsG_L", "VDsG_U"):
            out[name] = _rd_csr(os.path.join(wd, name + ".csr"))
        with open(os.path.join(wd, "pcg.bin"), "rb") as f:
            niter, maxit = np.fromfile(f, np.int32, 2)
            rerr = np.fromfile(f, np.float64, maxit)
            phi = np.fromfile(f, np.complex128, nni)
        out["pcg"] = {"niter": int(niter), "rerr": rerr, "phi": phi}
    if e2m is not None:
        out["e2m"] = np.fromfile(os.path.join(wd, "e2m.bin"), np.float64).reshape(
            (nx, ny, nze), order="F")
    _cleanup(wd, own)
    return out


def _cleanup(wd, own):
    if own:
        for fn in os.listdir(wd):
            os.remove(os.path.join(wd, fn))
        os.rmdir(wd)


def datafunc_available() -> bool:
    return os.path.isfile(DF_BIN) and os.access(DF_BIN, os.X_OK)


def run_datafunc(grid, origin, omega, dtype, x, e1, e2, d, r=None, az=None, exist=None,
                 timeout=120, binary=DF_BIN):
    """Runs the reference's dataResp (every site) and LmultT (all sites) for
    one transmitter and one data type (driver oracle/ref/ref_datafunc.f90).
    x, r: (nsite, 3); az: (nsite, 7) degrees; d, exist: (nsite, ncomp).
    Returns (resp (nsite, ncomp), comb1, comb2).  With binary=DF_DROPIN_BIN
    (oracle/ref/ref_dropin_datafunc.f90) it also returns the library's
    (resp, comb1, comb2, status) from the same Fortran host."""
    dx = np.asarray(grid["dx"], np.float64)
    dy = np.asarray(grid["dy"], np.float64)
    dz = np.asarray(grid["dz"], np.float64)
    nx, ny, nz = len(dx), len(dy), len(dz)
    nza = int(grid["nzAir"])
    x = np.asarray(x, np.float64).reshape(-1, 3)
    ns = x.shape[0]
    d = np.asarray(d, np.float64).reshape(ns, -1)
    nc = d.shape[1]
    r = np.zeros_like(x) if r is None else np.asarray(r, np.float64).reshape(ns, 3)
    if az is None:
        az = np.tile(np.array([0.0, 90.0, 0.0, 0.0, 90.0, 0.0, 90.0]), (ns, 1))
    az = np.asarray(az, np.float64).reshape(ns, 7)
    ex = np.ones((ns, nc), np.int32) if exist is None else np.asarray(exist, np.int32)
    wd = tempfile.mkdtemp(prefix="mtdf_")
    inp = os.path.join(wd, "input.bin")
    with open(inp, "wb") as f:
        np.array([nx, ny, nza, nz - nza], np.int32).tofile(f)
        dx.tofile(f); dy.tofile(f); dz.tofile(f)
        np.asarray(origin, np.float64).tofile(f)
        np.array([omega], np.float64).tofile(f)
        np.array([dtype, ns, nc], np.int32).tofile(f)
        x.tofile(f); r.tofile(f); az.tofile(f)     # row-major (nsite, k) = Fortran (k, nsite)
        d.tofile(f); ex.tofile(f)
        np.asarray(e1, np.complex128).tofile(f)
        np.asarray(e2, np.complex128).tofile(f)
    subprocess.run([binary, inp, wd], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    resp = np.fromfile(os.path.join(wd, "resp.bin"), np.float64).reshape(ns, nc)
    comb = np.fromfile(os.path.join(wd, "comb.bin"), np.complex128)
    ne = len(comb) // 2
    out = (resp, comb[:ne], comb[ne:])
    if binary == DF_DROPIN_BIN:
        lresp = np.fromfile(os.path.join(wd, "resp_lib.bin"), np.float64).reshape(ns, nc)
        lcomb = np.fromfile(os.path.join(wd, "comb_lib.bin"), np.complex128)
        st = np.fromfile(os.path.join(wd, "ierr.bin"), np.int32)
        out = out + (lresp, lcomb[:ne], lcomb[ne:], st)
    _cleanup(wd, True)
    return out


def bcx0_available() -> bool:
    return os.access(BCX0_BIN, os.X_OK)


def run_bcx0(grid, logcond, air_log, systems, timeout=300):
    """Runs the reference's ComputeBC -> BC_x0_WS (driver
    oracle/ref/ref_bcx0.f90) for systems = [(omega, imode), ...].  Returns
    (e0, bc): both (nsys, nEdge) complex, E0 and the boundary conditions as
    setBC writes them into a zero edge vector."""
    dx = np.asarray(grid["dx"], np.float64)
    dy = np.asarray(grid["dy"], np.float64)
    dz = np.asarray(grid["dz"], np.float64)
    nx, ny, nz = len(dx), len(dy), len(dz)
    nza = int(grid["nzAir"])
    lc = np.asarray(logcond, np.float64).reshape(nx, ny, nz - nza)
    wd = tempfile.mkdtemp(prefix="mtbc_")
    inp = os.path.join(wd, "input.bin")
    with open(inp, "wb") as f:
        np.array([nx, ny, nza, nz - nza], np.int32).tofile(f)
        dx.tofile(f); dy.tofile(f); dz.tofile(f)
        lc.ravel(order="F").tofile(f)
        np.array([air_log], np.float64).tofile(f)
        np.array([len(systems)], np.int32).tofile(f)
        for om, mode in systems:
            np.array([om], np.float64).tofile(f)
            np.array([mode], np.int32).tofile(f)
    subprocess.run([BCX0_BIN, inp, wd], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    ns = len(systems)
    e0 = np.fromfile(os.path.join(wd, "e0.bin"), np.complex128).reshape(ns, -1)
    bc = np.fromfile(os.path.join(wd, "bc.bin"), np.complex128).reshape(ns, -1)
    _cleanup(wd, True)
    return e0, bc


IO_BIN = os.path.join(HERE, "_ref", "ref_io")


def io_available() -> bool:
    return os.access(IO_BIN, os.X_OK)


def run_io_ws(path, air=None, timeout=120):
    """The reference's read_modelParam_WS of a model file (driver
    oracle/ref/ref_io.f90); air = (method, nz, max_height_km, min_top_dz_km,
    alpha) as a control file sets them, or None for the defaults.  Returns
    dict(dx, dy, dz, nzAir, logcond (nx, ny, nzEarth) LOGE, origin, rotdeg,
    air, param_type)."""
    wd = tempfile.mkdtemp(prefix="mtio_")
    out = os.path.join(wd, "ws.bin")
    extra = [] if air is None else [str(air[0])] + [repr(v) for v in air[1:]]
    subprocess.run([IO_BIN, "ws", os.path.abspath(path), out] + extra, check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, cwd=wd)
    b = open(out, "rb").read()
    nx, ny, nza, nze = (int(v) for v in np.frombuffer(b[:16], np.int32))
    o = [16]

    def take(n):
        a = np.frombuffer(b[o[0]:o[0] + 8 * n], np.float64).copy()
        o[0] += 8 * n
        return a
    dx, dy, dz = take(nx), take(ny), take(nza + nze)
    ox, oy, oz, rot = take(4)
    lc = take(nx * ny * nze).reshape(nze, ny, nx).transpose(2, 1, 0).copy()
    air = float(take(1)[0])
    pt = b[o[0]:o[0] + 80].decode().strip()
    _cleanup(wd, True)
    return dict(dx=dx, dy=dy, dz=dz, nzAir=nza, logcond=lc, origin=(ox, oy, oz), rotdeg=rot,
                air=air, param_type=pt)


def run_io_soln_write(path, dx, dy, dz, nzAir, omegas, fields, origin=(0.0, 0.0, 0.0), rotdeg=0.0,
                      timeout=120):
    """The reference's FileWriteInit + EfileWrite (write_solnVectorMTX's
    records) of fields (nPer, 2, nEdge) into path."""
    dx, dy, dz = (np.asarray(a, np.float64) for a in (dx, dy, dz))
    wd = tempfile.mkdtemp(prefix="mtio_")
    inp = os.path.join(wd, "in.bin")
    with open(inp, "wb") as f:
        np.array([len(dx), len(dy), nzAir, len(dz) - nzAir], np.int32).tofile(f)
        dx.tofile(f); dy.tofile(f); dz.tofile(f)
        np.array(list(origin) + [rotdeg], np.float64).tofile(f)
        np.array([len(omegas)], np.int32).tofile(f)
        for j, w in enumerate(omegas):
            np.array([w], np.float64).tofile(f)
            for k in range(2):
                np.asarray(fields[j][k], np.complex128).tofile(f)
    subprocess.run([IO_BIN, "solnw", inp, os.path.abspath(path)], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, cwd=wd)
    _cleanup(wd, True)


def run_io_soln_read(path, timeout=120):
    """The reference's FileReadInit + EfileRead of a solution file: dict(dx,
    dy, dz, nzAir, origin, rotdeg, omegas, fields (nPer, nMode, nEdge))."""
    wd = tempfile.mkdtemp(prefix="mtio_")
    out = os.path.join(wd, "soln.bin")
    subprocess.run([IO_BIN, "solnr", os.path.abspath(path), out], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, cwd=wd)
    b = open(out, "rb").read()
    nper, nmode, nx, ny, nz, nza = (int(v) for v in np.frombuffer(b[:24], np.int32))
    o = 24
    ox, oy, oz, rot = np.frombuffer(b[o:o + 32], np.float64)
    o += 32
    dx = np.frombuffer(b[o:o + 8 * nx], np.float64).copy(); o += 8 * nx
    dy = np.frombuffer(b[o:o + 8 * ny], np.float64).copy(); o += 8 * ny
    dz = np.frombuffer(b[o:o + 8 * nz], np.float64).copy(); o += 8 * nz
    ne = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    om = np.zeros(nper)
    fields = np.zeros((nper, nmode, ne), np.complex128)
    for j in range(nper):
        for k in range(nmode):
            om[j] = np.frombuffer(b[o:o + 8], np.float64)[0]
            o += 8
            fields[j, k] = np.frombuffer(b[o:o + 16 * ne], np.complex128)
            o += 16 * ne
    _cleanup(wd, True)
    return dict(dx=dx, dy=dy, dz=dz, nzAir=nza, origin=(ox, oy, oz), rotdeg=rot, omegas=om,
                fields=fields)


def run_io_zlist(path, timeout=120):
    """The reference's read_dataVectorMTX (read_Z_list) of a list data file:
    dict(periods, receivers [(id, x)], data [dict(tx, blocks [dict(type,
    ncomp, complex, rx, value, error, exist, azimuth)])])."""
    wd = tempfile.mkdtemp(prefix="mtio_")
    out = os.path.join(wd, "zl.bin")
    subprocess.run([IO_BIN, "zlist", os.path.abspath(path), out], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, cwd=wd)
    b = open(out, "rb").read()
    o = [0]

    def i32(n=1):
        a = np.frombuffer(b[o[0]:o[0] + 4 * n], np.int32).copy()
        o[0] += 4 * n
        return a

    def f64(n=1):
        a = np.frombuffer(b[o[0]:o[0] + 8 * n], np.float64).copy()
        o[0] += 8 * n
        return a
    ntx = int(i32()[0])
    periods = f64(ntx)
    nrx = int(i32()[0])
    receivers = []
    for _ in range(nrx):
        rid = b[o[0]:o[0] + 50].decode().rstrip()
        o[0] += 50
        receivers.append((rid, f64(3)))
    nd = int(i32()[0])
    data = []
    for _ in range(nd):
        tx, ndt = (int(v) for v in i32(2))
        blocks = []
        for _ in range(ndt):
            dt, nc, ns, cx = (int(v) for v in i32(4))
            rx = i32(ns)
            v = f64(nc * ns).reshape(ns, nc).T.copy()
            e = f64(nc * ns).reshape(ns, nc).T.copy()
            ex = i32(nc * ns).reshape(ns, nc).T.astype(bool)
            az = f64(6 * ns).reshape(ns, 6)
            blocks.append(dict(type=dt, ncomp=nc, complex=bool(cx), rx=rx, value=v, error=e, exist=ex,
                               azimuth=az))
        data.append(dict(tx=tx, blocks=blocks))
    _cleanup(wd, True)
    return dict(periods=periods, receivers=receivers, data=data)


def dropin_bc_available() -> bool:
    return os.access(BC_DROPIN_BIN, os.X_OK)


def run_dropin_bc(grid, logcond, air_log, systems, timeout=300):
    """One Fortran host (oracle/ref/ref_dropin_bc.f90) runs the reference's
    ComputeBC and the library's mtsp2_bc_x0 for systems = [(omega, mode)].
    Returns (ref, lib, status): (nsys, nEdge) fields in the form FWDsolve3D
    starts from, and the binding's status."""
    dx = np.asarray(grid["dx"], np.float64)
    dy = np.asarray(grid["dy"], np.float64)
    dz = np.asarray(grid["dz"], np.float64)
    nx, ny, nz = len(dx), len(dy), len(dz)
    nza = int(grid["nzAir"])
    lc = np.asarray(logcond, np.float64).reshape(nx, ny, nz - nza)
    wd = tempfile.mkdtemp(prefix="mtbc_")
    inp = os.path.join(wd, "input.bin")
    with open(inp, "wb") as f:
        np.array([nx, ny, nza, nz - nza], np.int32).tofile(f)
        dx.tofile(f); dy.tofile(f); dz.tofile(f)
        lc.ravel(order="F").tofile(f)
        np.array([air_log], np.float64).tofile(f)
        np.array([len(systems)], np.int32).tofile(f)
        for om, mode in systems:
            np.array([om], np.float64).tofile(f)
            np.array([mode], np.int32).tofile(f)
    subprocess.run([BC_DROPIN_BIN, inp, wd], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    ns = len(systems)
    ref = np.fromfile(os.path.join(wd, "ref.bin"), np.complex128).reshape(ns, -1)
    lib = np.fromfile(os.path.join(wd, "lib.bin"), np.complex128).reshape(ns, -1)
    st = int(np.fromfile(os.path.join(wd, "ierr.bin"), np.int32)[0])
    _cleanup(wd, True)
    return ref, lib, st


JMULT_BIN = os.path.join(HERE, "_ref", "ref_jmultT")


def jmultT_available() -> bool:
    return os.access(JMULT_BIN, os.X_OK)


def run_jmultT(grid, logcond, air_log, omega, sites, d, maxit=2000, tol=1e-10,
               origin=(0.0, 0.0, 0.0), timeout=600):
    """The reference's sensitivity chain for one MT transmitter (driver
    oracle/ref/ref_jmultT.f90): forward solves of both modes (BC_x0_WS
    boundary data, BICG to `tol`), LmultT of full-impedance data d (nsite, 8),
    the TRN solves (sensSolve) and PmultT, plus JmultT_TX itself.  Returns
    dict(fwd (2, nEdge), comb (2, nEdge), trn (2, nEdge), dsig, dsig_jt
    (nx, ny, nzEarth))."""
    dx = np.asarray(grid["dx"], np.float64)
    dy = np.asarray(grid["dy"], np.float64)
    dz = np.asarray(grid["dz"], np.float64)
    nx, ny, nz = len(dx), len(dy), len(dz)
    nza = int(grid["nzAir"])
    nze = nz - nza
    lc = np.asarray(logcond, np.float64).reshape(nx, ny, nze)
    x = np.ascontiguousarray(np.asarray(sites, np.float64).reshape(-1, 3))
    dd = np.ascontiguousarray(np.asarray(d, np.float64).reshape(len(x), 8))
    wd = tempfile.mkdtemp(prefix="mtjt_")
    inp = os.path.join(wd, "input.bin")
    with open(inp, "wb") as f:
        np.array([nx, ny, nza, nze], np.int32).tofile(f)
        dx.tofile(f); dy.tofile(f); dz.tofile(f)
        np.asarray(origin, np.float64).tofile(f)
        lc.ravel(order="F").tofile(f)
        np.array([air_log, omega], np.float64).tofile(f)
        np.array([maxit], np.int32).tofile(f)
        np.array([tol], np.float64).tofile(f)
        np.array([len(x)], np.int32).tofile(f)
        x.tofile(f)          # (nsite, 3) C order = x(3, nsite) Fortran order
        dd.tofile(f)         # d(8, nsite) Fortran order
    subprocess.run([JMULT_BIN, inp, wd], check=True, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, cwd=wd)
    out = {}
    for k in ("fwd", "comb", "trn"):
        out[k] = np.fromfile(os.path.join(wd, k + ".bin"), np.complex128).reshape(2, -1)
    for k in ("dsig", "dsig_jt"):
        out[k] = np.fromfile(os.path.join(wd, k + ".bin"), np.float64).reshape((nx, ny, nze),
                                                                              order="F")
    _cleanup(wd, True)
    return out
