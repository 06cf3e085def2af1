"""The file formats tied to the GPU path: a WS model file -> the library's
reader (checked against the reference's read_modelParam_WS) -> device
assembly, boundary fields (BC_x0_WS), right-hand sides and mixed solves ->
the all-edge fields -> a binary solution file, which must be the bytes the
reference's own writer (FileWriteInit + EfileWrite, through
oracle/_ref/ref_io) produces for the same fields, and must read back through
the reference's reader unchanged.  The GPU fields are checked against the
oracle's FP64 solves of the same systems (E rel L2 <= 1e-7 at tol 1e-10).
The reference build's binaries travel with the tree; without them the
reference-writer comparisons are skipped."""
import os

import numpy as np
import pytest

import mtsp2
import refdriver
import sp2oracle
from conftest import rel
from mtsp2 import formats as F

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "io")


def _model_file(tmp_path):
    """A WS model written in LOG10 (the reader's single-precision log(10.)
    conversion is exercised), blocky, with a conductive block."""
    rng = np.random.default_rng(12)
    nx, ny, nze = 22, 19, 14
    dx = np.round(rng.uniform(180.0, 420.0, nx), 3)
    dy = np.round(rng.uniform(180.0, 420.0, ny), 3)
    dze = np.round(40.0 * 1.25 ** np.arange(nze), 3)
    lc = np.full((nx, ny, nze), np.log(0.01))
    lc[6:14, 5:12, 4:9] = np.log(0.5)
    lc[15:, :, 9:] = np.log(0.002)
    path = str(tmp_path / "model.rho")
    F.write_ws_model(path, dx, dy, dze, lc, "LOG10", origin=(-3000.0, -2500.0, 0.0))
    return path


@pytest.mark.parametrize("source", ["generated", "fixture"])
def test_model_file_to_gpu_solve_to_solution_file(gpu_ctx, tmp_path, source):
    print(f"case io path source={source}", flush=True)
    path = _model_file(tmp_path) if source == "generated" else os.path.join(GOLD, "ws_loge.rho")
    m = F.read_ws_model(path)
    if refdriver.io_available():
        r = refdriver.run_io_ws(path)   # the reference's reader sees the same model
        for k in ("dx", "dy", "dz", "logcond"):
            assert np.array_equal(m[k], r[k]), k
        assert m["nzAir"] == int(r["nzAir"])
    dx, dy, dz, nza, lc = m["dx"], m["dy"], m["dz"], m["nzAir"], m["logcond"]
    gpu_ctx.assemble(dx, dy, dz, nza, lc)
    n, ne = gpu_ctx.n_int, gpu_ctx.n_edge
    fhz = np.array([0.05, 8.0])
    om = np.repeat(2 * np.pi * fhz, 2)
    modes = np.tile(np.array([1, 2], np.int32), len(fhz))
    ns = len(om)
    d_e = gpu_ctx.dev_alloc(ns * ne * 16)
    d_b, d_x = gpu_ctx.dev_alloc(ns * n * 16), gpu_ctx.dev_alloc(ns * n * 16)
    gpu_ctx.bc_x0_dev(om, modes, d_e)
    gpu_ctx.rhs_fwd_dev(ns, d_e, d_b, d_x)
    st = gpu_ctx.solve_batch_dev(om, d_b, d_x, tol=1e-10, maxit=4000, precision=mtsp2.PREC_MIXED,
                                 precond=mtsp2.PC_DILU_RB)
    assert all(s["status"] == 0 for s in st), st
    gpu_ctx.post_fwd_dev(ns, d_x, d_e)
    E = gpu_ctx.download(d_e, np.zeros((ns, ne), np.complex128))
    B = gpu_ctx.download(d_b, np.zeros((ns, n), np.complex128))
    # the GPU fields against the oracle's FP64 solves of the same systems
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    ei = o.index()[0]
    for s in range(ns):
        o.update_freq(om[s])
        xo, info = o.bicg(B[s], E[s][ei], maxit=4000, tol=1e-10, gpu_semantics=True)
        assert rel(E[s][ei], xo) <= 1e-7, s
    fields = E.reshape(len(fhz), 2, ne)
    ours = str(tmp_path / "gpu.soln")
    F.write_soln(ours, dx, dy, dz, nza, 2 * np.pi * fhz, fields, origin=m["origin"],
                 rotdeg=m["rotdeg"])
    back = F.read_soln(ours)
    assert np.array_equal(back["fields"], fields)
    if not refdriver.io_available():
        pytest.skip("reference build (oracle/_ref/ref_io) absent: writer comparison skipped")
    ref = str(tmp_path / "ref.soln")
    refdriver.run_io_soln_write(ref, dx, dy, dz, nza, 2 * np.pi * fhz, fields, origin=m["origin"],
                                rotdeg=m["rotdeg"])
    assert open(ours, "rb").read() == open(ref, "rb").read()
    r = refdriver.run_io_soln_read(ours)
    assert np.array_equal(r["fields"], fields)
    assert np.array_equal(r["omegas"], 2 * np.pi * fhz)
