"""CPU checks of the product's host side and C ABI (no GPU needed):
libmtsp2.so loads and exports every function include/mtsp2.h declares; the
library's own operator assembly and block-DILU pivots agree with the oracle;
argument errors come back as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import mtsp2
import sp2oracle
from conftest import ROOT, small_problem


def header_functions():
    txt = open(os.path.join(ROOT, "include", "mtsp2.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mt_\w+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = C.CDLL(mtsp2.LIB_PATH)
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(mtsp2.exported_symbols())
    mtsp2.lib()  # argtypes bind to every symbol


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_host_assembly_matches_oracle(seed):
    dx, dy, dz, nza, lc = small_problem(seed, nx=9, ny=8, nzAir=3, nzEarth=7)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    H = mtsp2.host_assemble(dx, dy, dz, nza, lc)
    n = o.nInt
    A = o.csr("AAii")
    Ao = sp.csr_matrix((A["val"], A["col"], A["row"]), shape=(n, n)).toarray()
    Ah = sp.csr_matrix((H["val"], H["col"], H["row"]), shape=(n, n)).toarray()
    scale = abs(Ao).max(axis=1, keepdims=True)
    assert (abs(Ao - Ah) / scale).max() <= 1e-14
    assert abs(o.vms() - H["vms1"]).max() <= 1e-14 * abs(o.vms()).max()
    for f in (1e-3, 1.0, 1e3):
        w = 2 * np.pi * f
        o.update_freq(w)
        piv = sp.csr_matrix((o.lu(1)["val"], o.lu(1)["col"], o.lu(1)["row"]), shape=(n, n)).diagonal()
        d1 = mtsp2.host_dilu_pivots((len(dx), len(dy), len(dz)), H["row"], H["col"], H["val"],
                                    H["vms1"], w)
        assert abs(1 / d1 - piv).max() <= 1e-13 * abs(piv).max()
        # the reference's own 1-based CSR through the drop-in conversion
        d2 = mtsp2.host_dilu_pivots((len(dx), len(dy), len(dz)), A["row"] + 1, A["col"] + 1,
                                    A["val"], o.vms() / w, w, base=1)
        assert abs(1 / d2 - piv).max() <= 1e-14 * abs(piv).max()


def test_non_stencil_csr_is_rejected():
    dx, dy, dz, nza, lc = small_problem(34, nx=6, ny=6, nzAir=2, nzEarth=4)
    o = sp2oracle.Problem(dx, dy, dz, nza, lc)
    A = o.csr("AAii")
    col = A["col"].copy()
    col[0] = o.nInt - 1  # row 0 coupled to the far corner: not a stencil neighbour
    with pytest.raises(mtsp2.MTError) as e:
        mtsp2.host_dilu_pivots((6, 6, 6), A["row"], col, A["val"], o.vms(), 1.0)
    assert e.value.code == mtsp2.MT_E_STENCIL


def test_bad_arguments_return_status():
    L = mtsp2.lib()
    ni, nz = C.c_int(), C.c_int()
    one = np.ones(1)
    rc = L.mt_host_assemble(1, 1, 0, 1, one, one, one, one, 0.0, C.byref(ni), C.byref(nz),
                            None, None, None, None)
    assert rc == mtsp2.MT_E_ARG
    h = C.c_void_p()
    assert L.mt_ctx_create(-1, C.byref(h)) == mtsp2.MT_E_DEVICE
    assert L.mt_ctx_create(999, C.byref(h)) == mtsp2.MT_E_DEVICE
    assert L.mt_bicg(None, 1.0, np.zeros(1, complex), np.zeros(1, complex), 1, 1e-7, 0,
                     None, np.zeros(1), None) == mtsp2.MT_E_ARG
