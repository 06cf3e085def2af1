"""The reference's file formats (mtsp2.formats): WS model files and binary
E-field solution files, against what the reference's own readers and
writers (oracle/ref/ref_io.f90) read and write -- committed fixtures, and
live where the reference build is present."""
import os

import numpy as np
import pytest

import refdriver
from mtsp2 import formats as F

GOLD = os.path.join(os.path.dirname(__file__), "golden", "io")
FIXED = dict(air_method="fixed height", nz_air=12, max_height=1.0e6, min_top_dz=30000.0, alpha=3.0)
REF_UT = "/root/reference/unit_testing/3D"


def _same(m, r, prefix=""):
    for k in ("dx", "dy", "dz", "logcond"):
        assert np.array_equal(m[k], r[prefix + k]), k
    assert tuple(m["origin"]) == tuple(np.asarray(r[prefix + "origin"]).tolist())
    assert m["rotdeg"] == float(r[prefix + "rotdeg"])
    assert m["nzAir"] == int(r[prefix + "nzAir"])
    assert m["param_type"] == str(r[prefix + "param_type"])


@pytest.mark.parametrize("name", ["ws_loge", "ws_log10", "ws_linear", "ws_quirks"])
def test_ws_reader_fixture(name):
    """Every value as the reference's read_modelParam_WS reads it, with the
    default (mirrored) and the fixed-height air layers; ws_quirks uses
    list-directed freedoms (repeats, commas, D exponents, split and surplus
    values, no origin line) and LOG10 values (converted with log(10.) in
    single precision, as setType_modelParam does)."""
    d = np.load(os.path.join(GOLD, "ref_io.npz"))
    path = os.path.join(GOLD, name + ".rho")
    _same(F.read_ws_model(path), d, f"{name}_default_")
    _same(F.read_ws_model(path, **FIXED), d, f"{name}_fixed_")


def test_ws_writer_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    dx, dy = rng.uniform(100, 900, 6), rng.uniform(100, 900, 5)
    dze = 30.0 * 1.2 ** np.arange(11)
    lc = np.log(10.0 ** rng.uniform(-3.5, 0.5, (6, 5, 11)))
    for pt in ("LOGE", "LOG10", "LINEAR"):
        p = str(tmp_path / f"m_{pt}.rho")
        F.write_ws_model(p, dx, dy, dze, lc, pt, origin=(-10.5, 20.25, 0.0), rotdeg=3.0)
        m = F.read_ws_model(p)
        assert m["param_type"] == pt
        assert np.allclose(m["dx"], dx, rtol=0, atol=5e-4)             # f12.3
        assert np.allclose(m["logcond"], lc, rtol=2e-5, atol=1e-5)      # es13.5
        assert m["origin"] == (-10.5, 20.25, 0.0) and m["rotdeg"] == 3.0
        if refdriver.io_available():
            _same(m, refdriver.run_io_ws(p))                              # the reference reads our file alike


def test_air_layers_mirror_minimum():
    dz = np.full(10, 50.0)
    a = F.air_layers(dz, "mirror", 10, 1.5, 30000.0)
    assert a[0] == 30000.0                                             # 1.5^9 * 50 m raised to MinTopDz
    assert np.array_equal(a[1:][::-1], 50.0 * 1.5 ** np.arange(9))     # alpha^(j-1) dz(j), exact here
    b = F.air_layers(dz, "mirror", 10, 3.0, 30000.0)
    assert b[0] == 3.0 ** 9 * 50.0                                     # already above it
    with pytest.raises(ValueError):
        F.air_layers(dz[:5], "mirror", 10)


def test_soln_fixture(tmp_path):
    """A solution file written by the reference's own writer: read back bit
    for bit, and the library's writer produces the same bytes."""
    d = np.load(os.path.join(GOLD, "ref_io.npz"))
    path = os.path.join(GOLD, "ref_small.soln")
    s = F.read_soln(path)
    assert np.array_equal(s["fields"], d["soln_fields"])
    assert np.array_equal(s["omegas"], d["soln_omegas"])
    assert np.array_equal(s["dx"], d["soln_dx"]) and np.array_equal(s["dz"], d["soln_dz"])
    assert s["modes"] == ["Ex", "Ey"] and s["nzAir"] == 2
    assert s["origin"] == (-600.0, -450.0, 0.0) and s["rotdeg"] == 12.5
    out = str(tmp_path / "ours.soln")
    F.write_soln(out, d["soln_dx"], d["soln_dy"], d["soln_dz"], 2, d["soln_omegas"], d["soln_fields"],
                 origin=(-600.0, -450.0, 0.0), rotdeg=12.5)
    assert open(out, "rb").read() == open(path, "rb").read()


@pytest.mark.skipif(not refdriver.io_available(), reason="reference build (oracle/_ref/ref_io) absent")
def test_soln_live(tmp_path):
    rng = np.random.default_rng(9)
    dx, dy = rng.uniform(100, 500, 6), rng.uniform(100, 500, 4)
    dz = np.concatenate([[5000.0, 900.0, 150.0], rng.uniform(40, 300, 5)])
    ne = sum(F.edge_counts(6, 4, 8))
    om = 2 * np.pi * np.array([1e-3, 0.5, 30.0])
    fl = rng.standard_normal((3, 2, ne)) + 1j * rng.standard_normal((3, 2, ne))
    ours, ref = str(tmp_path / "o.soln"), str(tmp_path / "r.soln")
    F.write_soln(ours, dx, dy, dz, 3, om, fl)
    refdriver.run_io_soln_write(ref, dx, dy, dz, 3, om, fl)
    assert open(ours, "rb").read() == open(ref, "rb").read()
    r = refdriver.run_io_soln_read(ours)
    assert np.array_equal(r["fields"], fl) and np.array_equal(r["omegas"], om)


@pytest.mark.skipif(not (refdriver.io_available() and os.path.isdir(REF_UT)),
                    reason="reference tree or build absent")
@pytest.mark.parametrize("name", ["CylinderModel+QuarterSpace.rho", "Halfspace.rho"])
def test_ws_reference_models_live(name):
    """The reference's own test models (48 x 68 x 44) read alike."""
    path = os.path.join(REF_UT, name)
    _same(F.read_ws_model(path), refdriver.run_io_ws(path))


def _same_list(m, r):
    assert np.array_equal(m["periods"], r["periods"])
    assert [a for a, _ in m["receivers"]] == [a for a, _ in r["receivers"]]
    assert all(np.array_equal(xa, xb) for (_, xa), (_, xb) in zip(m["receivers"], r["receivers"]))
    assert len(m["data"]) == len(r["data"])
    for dm, dr in zip(m["data"], r["data"]):
        assert dm["tx"] == dr["tx"] and len(dm["blocks"]) == len(dr["blocks"])
        for bm, br in zip(dm["blocks"], dr["blocks"]):
            assert bm["type"] == br["type"] and bm["ncomp"] == br["ncomp"]
            for k in ("rx", "value", "error", "exist", "azimuth"):
                assert np.array_equal(np.asarray(bm[k]), np.asarray(br[k])), (bm["type"], k)


def _unpack_list(d, p):
    z = dict(periods=d[p + "periods"],
             receivers=list(zip([str(s) for s in d[p + "rx_ids"]], d[p + "rx_x"])), data=[])
    for i in range(int(d[p + "nd"])):
        blocks = [{k: d[f"{p}d{i}_b{j}_{k}"] for k in
                   ("type", "ncomp", "rx", "value", "error", "exist", "azimuth")}
                  for j in range(int(d[f"{p}d{i}_nb"]))]
        for b in blocks:
            b["type"], b["ncomp"] = int(b["type"]), int(b["ncomp"])
        z["data"].append(dict(tx=int(d[f"{p}d{i}_tx"]), blocks=blocks))
    return z


def test_list_reader_fixture():
    """read_Z_list's MT branches: impedances with 11-15 columns (azimuths,
    the geographic orientation added), rho/phase (log10 of rho with its error
    propagation, phase to radians with the sign convention), phase tensor,
    vertical TFs, Ohm and [mV/km]/[nT] units (ImpUnits' default-real
    literals), conjugation of exp(+i omega t) data, transmitters and sites
    shared across blocks (and the reference's two site-id formats)."""
    d = np.load(os.path.join(GOLD, "ref_io.npz"))
    _same_list(F.read_data_list(os.path.join(GOLD, "zlist_mixed.dat")), _unpack_list(d, "zl_"))


@pytest.mark.skipif(not refdriver.io_available(), reason="reference build (oracle/_ref/ref_io) absent")
def test_list_reader_live():
    files = [os.path.join(GOLD, "zlist_mixed.dat")]
    if os.path.isdir(REF_UT):
        files += [os.path.join(REF_UT, f) for f in
                  ("CylinderModel_1per.dat", "CylinderModel_1per_Template.dat")]
    for p in files:
        _same_list(F.read_data_list(p), refdriver.run_io_zlist(p))
