"""The reference's file formats, for callers of the library that have no
ModEM host: WS model files (3D_MT/modelParam/modelParamIO/WS.inc) with the
air layers the reader adds (GridDef.f90:406-488), the binary E-field
solution files (3D_MT/ioMod/ioAscii.f90: FileWriteInit, EfileWrite,
FileReadInit, EfileRead, as write_solnVectorMTX / read_solnVectorMTX use
them), and list data files of MT transmitters (read_Z_list,
3D_MT/DataIO.f90:552-1209).  Values read are bit-identical to what the
reference's readers produce from the same file; solution files written are
byte-identical to the reference's."""
from __future__ import annotations

import math
import re
import struct

import numpy as np

# log(10.) with a default-real 10. is single precision; setType_modelParam
# (ModelSpace.f90:276-279) multiplies LOG10 values by it
LN10_SINGLE = float(np.float32(np.log(np.float32(10.0))))

# ------------------------------------------------------------- list-directed


class _Records:
    """Fortran list-directed input over text records: each read starts at a
    new record, takes values from as many records as it needs and discards
    the rest of the last one; blanks and commas separate values, r*c repeats
    c, a slash ends the read, D exponents are accepted."""

    def __init__(self, lines):
        self.lines = lines
        self.pos = 0

    def next_line(self):
        if self.pos >= len(self.lines):
            raise EOFError("end of file")
        ln = self.lines[self.pos]
        self.pos += 1
        return ln

    @staticmethod
    def tokens(line):
        out = []
        for t in re.split(r"[,\s]+", line.strip()):
            if not t:
                continue
            if t.startswith("/"):
                out.append("/")
                break
            if "*" in t:
                r, c = t.split("*", 1)
                out.extend([c] * int(r))
            else:
                out.append(t)
        return out

    def values(self, n, conv=float):
        vals = []
        while len(vals) < n:
            toks = self.tokens(self.next_line())
            for t in toks:
                if t == "/":
                    return vals
                vals.append(conv(t.replace("D", "E").replace("d", "e")) if conv is float else conv(t))
                if len(vals) == n:
                    break
        return vals


def _powi(a, b):
    """real**integer as compiled (repeated squaring)."""
    recip = b < 0
    b = abs(b)
    r = 1.0
    while True:
        if b & 1:
            r *= a
        b //= 2
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def air_layers(dz_earth, method="mirror", nz=10, alpha=3.0, min_top_dz=30000.0,
               max_height=1.0e6):
    """setup_airlayers (GridDef.f90:406-488): thicknesses of the nz air layers,
    top first.  'mirror': the earth's top layers grown by alpha upwards, the
    top one at least min_top_dz (m); 'fixed height': log-spaced up to
    max_height (m)."""
    dz_earth = np.asarray(dz_earth, np.float64)
    dz = np.zeros(nz)
    if method == "mirror":
        if len(dz_earth) < nz:
            raise ValueError(f"mirrored air layers need at least {nz} earth layers "
                             f"(the reference reads past its array otherwise)")
        for iz in range(nz, 0, -1):
            j = nz - iz + 1
            dz[iz - 1] = _powi(float(alpha), j - 1) * float(dz_earth[j - 1])
        if dz[0] < min_top_dz:
            dz[0] = min_top_dz
    elif method == "fixed height":
        z1 = math.log10(float(dz_earth[0]))
        dl = (math.log10(max_height) - z1) / nz
        z = z1
        for iz in range(nz, 0, -1):
            dz[iz - 1] = math.pow(10.0, z + dl) - math.pow(10.0, z)
            z = z + dl
    else:
        raise ValueError(f"unknown air-layer method {method!r}")
    return dz


def _fsum(a):
    s = 0.0
    for v in a:
        s += float(v)
    return s


def read_ws_model(path, air_method="mirror", nz_air=10, alpha=3.0, min_top_dz=30000.0,
                  max_height=1.0e6):
    """read_modelParam_WS: returns dict(dx, dy, dz (air + earth), nzAir,
    logcond (nx, ny, nzEarth) as LOGE conductivity, origin, rotdeg,
    param_type as the file declares it)."""
    with open(path, "r") as f:
        lines = f.read().split("\n")
    R = _Records(lines)
    R.next_line()                                    # comment
    head = R.next_line()[:80]
    nums = _Records([head]).values(4, int)
    nx, ny, nze, mapping = nums
    if mapping != 0:
        raise ValueError("WS model files with a mapping index are not supported (as in the reference)")
    ptype = "LOGE" if "LOGE" in head else ("LOG10" if "LOG10" in head else "LINEAR")
    dx = np.array(R.values(nx))
    dy = np.array(R.values(ny))
    dze = np.array(R.values(nze))
    rho = np.zeros((nx, ny, nze))
    try:
        for k in range(nze):
            for j in range(ny):
                vals = R.values(nx)
                rho[::-1, j, k][:len(vals)] = vals        # i = nx .. 1
    except EOFError:
        pass                                          # the reference reads with iostat: rest stays 0
    if ptype in ("LOGE", "LOG10"):
        cc = -rho
    else:
        cc = 1.0 / rho
    try:
        o = R.values(3)
        if len(o) < 3:
            raise ValueError
        origin = (o[0], o[1], o[2])
    except (EOFError, ValueError):
        origin = (-_fsum(dx) / 2.0, -_fsum(dy) / 2.0, 0.0)
    try:
        r = R.values(1)
        rot = r[0] if r else 0.0
    except (EOFError, ValueError):
        rot = 0.0
    if ptype == "LOG10":
        lc = cc * LN10_SINGLE
    elif ptype == "LINEAR":
        lc = np.vectorize(math.log, otypes=[np.float64])(cc)
    else:
        lc = cc
    dza = air_layers(dze, air_method, nz_air, alpha, min_top_dz, max_height)
    return dict(dx=dx, dy=dy, dz=np.concatenate([dza, dze]), nzAir=nz_air, logcond=lc,
                origin=origin, rotdeg=rot, param_type=ptype)


def write_ws_model(path, dx, dy, dz_earth, logcond, param_type="LOGE", origin=None, rotdeg=0.0,
                   comment="3D MT model written by ModEM in WS format"):
    """write_modelParam_WS: resistivities as -LOGE/-LOG10 values or ohm-m
    (es13.5, i from nx down to 1), cell sizes f12.3, origin 3f16.3, rotation
    f9.3.  logcond (nx, ny, nzEarth) is LOGE conductivity."""
    dx, dy, dze = (np.asarray(a, np.float64) for a in (dx, dy, dz_earth))
    nx, ny, nze = len(dx), len(dy), len(dze)
    lc = np.asarray(logcond, np.float64).reshape(nx, ny, nze)
    if param_type == "LOGE":
        rho = -lc
    elif param_type == "LOG10":
        rho = -(lc / LN10_SINGLE)
    elif param_type == "LINEAR":
        rho = 1.0 / np.exp(lc)
    else:
        raise ValueError(f"unknown parameter type {param_type!r}")
    if origin is None:
        origin = (-_fsum(dx) / 2.0, -_fsum(dy) / 2.0, 0.0)
    out = [f" # {comment}", f"{nx:5d}{ny:5d}{nze:5d}{0:5d} {param_type}",
           "".join(f"{v:12.3f}" for v in dx), "".join(f"{v:12.3f}" for v in dy),
           "".join(f"{v:12.3f}" for v in dze), ""]
    for k in range(nze):
        for j in range(ny):
            out.append("".join(f"{rho[i, j, k]:13.5E}" for i in range(nx - 1, -1, -1)))
        out.append("")
    out.append("".join(f"{v:16.3f}" for v in origin))
    out.append(f"{rotdeg:9.3f}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# ------------------------------------------------------------- solution files

MODE_NAMES = ("Ex", "Ey")     # Pol_name of an MT solution (SolnSpace.f90:200-204)


def _rec(f, payload):
    n = len(payload)
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def _read_rec(f):
    h = f.read(4)
    if len(h) < 4:
        raise EOFError("end of solution file")
    (n,) = struct.unpack("<i", h)
    data = f.read(n)
    (m,) = struct.unpack("<i", f.read(4))
    if m != n or len(data) != n:
        raise ValueError("corrupt record markers")
    return data


def edge_counts(nx, ny, nz):
    return nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz


def write_soln(path, dx, dy, dz, nzAir, omegas, fields, origin=(0.0, 0.0, 0.0), rotdeg=0.0,
               version=""):
    """write_solnVectorMTX: header (version, nPer, nMode = 2, nx, ny, nz,
    nzAir, origin, rotation), dx, dy, dz, then per period and mode the record
    (omega, iFreq, iMode, mode name) and the x-, y- and z-edge fields.
    fields: (nPer, 2, nEdge) complex, all-edge order."""
    dx, dy, dz = (np.ascontiguousarray(a, np.float64) for a in (dx, dy, dz))
    nx, ny, nz = len(dx), len(dy), len(dz)
    nxe, nye, nze = edge_counts(nx, ny, nz)
    fields = np.asarray(fields, np.complex128)
    nper = len(omegas)
    if fields.shape != (nper, 2, nxe + nye + nze):
        raise ValueError(f"fields must be (nPer, 2, {nxe + nye + nze})")
    with open(path, "wb") as f:
        _rec(f, version.ljust(20)[:20].encode() +
             struct.pack("<6i4d", nper, 2, nx, ny, nz, int(nzAir), *map(float, origin), float(rotdeg)))
        for a in (dx, dy, dz):
            _rec(f, a.tobytes())
        for j in range(nper):
            for k in range(2):
                _rec(f, struct.pack("<d2i", float(omegas[j]), j + 1, k + 1) +
                     MODE_NAMES[k].ljust(20).encode())
                e = fields[j, k]
                for part in (e[:nxe], e[nxe:nxe + nye], e[nxe + nye:]):
                    _rec(f, np.ascontiguousarray(part).tobytes())


def read_soln(path):
    """read_solnVectorMTX: returns dict(version, dx, dy, dz, nzAir, origin,
    rotdeg, omegas (nPer,), modes (names), fields (nPer, nMode, nEdge))."""
    with open(path, "rb") as f:
        h = _read_rec(f)
        version = h[:20].decode(errors="replace").rstrip()
        nper, nmode, nx, ny, nz, nza = struct.unpack("<6i", h[20:44])
        ox, oy, oz, rot = struct.unpack("<4d", h[44:76])
        dx = np.frombuffer(_read_rec(f), np.float64).copy()
        dy = np.frombuffer(_read_rec(f), np.float64).copy()
        dz = np.frombuffer(_read_rec(f), np.float64).copy()
        ne = sum(edge_counts(nx, ny, nz))
        om = np.zeros(nper)
        names = []
        fields = np.zeros((nper, nmode, ne), np.complex128)
        for j in range(nper):
            for k in range(nmode):
                r = _read_rec(f)
                om[j] = struct.unpack("<d", r[:8])[0]
                if j == 0:
                    names.append(r[16:36].decode(errors="replace").rstrip())
                fields[j, k] = np.concatenate(
                    [np.frombuffer(_read_rec(f), np.complex128) for _ in range(3)])
    return dict(version=version, dx=dx, dy=dy, dz=dz, nzAir=nza, origin=(ox, oy, oz), rotdeg=rot,
                omegas=om, modes=names, fields=fields)


# ------------------------------------------------------------- list data files
#
# read_Z_list (3D_MT/DataIO.f90:552-1209) for MT transmitters.  The data
# types follow dataTypes.f90 (numbering, units, component ids).

_PI = 3.14159265357898                       # math_constants.f90:12
_D2R = _PI / 180.0
_LARGE_REAL = float(np.float32(1.0e13))      # a default-real literal widened
_TOL6 = 0.000001
_ISIGN = -1

DATA_TYPES = {                               # name: (id, complex, units, component ids)
    "Full_Impedance": (1, True, "[V/m]/[T]", ("ZXX", "ZXY", "ZYX", "ZYY")),
    "Off_Diagonal_Impedance": (2, True, "[V/m]/[T]", ("ZXY", "ZYX")),
    "Full_Vertical_Components": (3, True, "[]", ("TX ", "TY ")),
    "Off_Diagonal_Rho_Phase": (5, False, "[]", ("RHOXY", "PHSXY", "RHOYX", "PHSYX")),
    "Phase_Tensor": (6, False, "[]", ("PTXX", "PTXY", "PTYX", "PTYY")),
    "Ex_Field": (7, True, "[V/m]", ("Ex_Field",)),
    "Ey_Field": (8, True, "[V/m]", ("Ey_Field",)),
    "Bx_Field": (9, True, "[T]", ("Bx_Field",)),
    "By_Field": (10, True, "[T]", ("By_Field",)),
    "Bz_Field": (11, True, "[T]", ("Bz_Field",)),
}


def _f32(x):
    return float(np.float32(x))


def _units_factor(old, new):
    """ImpUnits (dataTypes.f90:252-315), with its default-real literals."""
    if "[]" in old or "[]" in new:
        return 1.0
    four_pi = 4 * _PI
    if "[V/m]/[T]" in old:
        f1 = 1.0
    elif "[mV/km]/[nT]" in old:
        f1 = 1.0 * 1000.0
    elif "[V/m]/[A/m]" in old or "Ohm" in old:
        f1 = 1.0 * 1000.0 * 10000.0 / four_pi
    elif "[V/m]" in old or "[T]" in old:
        f1 = 1.0
    elif "[nT]" in old:
        f1 = 1.0 * _f32(1.0e-9)
    elif "[A/m]" in old:
        f1 = 1.0 * (four_pi * _f32(1.0e-7))
    else:
        raise ValueError(f"unknown input units {old.strip()!r}")
    if "[V/m]/[T]" in new:
        f2 = 1.0
    elif "[mV/km]/[nT]" in new:
        f2 = 1.0 / 1000.0
    elif "[V/m]/[A/m]" in new or "Ohm" in new:
        f2 = 1.0 / (1000.0 * 10000.0 / four_pi)
    elif "[V/m]" in new or "[T]" in new:
        f2 = 1.0
    elif "[nT]" in new:
        f2 = 1.0 * 1.0e9
    elif "[A/m]" in new:
        f2 = 1.0 / (four_pi * _f32(1.0e-7))
    else:
        raise ValueError(f"unknown output units {new.strip()!r}")
    return f1 * f2


def _compact(s):
    """compact (utilities.f90:431-469): blanks and tabs collapsed, control
    characters dropped, left-adjusted."""
    out, sp = [], False
    for ch in s.lstrip(" \t"):
        o = ord(ch)
        if o in (9, 32):
            if not sp:
                out.append(" ")
            sp = True
        elif o >= 33:
            out.append(ch)
            sp = False
    return "".join(out).strip()


def _fmt_f93(v):
    t = f"{v:9.3f}"
    return "*" * 9 if len(t) > 9 else t


def _site_id(code, lat, lon, width, with_latlon=True):
    """write(siteid, '(aW,2f9.3)') code[, lat, lon]: code as a character(40)."""
    c = code[:40].ljust(40)[:width]
    return c + (_fmt_f93(lat) + _fmt_f93(lon) if with_latlon else "")


def read_data_list(path):
    """read_Z_list for MT transmitters: returns dict(periods (transmitter
    dictionary order), receivers [(id, x(3))] (receiver dictionary order),
    data [dict(tx, blocks [dict(type, ncomp, rx, value (ncomp, nsite), error,
    exist, azimuth (nsite, 6: Hx, Hy, Ex, Ey, Hx_ref, Hy_ref))])], info per
    data type).  tx and rx are 1-based dictionary indices, as the
    reference's data vectors hold them."""
    with open(path, "r") as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    pos = [0]

    def nxt():
        if pos[0] >= len(lines):
            raise EOFError
        ln = lines[pos[0]]
        pos[0] += 1
        return ln

    periods, receivers, rxid = [], [], {}
    alld, info = [], {}
    while True:
        try:
            nxt()                                    # '# ' + info
            nxt()                                    # '# ' + header
            ln = nxt()
            if ln[:1] == "+":
                txtype = _compact(ln[2:102])
                ln = nxt()
            else:
                txtype = "MT"
        except EOFError:
            break
        if txtype != "MT":
            raise NotImplementedError(f"transmitter type {txtype!r}: only MT list data are read")
        tname = _compact(ln[2:102])
        if tname not in DATA_TYPES:
            raise ValueError(f"unknown or unsupported data type {tname!r}")
        idt, cplx, tunits, ids = DATA_TYPES[tname]
        nc = len(ids)
        sign_info = nxt()[2:22]
        conjugate = (-1 if "-" in sign_info else 1) != _ISIGN
        units = nxt()[2:22]
        si = _units_factor(units, tunits)
        orient = float(_Records.tokens(nxt())[1])
        t = _Records.tokens(nxt())
        origin = (float(t[1]), float(t[2]))
        t = _Records.tokens(nxt())
        ntx, nrx = int(t[1]), int(t[2])
        info[tname] = dict(sign=sign_info.strip(), units=units.strip(), orientation=orient,
                           origin=origin, conjugate=conjugate, si_factor=si)
        new_tx, new_rx = [0] * ntx, [0] * nrx
        value = np.zeros((ntx, nrx, nc), np.complex128)
        error = np.full((ntx, nrx, nc), _LARGE_REAL)
        exist = np.zeros((ntx, nrx, nc), bool)
        az = np.zeros((ntx, nrx, 6))
        while pos[0] < len(lines):
            raw = lines[pos[0]]
            if raw[:1] == "#":
                break
            toks = _Records.tokens(raw)
            nt = len(_compact(raw).split(" ")) if _compact(raw) else 0
            real_type = not cplx
            lo, hi = (10, 14) if real_type and idt in (5, 6) else (11, 15)
            if idt >= 7:
                lo, hi = 11, 11
            if nt < lo or nt > hi:
                break
            pos[0] += 1
            per = float(toks[0])
            code = toks[1]
            lat, lon = float(toks[2]), float(toks[3])
            x = [float(toks[4]), float(toks[5]), float(toks[6])]
            comp = toks[7][:50]
            if real_type:
                zr, zi, ze = float(toks[8]), 0.0, float(toks[9])
                angles = [float(v) for v in toks[10:nt]]
            else:
                zr, zi, ze = float(toks[8]), float(toks[9]), float(toks[10])
                angles = [float(v) for v in toks[11:nt]]
            # channel azimuths (the file's geographic orientation added)
            hx_ref, hy_ref = 0.0, 0.0 + 90.0
            if idt >= 7:
                hx = ex = hy = ey = 0.0
                hx_ref = hy_ref = 0.0
            elif not angles:
                hx = orient
                ex = orient
                hy, ey = hx + 90.0, ex + 90.0
            else:
                hx = angles[0] + orient
                hy = (angles[1] + orient) if len(angles) >= 2 else hx + 90.0
                ex = (angles[2] + orient) if len(angles) >= 3 else hx
                ey = (angles[3] + orient) if len(angles) >= 4 else ex + 90.0
            icomp = next((i for i, cid in enumerate(ids) if comp.strip() in cid.strip()), None)
            if icomp is None:
                raise ValueError(f"component {comp!r} not in {tname}")
            if idt in (5, 6):
                if "RHO" in comp:
                    ze = ze / zr / math.log(10.0)
                    zr = math.log10(zr)
                if "PHS" in comp:
                    zr = zr * _D2R if conjugate else -zr * _D2R
                    ze = ze * _D2R
            itx = next((i for i, p in enumerate(periods) if abs(p - per) < _TOL6), None)
            if itx is None:
                periods.append(per)
                itx = len(periods) - 1
            itx += 1
            if idt in (1, 2, 3):
                sid = _site_id(code, lat, lon, 20)
            elif idt in (5, 6):
                sid = _site_id(code, lat, lon, 22)
            else:
                sid = _site_id(code, lat, lon, 22, with_latlon=False)
            key = sid.rstrip()
            if key not in rxid:
                receivers.append((key, np.array(x)))
                rxid[key] = len(receivers)
            irx = rxid[key]
            i = next((q for q in range(ntx) if new_tx[q] in (itx, 0)), None)
            j = next((q for q in range(nrx) if new_rx[q] in (irx, 0)), None)
            if i is None or j is None:
                raise ValueError("more transmitters or receivers than the block header declares")
            new_tx[i], new_rx[j] = itx, irx
            if cplx:
                value[i, j, icomp] = complex(si * zr, si * (-zi if conjugate else zi))
            else:
                value[i, j, icomp] = si * zr
            error[i, j, icomp] = si * ze
            exist[i, j, icomp] = True
            az[i, j] = (hx, hy, ex, ey, hx_ref, hy_ref)
        # one block per transmitter of this data type, merged into the data
        ncomp = 2 * nc if cplx else nc
        for i in range(ntx):
            js = [j for j in range(nrx) if exist[i, j].any()]
            v = np.zeros((ncomp, len(js)))
            e = np.zeros((ncomp, len(js)))
            ex_ = np.zeros((ncomp, len(js)), bool)
            for k, j in enumerate(js):
                if cplx:
                    v[0::2, k], v[1::2, k] = value[i, j].real, value[i, j].imag
                    e[0::2, k] = e[1::2, k] = error[i, j]
                    ex_[0::2, k] = ex_[1::2, k] = exist[i, j]
                else:
                    v[:, k], e[:, k], ex_[:, k] = value[i, j].real, error[i, j], exist[i, j]
            blk = dict(type=idt, ncomp=ncomp, complex=cplx,
                       rx=np.array([new_rx[j] for j in js], np.int32), value=v, error=e, exist=ex_,
                       azimuth=np.array([az[i, j] for j in js]).reshape(len(js), 6))
            hit = next((d for d in alld if d["tx"] == new_tx[i]), None)
            if hit is None:
                alld.append(dict(tx=new_tx[i], blocks=[blk]))
            elif any(b["type"] == idt for b in hit["blocks"]):
                raise NotImplementedError("a data type given twice for one transmitter")
            else:
                hit["blocks"].append(blk)
    return dict(periods=np.array(periods), receivers=receivers, data=alld, info=info)
