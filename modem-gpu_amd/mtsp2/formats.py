"""The reference's file formats, for callers of the library that have no
ModEM host: WS model files (3D_MT/modelParam/modelParamIO/WS.inc) with the
air layers the reader adds (GridDef.f90:406-488), and the binary E-field
solution files (3D_MT/ioMod/ioAscii.f90: FileWriteInit, EfileWrite,
FileReadInit, EfileRead, as write_solnVectorMTX / read_solnVectorMTX use
them).  Values read are bit-identical to what the reference's readers
produce from the same file."""
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
