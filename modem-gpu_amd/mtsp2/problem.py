"""Synthetic SP2 problems: grids, seeded resistivity models, frequencies and
plane-wave boundary data for the benchmark and test workloads.

Grids follow the reference's conventions (grid_t, f90/3D_MT/GridDef.f90:34-91):
cell sizes dx(nx), dy(ny), dz(nzAir+nzEarth) with the air layers first (top
down), conductivity as LOGE (natural log of S/m) on earth cells in Fortran
order.  Edge vectors use the reference edge order [x ; y ; z], each
Fortran-ordered (vecTranslate.f90:41-78).

Workloads (sizes in cells, plus 10 air layers):
  cfg0  40x40x30   one 10 ohm-m block in 100 ohm-m, 1 Hz
  cfg1  100x100x60 200 seeded boxes, log10 rho ~ U(0,3), 8 f in 1e-3..1e2 Hz
  cfg3  120x120x80 as cfg1, 16 f in 1e-4..1e3 Hz
  cfg4  80x80x50   as cfg1, 12 f in 1e-3..1e2 Hz
Vertical cells start at 50 m and grow x1.2, capped at 10 km per cell: with
60-80 layers the uncapped growth would put the base thousands of km deep.
Air layers: 10 layers whose tops sit at heights log-spaced 100 m .. 1000 km.
"""
from __future__ import annotations

import numpy as np

PI = 3.14159265357898     # the reference's constant (math_constants.f90:12)
MU0 = PI * 4e-7
ISIGN = -1


def _pad(ncore, dcore, npad=8, grow=1.3):
    pad = dcore * grow ** np.arange(1, npad + 1)
    return np.concatenate([pad[::-1], np.full(ncore, dcore), pad])


def air_layers(nair=10, hmin=100.0, hmax=1e6):
    tops = np.geomspace(hmin, hmax, nair)          # height of each layer's top
    th = np.diff(np.concatenate([[0.0], tops]))    # bottom-up thicknesses
    return th[::-1].copy()                          # top-down


def earth_layers(n, d0=50.0, grow=1.2, dmax=1e4):
    return np.minimum(d0 * grow ** np.arange(n), dmax)


def make_grid(nx, ny, nzEarth, nair=10, dcore=500.0, npad=8):
    return {"dx": _pad(nx - 2 * npad, dcore, npad), "dy": _pad(ny - 2 * npad, dcore, npad),
            "dz": np.concatenate([air_layers(nair), earth_layers(nzEarth)]), "nzAir": nair}


CONFIGS = {
    "cfg0": dict(nx=40, ny=40, nzEarth=30, f=np.array([1.0])),
    "cfg1": dict(nx=100, ny=100, nzEarth=60, f=10.0 ** (-3 + 5 * np.arange(8) / 7)),
    "cfg3": dict(nx=120, ny=120, nzEarth=80, f=10.0 ** (-4 + 7 * np.arange(16) / 15)),
    "cfg4": dict(nx=80, ny=80, nzEarth=50, f=10.0 ** (-3 + 5 * np.arange(12) / 11)),
}


def config_grid(name):
    c = CONFIGS[name]
    return make_grid(c["nx"], c["ny"], c["nzEarth"])


def config_freqs(name):
    return CONFIGS[name]["f"].copy()


def block_model(grid, seed=0, rho_bg=100.0, rho_blk=10.0, size=(10, 10, 5), top_depth=1000.0):
    """cfg0 model: background with one conductive block inside the core,
    corners drawn from the seed, top at `top_depth` metres."""
    rng = np.random.default_rng(seed)
    nx, ny = len(grid["dx"]), len(grid["dy"])
    dze = grid["dz"][grid["nzAir"]:]
    lc = np.full((nx, ny, len(dze)), -np.log(rho_bg))
    depth_top = np.concatenate([[0.0], np.cumsum(dze)])[:-1]
    k0 = int(np.argmin(np.abs(depth_top - top_depth)))
    i0 = rng.integers(8, nx - 8 - size[0] + 1)
    j0 = rng.integers(8, ny - 8 - size[1] + 1)
    lc[i0:i0 + size[0], j0:j0 + size[1], k0:k0 + size[2]] = -np.log(rho_blk)
    return lc


def boxes_model(grid, nbox=200, seed=0, rho_bg=100.0, max_size=(12, 12, 8)):
    """Blocky seeded model: `nbox` random boxes with log10 rho ~ U(0, 3) over
    the background, placed inside the core (padding left at background)."""
    rng = np.random.default_rng(seed)
    nx, ny = len(grid["dx"]), len(grid["dy"])
    nze = len(grid["dz"]) - grid["nzAir"]
    lc = np.full((nx, ny, nze), -np.log(rho_bg))
    for _ in range(nbox):
        sx, sy, sz = (rng.integers(2, m + 1) for m in max_size)
        i0 = rng.integers(8, max(9, nx - 8 - sx))
        j0 = rng.integers(8, max(9, ny - 8 - sy))
        k0 = rng.integers(0, max(1, nze - sz - nze // 4))
        lc[i0:i0 + sx, j0:j0 + sy, k0:k0 + sz] = -np.log(10.0 ** rng.uniform(0.0, 3.0))
    return lc


def config_model(name, grid, seed=0):
    return block_model(grid, seed) if name == "cfg0" else boxes_model(grid, seed=seed)


# sites per axis on the core: 10 x 10 (cfg0), 20 x 20 = 400 (cfg1, cfg3), 12 x 12 = 144 (cfg4)
CONFIG_SITES = {"cfg0": 10, "cfg1": 20, "cfg3": 20, "cfg4": 12}


def core_sites(grid, nside, npad=8):
    """nside x nside receivers on the surface (z = 0) spread over the core
    cells (the padding excluded), in the grid frame with its origin at the
    grid's corner (x = 0 at the first x node).  Each site sits 0.37 of a cell
    from its cell's low corner, so no site lies on a node or a cell centre
    (where the interpolation rule would meet ties)."""
    out = []
    for d in (grid["dx"], grid["dy"]):
        nodes = np.concatenate([[0.0], np.cumsum(d)])
        ncore = len(d) - 2 * npad
        cells = npad + ((np.arange(nside) + 0.5) * ncore / nside).astype(int)
        out.append(nodes[cells] + 0.37 * d[cells])
    X, Y = np.meshgrid(out[0], out[1], indexing="ij")
    return np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])


def config_sites(name, grid):
    return core_sites(grid, CONFIG_SITES[name])


# ------------------------------------------------------------- edge layout
def edge_counts(nx, ny, nz):
    nXe, nYe, nZe = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz
    return nXe, nYe, nZe


def interior_edges(nx, ny, nz):
    """EDGEi (0-based) in the reference order (vecTranslate.f90:436-527)."""
    ix = np.arange(nx)
    jx, kx = np.arange(1, ny), np.arange(1, nz)
    nXe, nYe, _ = edge_counts(nx, ny, nz)
    ex = (ix[None, None, :] + nx * (jx[None, :, None] + (ny + 1) * kx[:, None, None])).ravel()
    iy, jy, ky = np.arange(1, nx), np.arange(ny), np.arange(1, nz)
    ey = nXe + (iy[None, None, :] + (nx + 1) * (jy[None, :, None] + ny * ky[:, None, None])).ravel()
    iz, jz, kz = np.arange(1, nx), np.arange(1, ny), np.arange(nz)
    ez = nXe + nYe + (iz[None, None, :] + (nx + 1) * (jz[None, :, None] + (ny + 1) * kz[:, None, None])).ravel()
    return np.concatenate([ex, ey, ez]).astype(np.int64)


def node_depths(grid):
    dz = grid["dz"]
    air = dz[:grid["nzAir"]].sum()
    return np.concatenate([[0.0], np.cumsum(dz)]) - air


def plane_wave(grid, omega, pol, rho=100.0):
    """All-edge field of a unit plane wave over a `rho` halfspace (e^{-i w t},
    ISIGN = -1): E = exp(i k z) in the earth, 1 + i k z in the air, with
    k = sqrt(i w mu0 / rho).  pol 0 fills x-edges (Ex source), pol 1 y-edges."""
    nx, ny, nz = len(grid["dx"]), len(grid["dy"]), len(grid["dz"])
    nXe, nYe, nZe = edge_counts(nx, ny, nz)
    zn = node_depths(grid)
    k = np.sqrt(1j * omega * MU0 / rho)
    ez = 1 + 1j * k * zn
    below = zn >= 0
    ez[below] = np.exp(1j * k * zn[below])
    e = np.zeros(nXe + nYe + nZe, np.complex128)
    if pol == 0:
        e[:nXe] = np.broadcast_to(ez[:, None, None], (nz + 1, ny + 1, nx)).ravel()
    else:
        e[nXe:nXe + nYe] = np.broadcast_to(ez[:, None, None], (nz + 1, ny, nx + 1)).ravel()
    return e


def forward_systems(ctx, grid, freqs_hz, pols=(0, 1)):
    """Right-hand sides b = -CCib e_b (FWDsolve3D, EMsolve3D.f90:261-274) and
    initial guesses x0 = e(EDGEi) for every (frequency, polarization), in
    frequency-major order.  Returns omegas, B (nsys, nInt), X0."""
    nx, ny, nz = len(grid["dx"]), len(grid["dy"]), len(grid["dz"])
    ei = interior_edges(nx, ny, nz)
    om, B, X0 = [], [], []
    for f in freqs_hz:
        w = 2 * PI * f
        for p in pols:
            e = plane_wave(grid, w, p)
            om.append(w)
            B.append(ctx.rhs_fwd(e))
            X0.append(e[ei])
    return np.array(om), np.array(B), np.array(X0)
